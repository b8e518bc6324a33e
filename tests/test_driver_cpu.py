"""The run loop and the input-file reader on the CPU oracle (the GPU runs are in
test_gpu_run.py): the reference's example inputs build the configured solver, and a few
steps produce the monitors and writers of manage_output."""
import pathlib

import numpy as np
import pytest

import oracle_backend

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def be():
    return oracle_backend.OracleBackend()


def test_read_ini_sections():
    pkg_ini = GOLD / "reactive_shock_tube.ini"
    import prop_streams2_amd.driver as drv
    ini = drv.read_ini(pkg_ini)
    assert ini["grid"]["nxmax"] == "400"
    assert ini["controls"]["cfl"] == "-2E-8"
    assert ini["numerics"]["sensor_threshold"] == "-0.2"
    assert ini["fluid"]["enable_chemistry"] == "2"
    assert float(ini["ref_state"]["init_y_09"]) == pytest.approx(0.88594456906)


def test_batch_reactor_example_steps(pkg, be, tmp_path):
    run = pkg.driver.Run.from_ini(GOLD / "batch_reactor.ini", be,
                                  mechanism=str(pkg.cases.mechanism_path("keromnes2013_14sp.yaml")),
                                  outdir=str(tmp_path))
    assert (run.s.nx, run.s.ny, run.s.nz) == (1, 1, 1)
    assert run.s.dt == 0.0
    run.c.print_control = 10
    run.update_start()
    assert run.s.dt == 2e-7
    for _ in range(20):
        run.step()
    rows = np.loadtxt(tmp_path / "output_batch.dat")
    gold = np.loadtxt(GOLD / "batch_reactor_cantera.dat")
    assert rows.shape == (2, 3)
    np.testing.assert_allclose(rows[:, 0], gold[:2, 0], rtol=0, atol=1e-12)
    assert np.abs(rows[:, 2] - gold[:2, 2]).max() < 1e-3
    prog = np.loadtxt(tmp_path / "progress.out")
    assert prog.shape == (2, 11)
    assert prog[-1, 0] == 20 and prog[-1, 3] == pytest.approx(4e-6)
    # a uniform periodic box has no transport: the residual is round-off (the ghost copies'
    # Newton temperatures start from other guesses than the cell's), the extrema are equal
    assert prog[-1, 4] < 1e-6 and prog[-1, 7] == prog[-1, 8]


def test_shock_tube_example_steps(pkg, be, tmp_path):
    run = pkg.driver.Run.from_ini(GOLD / "reactive_shock_tube.ini", be,
                                  mechanism=str(pkg.cases.mechanism_path("h2_shocktube_9sp.yaml")),
                                  outdir=str(tmp_path), nx=60)
    s = run.s
    assert s.bctags == [8, 2, 0, 0, 0, 0]
    assert s.num.sensor_threshold == -0.2 and s.num.rosenbrock_version == 2
    # u0 = 0 in the input: set_flow_params replaces it by the sound speed
    assert 300.0 < s.u0 < 500.0 and s.eps_sensor == pytest.approx(s.u0 ** 2)
    run.c.print_control = 1
    run.update_start()
    for _ in range(3):
        run.step()
    e = run.extrema
    assert e["pmin"] == pytest.approx(7173.0, rel=1e-6) and e["pmax"] > 30000.0
    assert np.isfinite(run.residual_rhou) and run.residual_rhou > 0.0


def test_nan_residual_stops_the_run(pkg, be):
    """A NaN planted in the state (the x momentum of one cell; a species NaN would be
    clipped to zero by eval_aux) makes the residual NaN at the next monitor: the run stops
    with the location found by find_boom (the reference aborts with 'BOOM')."""
    run = pkg.driver.Run.from_ini(GOLD / "reactive_shock_tube.ini", be,
                                  mechanism=str(pkg.cases.mechanism_path("h2_shocktube_9sp.yaml")), nx=40)
    run.c.print_control = 1
    run.update_start()
    run.step()
    run.s.w[run.s.ns, 3, 3, 3 + 16] = np.nan
    with pytest.raises(pkg.driver.BoomError, match="BOOM"):
        run.step()


@pytest.mark.parametrize("io_type", [1, 2])
def test_restart_round_trip_continues_the_run(pkg, be, tmp_path, io_type):
    """write_field + write_field_info after 4 steps, read back into a fresh run (rst1 is
    renamed rst0 for io_type 1, as the reference's next run reads rst0), 3 more steps: the
    same state as 7 uninterrupted steps.  Explicit chemistry (enable_chemistry = 1) keeps the
    comparison free of the adaptive integrator's step decisions; the Newton temperature
    guesses after the restart come from the initial state, so the agreement is to the
    Newton tolerance."""
    import shutil
    mech = str(pkg.cases.mechanism_path("h2_shocktube_9sp.yaml"))

    def make():
        r = pkg.driver.Run.from_ini(GOLD / "reactive_shock_tube.ini", be, mechanism=mech, nx=48)
        r.s.num.enable_chemistry = 1
        return r

    ref = make()
    ref.update_start()
    for _ in range(7):
        ref.step()
    a = make()
    a.update_start()
    for _ in range(4):
        a.step()
    a.write_restart(tmp_path, io_type)
    if io_type == 1:
        shutil.copy(tmp_path / "rst1_0000_0000_0000.bin", tmp_path / "rst0_0000_0000_0000.bin")
        shutil.copy(tmp_path / "field_info1.dat", tmp_path / "field_info0.dat")
        raw = (tmp_path / "rst1_0000_0000_0000.bin").read_bytes()
        assert int.from_bytes(raw[:4], "little") == 8 * 13 * 48 == len(raw) - 8
    else:
        assert (tmp_path / "rst.bin").stat().st_size == 8 * 13 * 48
    text = (tmp_path / ("field_info0.dat")).read_text()
    assert text.startswith("[field_info]\n") and "icyc0 = 4\n" in text
    b = make()
    info = b.read_restart(tmp_path, io_type)
    assert info["icyc0"] == 4 and info["time0"] == a.s.time and b.s.time == a.s.time
    b.update_start()
    for _ in range(3):
        b.step()
    assert b.icyc == 7 and b.s.time == pytest.approx(ref.s.time, rel=1e-15)
    wi, wr = b.s.w[:, 3:-3, 3:-3, 3:-3], ref.s.w[:, 3:-3, 3:-3, 3:-3]
    err = (np.abs(wi - wr).max(axis=(1, 2, 3)) / np.maximum(np.abs(wr).max(axis=(1, 2, 3)), 1e-300))
    assert err.max() < 1e-10, err


def test_field_info_values_round_trip(tmp_path):
    import prop_streams2_amd.restart as rs
    info = dict(icyc0=12000, time0=2.4000000000000001e-4, itav=0, itslice_vtr=3, time_from_last_rst=1.0 / 3.0,
                time_from_last_write=5e-310, time_from_last_stat=0.0, time_from_last_slice=2.5,
                time_from_last_slice_vtr=2.5, istore=7)
    rs.write_field_info(tmp_path / "f.dat", info)
    back = rs.read_field_info(tmp_path / "f.dat")
    assert back == info
    assert "time_from_last_rst = 3.3333333333333331E-001" in (tmp_path / "f.dat").read_text()


def test_batched_chemistry_options_are_checked(pkg, be):
    """chem_batch_form >= 0 runs the HIP reactor batch: a backend without one, the simpler
    splitting (which needs the transport forcing the batch does not take) and the PaSR
    closure are refused before any work is done."""
    tb = pkg.cases.h2_air_mechanism().tables()
    w, tg, info = pkg.cases.turbulence_box(tb, 6, seed=3, ng=3)

    def solver(**kw):
        num = pkg.solver.Numerics(enable_chemistry=2, chem_batch_form=1, **kw)
        s = pkg.solver.Solver(be, tb, 6, 6, 6, 3, [info["length"]] * 3, numerics=num, rho0=info["rho0"],
                              u0=info["u_rms"])
        s.set_state(w, tg)
        s.compute_aux()
        s.dt = 1e-9
        return s

    w0 = solver().w.copy()
    with pytest.raises(ValueError, match="HIP reactor batch"):
        s = solver()
        s.advance()
    assert np.array_equal(s.w, w0)
    with pytest.raises(ValueError, match="simpler splitting"):
        solver(operator_splitting=1, rosenbrock_version=2).advance()
    with pytest.raises(ValueError, match="PaSR"):
        solver(enable_les=1, les_model=2).chemistry_batch(1e-9)
