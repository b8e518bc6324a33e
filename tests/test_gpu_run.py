"""The run loop's monitors and a whole run of the reference's 0-D reactor example on the GPU.

* compute_residual, compute_rho_t_p_minmax and find_boom (kernels_gpu.F90:5991-6012,
  6038-6073, 10334-10358) against the oracle on a reacting box, with and without a fluid
  mask: the residual within 1e-12 relative (the sums are taken in different orders), the
  extrema of the same auxiliary field and the NaN search exactly.
* examples/batch_reactor/multideal.ini (committed as tests/golden/batch_reactor.ini): the
  Keromnes-2013 mixture at 1000 K and 1 atm in a 1 x 1 x 1 periodic box, fixed dt = 2e-7 s,
  Rosenbrock-Krylov chemistry with Strang splitting, 10,000 steps, run by driver.Run through
  the whole field solver (fused substeps, halo copies, WENO on every face, monitors every
  1,000 steps); its output_batch.dat rows against Cantera's results with the example's own
  tolerances (|dt| <= 1e-7 s, |drho| <= 1e-5 kg/m^3, |dT| <= 1 K; result_comp.py:7-9).
"""
import pathlib

import numpy as np
import pytest

from parity_util import box_state

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
NG = 3


def test_monitors_match_oracle(pkg, h2_tables, oracle):
    import oracle_backend
    hip = pkg.solver.HipBackend()
    cpu = oracle_backend.OracleBackend(oracle)
    shape = (14, 10, 8)
    w, tg, info = box_state(pkg, h2_tables, shape, seed=9, temperature=1400.0, perturb_y=0.2)
    nv = h2_tables.n_species + 4
    rng = np.random.default_rng(4)
    fln = rng.standard_normal((nv, shape[2], shape[1], shape[0])) * 1e-3
    mask = np.zeros(w.shape[1:], dtype=np.int32)
    mask[NG + 2, NG + 3, NG + 4:NG + 9] = 1
    boom = fln.copy()
    boom[1, 5, 7, 9] = np.nan                 # variable 2, cell (i,j,k) = (10, 8, 6)
    boom[1, 2, 3, 4] = np.nan                 # cell (5, 4, 3): the first in i-fastest order
    out = {}
    aux = None
    for be in (hip, cpu):
        s = pkg.solver.Solver(be, h2_tables, *shape, NG, [info["length"]] * 3, rho0=info["rho0"], u0=info["u_rms"])
        s.set_state(w, tg)
        s.compute_aux()
        if aux is None:
            be.sync()
            aux = be.to_numpy(s.w_aux).copy()
        else:
            # both sides search the same auxiliary field (compute_aux itself agrees to ~1e-15)
            s.w_aux = be.asarray(aux)
        r = {}
        for name, m in (("plain", None), ("masked", be.asarray(mask, "i4"))):
            res = np.zeros(1)
            be.call("compute_residual", *shape, NG, nv, be.asarray(fln), 3e-8, res, m, be.stream)
            ext = [np.zeros(1) for _ in range(6)]
            be.call("compute_rho_t_p_minmax", s.ns, *shape, NG, s.nv_aux, s.w_aux, *ext, m, be.stream)
            n, ijk = np.zeros(1, dtype=np.int32), np.zeros(3, dtype=np.int32)
            be.call("find_boom", *shape, NG, nv, be.asarray(boom), 3e-8, m, 2, 1, n, ijk, be.stream)
            n0, ijk0 = np.zeros(1, dtype=np.int32), np.zeros(3, dtype=np.int32)
            be.call("find_boom", *shape, NG, nv, be.asarray(fln), 3e-8, m, 0, 0, n0, ijk0, be.stream)
            r[name] = (res[0], [float(e[0]) for e in ext], int(n[0]), ijk.tolist(), int(n0[0]), ijk0.tolist())
        out[be.name] = r
    for name in ("plain", "masked"):
        g, c = out["hip"][name], out["oracle"][name]
        print(f"[monitors] {name}: residual {g[0]:.6e} vs {c[0]:.6e}", flush=True)
        assert abs(g[0] - c[0]) <= 1e-12 * c[0]
        assert g[1] == c[1]
        assert g[2:] == c[2:]
    # the planted NaNs: two cells, the first at (5 + 2*14, 4, 3 + 1*8) with ncoords (2, 1)
    assert out["oracle"]["plain"][2:4] == (2, [5 + 2 * 14, 4, 3 + 8])
    assert out["oracle"]["plain"][4:] == (0, [0, 0, 0])
    # masking changes the extrema and the residual
    assert out["oracle"]["masked"][0] < out["oracle"]["plain"][0]


def test_batch_reactor_example_through_the_field_solver(pkg):
    run = pkg.driver.Run.from_ini(GOLD / "batch_reactor.ini", pkg.solver.HipBackend(),
                                  mechanism=str(pkg.cases.mechanism_path("keromnes2013_14sp.yaml")))
    assert run.c.flow_init == 8 and run.c.num_iter == 10000 and run.c.cfl == -2e-7
    assert run.s.num.rosenbrock_version == 2 and run.s.num.operator_splitting == 2
    assert run.s.fused_supported()
    run.update_start()
    for k in range(run.c.num_iter):
        if k % 1000 == 0:
            print(f"[run] batch reactor step {k}", flush=True)
        run.step()
    rows = np.array(run.batch_rows)
    gold = np.loadtxt(GOLD / "batch_reactor_cantera.dat")
    assert rows.shape == gold.shape
    d = np.abs(rows - gold).max(axis=0)
    print(f"[run] output_batch vs Cantera: max |dt| {d[0]:.2e} s, |drho| {d[1]:.2e}, |dT| {d[2]:.3f} K; "
          f"{len(run.progress)} progress lines, residual {run.residual_rhou:.3e}", flush=True)
    assert d[0] <= 1e-7 and d[1] <= 1e-5 and d[2] <= 1.0
    assert len(run.progress) == 10
    # the last monitors: the uniform box has T and rho extrema equal to the reactor's
    e = run.extrema
    assert e["tmin"] == e["tmax"] and abs(e["tmax"] - gold[-1, 2]) < 1.0


def test_field_copies_and_restart_files_are_bit_exact(pkg, h2_tables, tmp_path):
    """copy_gpu_cpu / copy_cpu_gpu (the layout change runs on the device) and the restart
    files written from the GPU field: bit for bit the host layouts, read back by the
    oracle's solver into the same state."""
    import oracle_backend
    hip = pkg.solver.HipBackend()
    shape = (20, 12, 6)
    w, tg, info = box_state(pkg, h2_tables, shape, seed=21, temperature=1300.0, perturb_y=0.2)
    g = pkg.solver.Solver(hip, h2_tables, *shape, NG, [info["length"]] * 3)
    g.set_state(w, tg)
    nv = g.nv
    host = np.empty(w.shape[1:] + (nv,))
    hip.call("copy_gpu_cpu", *shape, NG, nv, g.w, host, hip.stream)
    assert np.array_equal(host, np.moveaxis(w, 0, -1))
    host2 = 2.0 * host
    hip.call("copy_cpu_gpu", *shape, NG, nv, host2, g.w, hip.stream)
    assert np.array_equal(hip.to_numpy(g.w), 2.0 * w)
    g.set_state(w, tg)
    c = pkg.solver.Solver(oracle_backend.OracleBackend(), h2_tables, *shape, NG, [info["length"]] * 3)
    for io_type, name in ((2, "rst.bin"), (1, "rst1_0000_0000_0000.bin")):
        d = tmp_path / f"io{io_type}"
        g.save_field(d, io_type)
        raw = (d / name).read_bytes()
        interior = w[:, NG:-NG, NG:-NG, NG:-NG]
        want = interior.tobytes() if io_type == 2 else np.moveaxis(interior, 0, -1).tobytes()
        body = raw if io_type == 2 else raw[4:-4]
        assert body == want
        if io_type == 1:
            (d / "rst0_0000_0000_0000.bin").write_bytes(raw)
        c.set_state(np.zeros_like(w), tg)
        c.load_field(d, io_type)
        assert np.array_equal(c.w[:, NG:-NG, NG:-NG, NG:-NG], interior)
