"""GPU parity: each HIP entry point against the CPU restatement on identical seeded inputs.

Tolerance: RHS agreement within 1e-12 relative in the max-norm of each
conserved variable; species and temperature after N steps within 1e-9 relative.
Auxiliary variables are held to the same 1e-12.  Copies (halo exchange) are bit-exact.
Shapes are deliberately non-cubic so that a transposed stride or a swapped extent fails.
"""
import numpy as np
import pytest

from parity_util import box_state, faces_only_mask, interior, rel_maxnorm

pytestmark = pytest.mark.gpu

TOL = 1e-12
SHAPE = (18, 12, 10)   # nx, ny, nz
NG = 3


@pytest.fixture(scope="module")
def hip_backend(pkg):
    return pkg.solver.HipBackend()


@pytest.fixture(scope="module")
def cpu_backend(oracle):
    import oracle_backend
    return oracle_backend.OracleBackend(oracle)


def make_pair(pkg, tables, hip_backend, cpu_backend, shape=SHAPE, numerics=None, seed=11, **kw):
    w, tg, info = box_state(pkg, tables, shape, seed=seed, **kw)
    out = []
    for be in (hip_backend, cpu_backend):
        s = pkg.solver.Solver(be, tables, *shape, NG, [info["length"]] * 3, numerics=numerics,
                              rho0=info["rho0"], u0=info["u_rms"])
        s.set_state(w, tg)
        out.append(s)
    return out[0], out[1], info


def sync_aux(g, c):
    """Give the GPU solver the CPU solver's auxiliaries (isolates one kernel)."""
    g.be.copy_(g.w_aux, g.be.asarray(c.w_aux))
    g.be.copy_(g.w, g.be.asarray(c.w))


def prepare_aux(s):
    s.compute_aux()
    jduc, jdiv = s.ns + 9, s.ns + 10
    s.be.call("bcextr_var", s.nx, s.ny, s.nz, s.ng, s.w_aux[jdiv - 1], s.be.stream)
    s.bcswap_var(jduc)
    s.bcswap_var(jdiv)


def test_library_reports_species(pkg):
    assert 9 in pkg.capi.species_supported()
    assert 14 in pkg.capi.species_supported()


def test_eval_aux(pkg, h2_tables, hip_backend, cpu_backend):
    g, c, _ = make_pair(pkg, h2_tables, hip_backend, cpu_backend, perturb_y=0.3)
    g.eval_aux()
    c.eval_aux()
    g.be.sync()
    ns = h2_tables.n_species
    written = list(range(ns)) + list(range(ns, ns + 8)) + [ns + 10, ns + 11, ns + 12] + \
        list(range(ns + 13, 2 * ns + 13))
    ga, ca = g.be.to_numpy(g.w_aux)[written], c.w_aux[written]
    err = rel_maxnorm(ga, ca)
    assert err.max() < TOL, err
    assert np.array_equal(g.be.to_numpy(g.w), c.w)


def test_eval_aux_subbox_and_clip(pkg, h2_tables, hip_backend, cpu_backend):
    """A sub-box of the padded field (compute_aux central=1, ghost=0) and the rho*Y >= 0 clip."""
    g, c, _ = make_pair(pkg, h2_tables, hip_backend, cpu_backend)
    wn = c.w.copy()
    wn[3, 5, 6, 7] = -1e-9                    # a negative partial density is clipped in w
    for s in (g, c):
        s.be.copy_(s.w, s.be.asarray(wn))
        s.compute_aux(central=1, ghost=0)
    g.be.sync()
    assert np.array_equal(g.be.to_numpy(g.w), c.w)
    assert c.w[3, 5, 6, 7] == 0.0
    err = rel_maxnorm(interior(g.be.to_numpy(g.w_aux), NG)[: 2 * h2_tables.n_species + 13],
                      interior(c.w_aux, NG)[: 2 * h2_tables.n_species + 13])
    err[h2_tables.n_species + 8: h2_tables.n_species + 10] = 0.0   # sensor slots are not written
    assert err.max() < TOL, err


@pytest.mark.parametrize("sensor_type", [0, 1])
def test_eval_aux2(pkg, h2_tables, hip_backend, cpu_backend, sensor_type):
    num = pkg.solver.Numerics(sensor_type=sensor_type)
    g, c, _ = make_pair(pkg, h2_tables, hip_backend, cpu_backend, numerics=num)
    c.eval_aux()
    sync_aux(g, c)
    g.eval_aux2()
    c.eval_aux2()
    ns = h2_tables.n_species
    err = rel_maxnorm(interior(g.be.to_numpy(g.w_aux), NG)[ns + 8: ns + 10], interior(c.w_aux, NG)[ns + 8: ns + 10])
    assert err.max() < TOL, err


def test_eval_aux2_div_only(pkg, h2_tables, hip_backend):
    """sensor_type = -1 (the fused substeps with the sensor off): div/3 with the same bits as
    the full eval_aux2, and the Ducros column left as it was."""
    g, _, _ = make_pair(pkg, h2_tables, hip_backend, hip_backend)
    g.eval_aux()
    ns = h2_tables.n_species
    g.eval_aux2()
    hip_backend.sync()
    full = hip_backend.to_numpy(g.w_aux).copy()
    g.w_aux[ns + 8].fill_(-7.0)
    g.w_aux[ns + 9].fill_(0.0)
    g.eval_aux2(ducros=False)
    hip_backend.sync()
    part = hip_backend.to_numpy(g.w_aux)
    assert np.array_equal(interior(part, NG)[ns + 9], interior(full, NG)[ns + 9])
    assert (part[ns + 8] == -7.0).all()


@pytest.mark.parametrize("periodic", [(1, 1, 1), (1, 0, 1), (0, 1, 0)])
def test_bcswap_bit_exact(pkg, hip_backend, cpu_backend, periodic):
    rng = np.random.default_rng(5)
    nx, ny, nz = SHAPE
    a = rng.standard_normal((4, nz + 2 * NG, ny + 2 * NG, nx + 2 * NG))
    per = np.array(periodic, dtype=np.int32)
    ga = hip_backend.asarray(a)
    ca = a.copy()
    hip_backend.call("bcswap", nx, ny, nz, NG, 4, ga, per, hip_backend.stream)
    cpu_backend.call("bcswap", nx, ny, nz, NG, 4, ca, per, None)
    assert np.array_equal(hip_backend.to_numpy(ga), ca)


def test_bcswap_short_axis_bit_exact(pkg, hip_backend, cpu_backend):
    """An axis shorter than ng (the 1-D tube's y and z) takes the staged path."""
    rng = np.random.default_rng(6)
    nx, ny, nz = 9, 1, 2
    a = rng.standard_normal((3, nz + 2 * NG, ny + 2 * NG, nx + 2 * NG))
    per = np.array([0, 1, 1], dtype=np.int32)
    ga = hip_backend.asarray(a)
    ca = a.copy()
    hip_backend.call("bcswap", nx, ny, nz, NG, 3, ga, per, hip_backend.stream)
    cpu_backend.call("bcswap", nx, ny, nz, NG, 3, ca, per, None)
    assert np.array_equal(hip_backend.to_numpy(ga), ca)


@pytest.mark.parametrize("ilat", [1, 2, 3, 4, 5, 6])
def test_physical_bc_bit_exact(pkg, hip_backend, cpu_backend, ilat):
    rng = np.random.default_rng(ilat)
    nx, ny, nz = SHAPE
    ns = 9
    a = rng.standard_normal((ns + 4, nz + 2 * NG, ny + 2 * NG, nx + 2 * NG))
    for name, args in (("bcextr", (ilat, nx, ny, nz, NG, ns + 4)), ("bcwall_ad", (ns, ilat, nx, ny, nz, NG))):
        ga = hip_backend.asarray(a)
        ca = a.copy()
        hip_backend.call(name, *args, ga, hip_backend.stream)
        cpu_backend.call(name, *args, ca, None)
        assert np.array_equal(hip_backend.to_numpy(ga), ca), name
    ga = hip_backend.asarray(a[0])
    ca = a[0].copy()
    hip_backend.call("bcextr_var", nx, ny, nz, NG, ga, hip_backend.stream)
    cpu_backend.call("bcextr_var", nx, ny, nz, NG, ca, None)
    assert np.array_equal(hip_backend.to_numpy(ga), ca)


def _flux_case(pkg, h2_tables, hip_backend, cpu_backend, numerics, stage, shape=SHAPE, **kw):
    g, c, _ = make_pair(pkg, h2_tables, hip_backend, cpu_backend, numerics=numerics, shape=shape, **kw)
    prepare_aux(c)
    sync_aux(g, c)
    for s in (g, c):
        stage(s)
    g.be.sync()
    return rel_maxnorm(g.be.to_numpy(g.fl), c.fl)


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("nkeep", [0, 1, -1])
def test_euler_central(pkg, h2_tables, hip_backend, cpu_backend, d, nkeep):
    num = pkg.solver.Numerics(nkeep=nkeep, sensor_threshold=2.0)
    err = _flux_case(pkg, h2_tables, hip_backend, cpu_backend, num, lambda s: s.euler(d))
    assert err.max() < TOL, err


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("splitting,version", [(0, 1), (0, 0), (1, 1)])
def test_euler_weno(pkg, h2_tables, hip_backend, cpu_backend, d, splitting, version):
    """Threshold 0.05 with the classic sensor puts a large share of faces on the WENO branch."""
    num = pkg.solver.Numerics(sensor_threshold=0.05, sensor_type=0, flux_splitting=splitting,
                              weno_version=version)
    err = _flux_case(pkg, h2_tables, hip_backend, cpu_backend, num, lambda s: s.euler(d), mach_t=0.6)
    assert err.max() < TOL, err


@pytest.mark.parametrize("d", [0, 2])
def test_euler_boundary_order_reduction(pkg, h2_tables, hip_backend, cpu_backend, d):
    """Non-periodic faces with correct_bound_ord: reduced KEEP order and zero-flux faces."""
    num = pkg.solver.Numerics(correct_bound_ord=1, sensor_threshold=0.1, sensor_type=0)
    w, tg, info = box_state(pkg, h2_tables, SHAPE, seed=3, mach_t=0.5)
    out = []
    for be in (hip_backend, cpu_backend):
        s = pkg.solver.Solver(be, h2_tables, *SHAPE, NG, [info["length"]] * 3, numerics=num,
                              bctags=(2, 8, 0, 0, 2, 2), rho0=info["rho0"], u0=info["u_rms"])
        s.set_state(w, tg)
        s.force_zero_flux = [1, 0, 0, 0, 0, 1]
        out.append(s)
    g, c = out
    prepare_aux(c)
    sync_aux(g, c)
    for s in (g, c):
        s.euler(d)
    err = rel_maxnorm(g.be.to_numpy(g.fl), c.fl)
    assert err.max() < TOL, err


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_visflx(pkg, h2_tables, hip_backend, cpu_backend, mode):
    err = _flux_case(pkg, h2_tables, hip_backend, cpu_backend, None, lambda s: s.visflx(mode), perturb_y=0.3)
    assert err.max() < TOL, err


def test_laplacian_viscous_substeps_on_a_stretched_grid(pkg, h2_tables, hip_backend, cpu_backend):
    """conservative_viscous = 0 (visflx_cuf in rk_sync) on a grid stretched in x and z, so
    that the second-derivative metrics are not zero: three substeps routine by routine."""
    L = 2 * np.pi * 1e-3
    maps = [lambda xi: xi + 0.2 * L / (2 * np.pi) * np.sin(2 * np.pi * xi / L), None,
            lambda xi: xi - 0.15 * L / (2 * np.pi) * np.sin(2 * np.pi * xi / L)]
    num = pkg.solver.Numerics(conservative_viscous=0)
    w, tg, info = box_state(pkg, h2_tables, SHAPE, seed=13, temperature=1400.0, perturb_y=0.2)
    out = []
    for be in (hip_backend, cpu_backend):
        s = pkg.solver.Solver(be, h2_tables, *SHAPE, NG, [info["length"]] * 3, numerics=num, rho0=info["rho0"],
                              u0=info["u_rms"], mappings=maps)
        s.set_state(w, tg)
        s.compute_aux()
        out.append(s)
    g, c = out
    c.compute_dt(0.5)
    g.dt = c.dt
    assert not g.fused_supported()
    g.rk_substep(0)
    c.rk_substep(0)
    g.be.sync()
    err_fl = rel_maxnorm(g.be.to_numpy(g.fl), c.fl)
    for q in (1, 2):
        g.rk_substep(q)
        c.rk_substep(q)
    g.be.sync()
    err_w = rel_maxnorm(interior(g.be.to_numpy(g.w), NG), interior(c.w, NG))
    assert err_fl.max() < TOL, err_fl
    assert err_w.max() < 1e-13, err_w


@pytest.mark.parametrize("nr_type", [1, 2, 3, 6])
def test_bc_nr_lat_all_faces(pkg, h2_tables, hip_backend, cpu_backend, nr_type):
    """bc_nr_lat on all six faces of a non-periodic box (one call per face, as bc_nr does)."""
    w, tg, info = box_state(pkg, h2_tables, SHAPE, seed=31, mach_t=0.5, perturb_y=0.2)
    winf = w[:, NG + 2, NG + 3, NG + 4].copy() * 1.01
    out = []
    for be in (hip_backend, cpu_backend):
        s = pkg.solver.Solver(be, h2_tables, *SHAPE, NG, [info["length"]] * 3, bctags=(2,) * 6, bctags_nr=(nr_type,) * 6,
                              winf=winf, rho0=info["rho0"], u0=info["u_rms"])
        s.set_state(w, tg)
        out.append(s)
    g, c = out
    c.compute_aux()
    sync_aux(g, c)
    for s in (g, c):
        s.be.copy_(s.fl, s.be.asarray(np.zeros(c.fl.shape)))
        s.bc_nr()
    g.be.sync()
    err = rel_maxnorm(g.be.to_numpy(g.fl), c.fl)
    assert np.abs(c.fl).max() > 0.0
    assert err.max() < TOL, err


def test_substeps_with_characteristic_faces(pkg, h2_tables, hip_backend, cpu_backend):
    """rk_sync routine by routine with a non-reflecting xmin face, a freestream-relaxation
    xmax face and a reflective-wall zmax face: the face cells leave the convective update
    (bc_preproc) and take bc_nr's term instead."""
    w, tg, info = box_state(pkg, h2_tables, SHAPE, seed=33, mach_t=0.4, perturb_y=0.2, temperature=1300.0)
    winf = w[:, NG, NG, NG].copy()
    out = []
    for be in (hip_backend, cpu_backend):
        s = pkg.solver.Solver(be, h2_tables, *SHAPE, NG, [info["length"]] * 3, bctags=(2, 2, 0, 0, 2, 2),
                              bctags_nr=(1, 3, 0, 0, 0, 6), winf=winf, rho0=info["rho0"], u0=info["u_rms"])
        s.set_state(w, tg)
        s.compute_aux()
        out.append(s)
    g, c = out
    assert g.eul == (2, SHAPE[0] - 1, 1, SHAPE[1], 1, SHAPE[2] - 1) and not g.fused_supported()
    c.compute_dt(0.5)
    g.dt = c.dt
    for q in range(3):
        g.rk_substep(q)
        c.rk_substep(q)
        if q == 0:
            g.be.sync()
            err_fl = rel_maxnorm(g.be.to_numpy(g.fl), c.fl)
    g.be.sync()
    err_w = rel_maxnorm(interior(g.be.to_numpy(g.w), NG), interior(c.w, NG))
    assert err_fl.max() < TOL, err_fl
    assert err_w.max() < 1e-13, err_w


def test_chemistry(pkg, h2_tables, hip_backend, cpu_backend):
    """Explicit source at a temperature where every reaction class is active."""
    err = _flux_case(pkg, h2_tables, hip_backend, cpu_backend, None, lambda s: s.compute_chemistry(),
                     temperature=1500.0, perturb_y=0.5)
    assert err.max() < TOL, err


def test_compute_dt(pkg, h2_tables, hip_backend, cpu_backend):
    g, c, _ = make_pair(pkg, h2_tables, hip_backend, cpu_backend)
    c.compute_aux()
    sync_aux(g, c)
    dg, dc = g.compute_dt(0.5), c.compute_dt(0.5)
    assert abs(dg - dc) <= 1e-14 * dc


def test_rk_step_full(pkg, h2_tables, hip_backend, cpu_backend):
    """Three Wray substeps of the whole RHS (halo, aux, sweeps, viscous, chemistry, update)."""
    g, c, _ = make_pair(pkg, h2_tables, hip_backend, cpu_backend, perturb_y=0.2, temperature=1400.0)
    for s in (g, c):
        s.compute_aux()
        s.compute_dt(0.5)
    g.dt = c.dt
    for s in (g, c):
        s.rk_substep(0)
    g.be.sync()
    err = rel_maxnorm(g.be.to_numpy(g.fl), c.fl)
    assert err.max() < TOL, err
    for s in (g, c):
        s.rk_substep(1)
        s.rk_substep(2)
    g.be.sync()
    ns = h2_tables.n_species
    err = rel_maxnorm(interior(g.be.to_numpy(g.w), NG), interior(c.w, NG))
    assert err.max() < 1e-13, err
    for s in (g, c):
        s.compute_aux(central=1, ghost=0)
    tg = interior(g.be.to_numpy(g.w_aux), NG)[ns + 4]
    tc = interior(c.w_aux, NG)[ns + 4]
    assert np.abs(tg - tc).max() / tc.max() < 1e-12


def test_n_steps_species_and_temperature(pkg, h2_tables, hip_backend, cpu_backend):
    """N full RK3 steps: species and temperature within 1e-9 relative."""
    g, c, _ = make_pair(pkg, h2_tables, hip_backend, cpu_backend, shape=(12, 10, 8), temperature=1500.0,
                        perturb_y=0.2)
    for s in (g, c):
        s.compute_aux()
        s.compute_dt(0.5)
    g.dt = c.dt
    for _ in range(5):
        g.advance()
        c.advance()
    for s in (g, c):
        s.compute_aux(central=1, ghost=0)
    ns = h2_tables.n_species
    ga, ca = interior(g.be.to_numpy(g.w_aux), NG), interior(c.w_aux, NG)
    err = rel_maxnorm(ga[: ns], ca[: ns])
    assert err.max() < 1e-9, err
    assert np.abs(ga[ns + 4] - ca[ns + 4]).max() / ca[ns + 4].max() < 1e-9


def test_rosenbrock_cells(pkg, h2_tables, hip_backend, cpu_backend):
    """Implicit per-cell chemistry (rosenbrock_kernel) over one time step on a small box."""
    num = pkg.solver.Numerics(enable_chemistry=2)
    g, c, _ = make_pair(pkg, h2_tables, hip_backend, cpu_backend, shape=(6, 5, 4), numerics=num,
                        temperature=1500.0, perturb_y=0.4)
    for s in (g, c):
        s.compute_aux()
        s.rosenbrock(2e-6)
    g.be.sync()
    err = rel_maxnorm(interior(g.be.to_numpy(g.w), NG), interior(c.w, NG))
    assert err.max() < 1e-9, err


def test_rosenbrock_krylov_cells(pkg, h2_tables, hip_backend, cpu_backend):
    """Rosenbrock-Krylov chemistry (rosenbrock_krylov_kernel) over one time step on a small
    box, plain and with the simpler-splitting transport forcing fl_sav.  With the forcing the
    method's accept/reject and Krylov-size decisions make it sensitive to rounding (the
    oracle itself moves by ~1e-5 when its input moves by 1e-15): there the bound is ten
    times the oracle's own response to that perturbation."""
    num = pkg.solver.Numerics(enable_chemistry=2, rosenbrock_version=2)
    for forcing in (False, True):
        g, c = make_pair(pkg, h2_tables, hip_backend, cpu_backend, shape=(6, 5, 4), numerics=num,
                         temperature=1500.0, perturb_y=0.4)[:2]
        w0 = c.w.copy()
        f = None
        if forcing:
            # a transport-like forcing: moves each variable by ~0.1 % of itself over the step
            rng = np.random.default_rng(5)
            f = rng.standard_normal(c.fl.shape) * 1e-3 * np.abs(interior(c.w, NG)) / 2e-6
        for s, fs in ((g, g.be.asarray(f) if forcing else None), (c, f)):
            s.compute_aux()
            s.rosenbrock_krylov(2e-6, fs)
        g.be.sync()
        err = rel_maxnorm(interior(g.be.to_numpy(g.w), NG), interior(c.w, NG))
        bound = 1e-9
        if forcing:
            ref = c.w.copy()
            c.w[...] = w0 * (1.0 + 1e-15)
            c.compute_aux()
            c.rosenbrock_krylov(2e-6, f)
            bound = max(bound, 10.0 * rel_maxnorm(interior(c.w, NG), interior(ref, NG)).max())
        assert err.max() < bound, (forcing, bound, err)


def _advance_pair(pkg, tables, hip_backend, cpu_backend, numerics, shape, steps, seed=17, **kw):
    """`steps` calls of advance() on both sides from the same state, plus the oracle's own
    result from that state perturbed by 1e-15 (relative), all as interior (w, T)."""
    w, tg, info = box_state(pkg, tables, shape, seed=seed, **kw)
    probe = pkg.solver.Solver(cpu_backend, tables, *shape, NG, [info["length"]] * 3, numerics=numerics,
                              rho0=info["rho0"], u0=info["u_rms"])
    probe.set_state(w, tg)
    probe.compute_aux()
    dt = probe.compute_dt(0.5)

    def run(be, w0):
        s = pkg.solver.Solver(be, tables, *shape, NG, [info["length"]] * 3, numerics=numerics,
                              rho0=info["rho0"], u0=info["u_rms"])
        s.set_state(w0, tg)
        s.compute_aux()
        s.dt = dt
        for _ in range(steps):
            s.advance()
        s.compute_aux(central=1, ghost=0)
        s.be.sync()
        ns = tables.n_species
        return (interior(s.be.to_numpy(s.w), NG).copy(), interior(s.be.to_numpy(s.w_aux), NG)[ns + 4].copy())

    wp = w * (1.0 + 1e-15 * np.random.default_rng(2).standard_normal(w.shape))
    return run(hip_backend, w), run(cpu_backend, w), run(cpu_backend, wp)


def test_simpler_splitting_advance(pkg, h2_tables, hip_backend, cpu_backend):
    """advance_solution with operator_splitting = 1 (multideal_gpu.F90:1875-1893): sav_flx,
    Rosenbrock-Krylov chemistry over dt with the saved transport RHS as forcing, then
    rk_sync over dt/2 with init_flux_simpler / update_simpler_flux.  The forcing makes the
    Krylov accept/reject and subspace-size decisions sensitive to rounding (as in
    test_rosenbrock_krylov_cells), so species and T are held to ten times the oracle's own
    response to a 1e-15 input perturbation, never below 1e-9."""
    num = pkg.solver.Numerics(enable_chemistry=2, rosenbrock_version=2, operator_splitting=1)
    (gw, gt), (cw, ct), (pw, pt) = _advance_pair(pkg, h2_tables, hip_backend, cpu_backend, num, (8, 6, 5), 2,
                                                 temperature=1500.0, perturb_y=0.3)
    bound_w = max(1e-9, 10.0 * rel_maxnorm(pw, cw).max())
    bound_t = max(1e-9, 10.0 * np.abs(pt - ct).max() / ct.max())
    err = rel_maxnorm(gw, cw)
    assert err.max() < bound_w, (bound_w, err)
    assert np.abs(gt - ct).max() / ct.max() < bound_t


@pytest.mark.parametrize("version", [1, 2])
def test_strang_advance_on_the_fused_substeps(pkg, h2_tables, hip_backend, cpu_backend, version):
    """advance_solution with operator_splitting = 2 (Strang, multideal_gpu.F90:1894-1917):
    implicit chemistry over dt/2, the fused RK substeps on the GPU (routine by routine on
    the oracle), implicit chemistry over dt/2; ROS4 and Rosenbrock-Krylov.  Species and T
    after three steps within 1e-9."""
    num = pkg.solver.Numerics(enable_chemistry=2, rosenbrock_version=version, operator_splitting=2)
    (gw, gt), (cw, ct), _ = _advance_pair(pkg, h2_tables, hip_backend, cpu_backend, num, (8, 6, 5), 3,
                                          temperature=1500.0, perturb_y=0.3)
    err = rel_maxnorm(gw, cw)
    assert err.max() < 1e-9, err
    assert np.abs(gt - ct).max() / ct.max() < 1e-9


# ---------------------------------------------------------------- fused substep stages
def _fused_vs_sequence(pkg, tables, hip_backend, cpu_backend, shape=SHAPE, numerics=None, bctags=(0,) * 6,
                       steps=1, **kw):
    w, tg, info = box_state(pkg, tables, shape, seed=21, **kw)
    out = []
    for be in (hip_backend, cpu_backend):
        s = pkg.solver.Solver(be, tables, *shape, NG, [info["length"]] * 3, numerics=numerics, bctags=bctags,
                              rho0=info["rho0"], u0=info["u_rms"])
        s.set_state(w, tg)
        s.compute_aux()
        out.append(s)
    g, c = out
    c.compute_dt(0.5)
    g.dt = c.dt
    assert g.fused_supported()
    g.rk_substep_fused(0)
    c.rk_substep(0)
    g.be.sync()
    err_fl = rel_maxnorm(g.be.to_numpy(g.fl), c.fl)
    for _ in range(steps):
        for q in (1, 2):
            g.rk_substep_fused(q)
            c.rk_substep(q)
    g.be.sync()
    err_w = rel_maxnorm(interior(g.be.to_numpy(g.w), NG), interior(c.w, NG))
    return err_fl, err_w


def test_fused_substep_periodic(pkg, h2_tables, hip_backend, cpu_backend):
    err_fl, err_w = _fused_vs_sequence(pkg, h2_tables, hip_backend, cpu_backend, temperature=1500.0,
                                       perturb_y=0.3)
    assert err_fl.max() < TOL, err_fl
    assert err_w.max() < 1e-13, err_w


def test_fused_substep_weno_and_walls(pkg, h2_tables, hip_backend, cpu_backend):
    num = pkg.solver.Numerics(sensor_threshold=0.05, sensor_type=0, correct_bound_ord=1, nkeep=1)
    err_fl, err_w = _fused_vs_sequence(pkg, h2_tables, hip_backend, cpu_backend, numerics=num,
                                       bctags=(8, 2, 0, 0, 2, 2), mach_t=0.6, perturb_y=0.2)
    assert err_fl.max() < TOL, err_fl
    assert err_w.max() < 1e-13, err_w


def test_fused_substep_tube_like(pkg, h2_tables, hip_backend, cpu_backend):
    """nx x 1 x 1 with walls: the fused path skips the degenerate y and z sweeps as rk_sync does."""
    num = pkg.solver.Numerics(sensor_threshold=0.2, sensor_type=0)
    err_fl, err_w = _fused_vs_sequence(pkg, h2_tables, hip_backend, cpu_backend, shape=(40, 1, 1), numerics=num,
                                       bctags=(8, 2, 0, 0, 0, 0), mach_t=0.5)
    assert err_fl.max() < TOL, err_fl
    assert err_w.max() < 1e-13, err_w


def test_folded_substeps_are_bit_identical(pkg, h2_tables, hip_backend):
    """rk_sync folds each substep's init_flux into the previous substep's tail
    (streams_rhs_finish_next, then an x sweep with init=2).  Two RK steps give the same bits
    in w, fl and fln as the unfolded chain (streams_rhs_finish, then init=1), with explicit
    chemistry, and again with walls and WENO faces."""
    assert hip_backend.has("rhs_finish_next")
    cases = ((None, (0,) * 6),
             (pkg.solver.Numerics(sensor_threshold=0.05, sensor_type=0, nkeep=1), (8, 2, 0, 0, 2, 2)))
    for numerics, bctags in cases:
        w, tg, info = box_state(pkg, h2_tables, SHAPE, seed=31, temperature=1500.0, perturb_y=0.3)
        out = []
        dt = None
        for fold in (True, False):
            s = pkg.solver.Solver(hip_backend, h2_tables, *SHAPE, NG, [info["length"]] * 3, numerics=numerics,
                                  bctags=bctags, rho0=info["rho0"], u0=info["u_rms"])
            s.fold_substeps = fold
            s.set_state(w, tg)
            s.compute_aux()
            dt = s.compute_dt(0.5) if dt is None else dt
            s.dt = dt
            print(f"[fold] bctags {bctags} fold {fold}", flush=True)
            for _ in range(2):
                s.rk_sync()
            s.be.sync()
            out.append([s.be.to_numpy(a).copy() for a in (s.w, s.fl, s.fln)])
        for a, b in zip(*out):
            assert np.array_equal(a, b)


def test_fused_substep_syngas_mixing_layer(pkg, hip_backend, cpu_backend):
    """The 14-species Keromnes-2013 syngas mixing layer (the C4 configuration at 24 x 16 x 12)
    with the sensor threshold set so that about 10 % of the faces take the WENO flux: one
    fused RK step against the oracle's routine-by-routine sequence.  Across the layer the
    WENO weights amplify rounding: the oracle's own RHS moves by ~1e-9 (relative to the
    field max-norm) when its input moves by 1e-15, so the bound is ten times that response,
    measured here, and never below the 1e-12 of the other cases."""
    tb = pkg.cases.keromnes_mechanism().tables()
    shape = (24, 16, 12)
    w, tg, info = pkg.cases.mixing_layer(tb, *shape, seed=5)
    probe = pkg.solver.Solver(cpu_backend, tb, *shape, NG, info["length"], rho0=info["rho0"], u0=info["u0"])
    probe.set_state(w, tg)
    probe.compute_aux()
    duc = interior(probe.w_aux[tb.n_species + 8], NG)
    num = pkg.solver.Numerics(weno_scheme=3)
    num.sensor_threshold = pkg.cases.weno_threshold_for_fraction(duc, num.weno_scheme, 0.10)
    assert 0.0 < num.sensor_threshold < 1.0

    def run(be, w0, fused):
        s = pkg.solver.Solver(be, tb, *shape, NG, info["length"], numerics=num, rho0=info["rho0"], u0=info["u0"])
        s.set_state(w0, tg)
        s.compute_aux()
        s.dt = dt
        step = s.rk_substep_fused if fused else s.rk_substep
        step(0)
        s.be.sync()
        fl = s.be.to_numpy(s.fl).copy()
        for q in (1, 2):
            step(q)
        s.be.sync()
        return fl, interior(s.be.to_numpy(s.w), NG).copy()

    probe.compute_dt(0.5)
    dt = probe.dt
    g_fl, g_w = run(hip_backend, w, True)
    c_fl, c_w = run(cpu_backend, w, False)
    wp = w * (1.0 + 1e-15 * np.random.default_rng(1).standard_normal(w.shape))
    p_fl, p_w = run(cpu_backend, wp, False)
    b_fl = max(TOL, 10.0 * rel_maxnorm(p_fl, c_fl).max())
    b_w = max(1e-13, 10.0 * rel_maxnorm(p_w, c_w).max())
    err_fl = rel_maxnorm(g_fl, c_fl)
    err_w = rel_maxnorm(g_w, c_w)
    assert err_fl.max() < b_fl, (b_fl, err_fl)
    assert err_w.max() < b_w, (b_w, err_w)
