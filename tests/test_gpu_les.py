"""GPU parity of the LES path against the oracle: compute_aux_les (eval_velaux, les_wale_mut,
the ghost exchange of the sensor, div/3 and the eddy viscosity, eval_aux_les;
multideal_gpu.F90:445-498), the output-time chemistry diagnostics with the PaSR factor
(eval_chem_aux, kernels_gpu.F90:9676-9877), the PaSR closure inside both implicit
integrators, and whole Strang-split steps of an LES with the PaSR closure.

Tolerances: the auxiliaries within 1e-12 relative (the GPU's transport uses a reciprocal
with two Newton steps, a few units in the last place); the production rates within 1e-11
of each species' largest rate; the integrators over a step short enough for a handful of
accepted steps, and the LES steps, within 1e-9 (species and T), as the other substep tests.
"""
import numpy as np
import pytest

from parity_util import box_state, interior, rel_maxnorm

pytestmark = pytest.mark.gpu
NG = 3
LES = dict(enable_les=1, les_model=2, les_pr=0.9, les_sc=0.7, les_c_wale=0.5, les_c_yoshi=0.0886, les_c_eps=1.048,
           les_c_mix=0.1)


@pytest.fixture(scope="module")
def hip(pkg):
    return pkg.solver.HipBackend()


@pytest.fixture(scope="module")
def cpu(oracle):
    import oracle_backend
    return oracle_backend.OracleBackend(oracle)


def _pair(pkg, tables, hip, cpu, shape, seed=5, numerics=None, **kw):
    w, tg, info = box_state(pkg, tables, shape, seed=seed, **kw)
    out = []
    for be in (hip, cpu):
        s = pkg.solver.Solver(be, tables, *shape, NG, [info["length"]] * 3, numerics=numerics, rho0=info["rho0"],
                              u0=info["u_rms"])
        s.set_state(w, tg)
        out.append(s)
    return out


def _num(pkg, **kw):
    d = dict(enable_chemistry=2, rosenbrock_version=2, operator_splitting=2, **LES)
    d.update(kw)
    return pkg.solver.Numerics(**d)


def test_les_auxiliaries_match_oracle(pkg, h2_tables, hip, cpu):
    g, c = _pair(pkg, h2_tables, hip, cpu, (24, 20, 16), numerics=_num(pkg), temperature=1400.0, perturb_y=0.3)
    print("[les] compute_aux_les 24x20x16", flush=True)
    for s in (g, c):
        s.compute_aux()
    hip.sync()
    ga, ca = hip.to_numpy(g.w_aux), c.w_aux
    ns = h2_tables.n_species
    cols = list(range(ns + 11)) + [ns + 12] + list(range(ns + 13, 2 * ns + 14))   # all but J_Z
    worst = 0.0
    for m in cols:
        scale = max(np.abs(ca[m]).max(), 1e-300)
        err = np.abs(ga[m] - ca[m]).max() / scale
        worst = max(worst, err)
        assert err < 1e-12, (m + 1, err)
    assert np.array_equal(hip.to_numpy(g.w), c.w)          # the species clipping of eval_velaux
    print(f"[les] worst aux column {worst:.2e}", flush=True)


def test_chemistry_diagnostics_with_pasr_match_oracle(pkg, h2_tables, hip, cpu):
    g, c = _pair(pkg, h2_tables, hip, cpu, (20, 12, 10), numerics=_num(pkg), temperature=1500.0, perturb_y=0.4)
    for s in (g, c):
        s.compute_aux()
        s.compute_chem_aux()
    hip.sync()
    ns = h2_tables.n_species
    ga, ca = interior(hip.to_numpy(g.w_aux), NG), interior(c.w_aux, NG)
    wd_g, wd_c = ga[2 * ns + 14:3 * ns + 14], ca[2 * ns + 14:3 * ns + 14]
    scale = np.abs(wd_c).max(axis=(1, 2, 3)) + 1e-300
    err = (np.abs(wd_g - wd_c).max(axis=(1, 2, 3)) / scale).max()
    hrr = np.abs(ga[3 * ns + 14] - ca[3 * ns + 14]).max() / np.abs(ca[3 * ns + 14]).max()
    print(f"[les] wdot {err:.2e}, heat release {hrr:.2e}", flush=True)
    assert err < 1e-11 and hrr < 1e-11


@pytest.mark.parametrize("version", [1, 2])
def test_integrators_with_pasr_match_oracle(pkg, h2_tables, hip, cpu, version):
    g, c = _pair(pkg, h2_tables, hip, cpu, (10, 8, 6), numerics=_num(pkg, rosenbrock_version=version),
                 temperature=1500.0, perturb_y=0.4)
    ns = h2_tables.n_species
    for s in (g, c):
        s.compute_aux()
        if version == 2:
            z = s.be.to_numpy(s.w_aux)[ns + 11] * 0.0 + 0.5
            s.be.copy_(s.w_aux[ns + 11], s.be.asarray(z))
        (s.rosenbrock if version == 1 else s.rosenbrock_krylov)(2e-9)
    hip.sync()
    gw, cw = interior(hip.to_numpy(g.w), NG), interior(c.w, NG)
    err = rel_maxnorm(gw, cw)
    print(f"[les] integrator {version} with PaSR: {err.max():.2e}", flush=True)
    assert err.max() < 1e-9, err


def test_les_pasr_strang_steps_match_oracle(pkg, h2_tables, hip, cpu):
    g, c = _pair(pkg, h2_tables, hip, cpu, (8, 6, 5), seed=17, numerics=_num(pkg), temperature=1500.0,
                 perturb_y=0.3)
    assert not g.fused_supported()
    for s in (g, c):
        s.compute_aux()
    dt = c.compute_dt(0.5)
    out = []
    for s in (g, c):
        s.dt = dt
        for k in range(3):
            print(f"[les] {s.be.name} step {k}", flush=True)
            s.advance()
        s.compute_aux(central=1, ghost=0)
        s.be.sync()
        out.append((interior(s.be.to_numpy(s.w), NG).copy(), interior(s.be.to_numpy(s.w_aux), NG)[h2_tables.n_species + 4]))
    (gw, gt), (cw, ct) = out
    err = rel_maxnorm(gw, cw)
    print(f"[les] 3 LES+PaSR steps: species {err.max():.2e}, T {np.abs(gt - ct).max() / ct.max():.2e}", flush=True)
    assert err.max() < 1e-9, err
    assert np.abs(gt - ct).max() / ct.max() < 1e-9
