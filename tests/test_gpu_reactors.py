"""GPU parity of the batched reactor integrator (streams_reactor_batch) against the oracle's
per-reactor Rosenbrock restatement (orc_rosenbrock_batch, rosenbrock_kernel of
kernels_gpu.F90:8928-9193) on the same seeded reactors of uneven stiffness.

Tolerance (north_star): species and temperature after integration within 1e-9 relative
(max-norm of each conserved variable over the batch; temperature relative per reactor).
"""
import numpy as np
import pytest

import oracle_backend
import reactor_util

pytestmark = pytest.mark.gpu
TOL = 1e-9


@pytest.fixture(scope="module")
def hip(pkg):
    return pkg.solver.HipBackend()


def _run_both(pkg, hip, oracle, tables, n, seed, dt, t_range=(800.0, 2500.0)):
    w, tg, info = pkg.cases.reactor_batch(tables, n, seed=seed, t_range=t_range)
    dtt = np.full(n, dt)
    # oracle
    wc = np.ascontiguousarray(w.copy())
    tc = tg.copy()
    R = reactor_util.ReactorBatch(oracle, tables, tol=1e-8, maxsteps=10000, maxtry=40)
    na_c, nr_c = R.integrate(wc, tc, dtt, dtt)
    # GPU
    G = pkg.reactors.ReactorBatchGPU(hip, tables, tol=1e-8)
    wd, td = hip.asarray(w), hip.asarray(tg)
    dd = hip.asarray(dtt)
    na = hip.zeros((n,), "i4")
    nr = hip.zeros((n,), "i4")
    print(f"[reactors] n={n} seed={seed} ns={tables.n_species}", flush=True)
    G.integrate(wd, td, dd, dd, na, nr)
    hip.sync()
    return (hip.to_numpy(wd), hip.to_numpy(td), hip.to_numpy(na), hip.to_numpy(nr)), (wc, tc, na_c, nr_c)


def _check(g, c):
    wg, tgg, na, nr = g
    wc, tc, na_c, nr_c = c
    for m in range(wc.shape[0]):
        scale = np.abs(wc[m]).max()
        if scale > 0:
            assert np.abs(wg[m] - wc[m]).max() <= TOL * scale, m
    assert (np.abs(tgg - tc) <= TOL * tc).all()
    same = (na == na_c).mean()
    assert same >= 0.99, same
    assert (na >= 1).all()


def test_reactor_batch_h2air_matches_oracle(pkg, hip, oracle):
    tables = pkg.cases.h2_air_mechanism().tables()
    g, c = _run_both(pkg, hip, oracle, tables, 1000, seed=7, dt=1e-6)
    _check(g, c)


def test_reactor_batch_keromnes_matches_oracle(pkg, hip, oracle):
    tables = pkg.mechanism.Mechanism(str(oracle_backend.ROOT / "tests/golden/mechanisms/keromnes2013_14sp.yaml")
                                     ).tables()
    g, c = _run_both(pkg, hip, oracle, tables, 300, seed=3, dt=2e-6, t_range=(900.0, 2000.0))
    _check(g, c)


def test_reactor_batch_ignition_against_cantera(pkg, hip):
    """The Keromnes constant-volume reactor of the reference's batch-reactor example on the
    GPU: 1000 output intervals of 2e-6 s, each advanced as the reference's Strang step
    (two rosenbrock calls of dt/2 per 2e-7 s step); T within 1 K of Cantera."""
    tables = pkg.mechanism.Mechanism(str(oracle_backend.ROOT / "tests/golden/mechanisms/keromnes2013_14sp.yaml")
                                     ).tables()
    gold = np.loadtxt(oracle_backend.ROOT / "tests/golden/batch_reactor_cantera.dat")
    y0 = dict(H2=0.01495, O2=0.1149, H2O=0.00375, N2=0.58285, AR=0.00635, CO=0.27715)
    y = np.array([y0.get(nm.upper(), 0.0) for nm in tables.names])
    rho = 101325.0 / (tables.rmix(y) * 1000.0)
    w = np.zeros((tables.n_species + 4, 1))
    w[:tables.n_species, 0] = rho * y
    w[-1, 0] = rho * tables.e_from_T(1000.0, y)
    G = pkg.reactors.ReactorBatchGPU(hip, tables, tol=1e-8)
    wd, td = hip.asarray(w), hip.asarray(np.array([1000.0]))
    half = hip.asarray(np.array([1e-7]))
    temps = np.empty(1000)
    for row in range(1000):
        for _ in range(20):
            G.integrate(wd, td, half, half)
        wv = hip.to_numpy(wd)[:, 0]
        rho_n = wv[:tables.n_species].sum()
        temps[row] = reactor_util.t_from_e(tables, wv[-1] / rho_n, wv[:tables.n_species] / rho_n)
    assert np.abs(temps - gold[:, 2]).max() < 1.0
