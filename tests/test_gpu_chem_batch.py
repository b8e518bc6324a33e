"""GPU tests of the field solver's batched implicit chemistry (Numerics.chem_batch_form):
the interior cells of a field integrated as one batch of constant-volume reactors
(streams_reactor_batch) in the Strang half-steps of advance_solution
(multideal_gpu.F90:1853-1931).

* form 0 (FP64 in the reference's arithmetic) against the oracle's field rosenbrock_kernel
  (kernels_gpu.F90:8928-9193), over one chemistry interval and over whole Strang steps:
  species and T within 1e-9 relative, the tolerance of the other integrator tests;
* form 1 (mixed precision) against form 0 over the same Strang steps, on the 9-species
  H2-air and the 14-species Keromnes-2013 mechanisms: T within 1e-6 relative (the
  mixed-precision target for final T) and the species on the mass-fraction scale,
  |d(rho Y_s)| / max(rho) within 1e-6.  Trace radicals differ by more than that relative to
  their own maxima: the two forms take different step sequences, each within the
  integrator's tolerance.
"""
import numpy as np
import pytest

from parity_util import box_state, interior, rel_maxnorm

pytestmark = pytest.mark.gpu
NG = 3
SHAPE = (10, 8, 6)


@pytest.fixture(scope="module")
def hip(pkg):
    return pkg.solver.HipBackend()


@pytest.fixture(scope="module")
def cpu(oracle):
    import oracle_backend
    return oracle_backend.OracleBackend(oracle)


def _num(pkg, form):
    return pkg.solver.Numerics(enable_chemistry=2, rosenbrock_version=1, operator_splitting=2, chem_batch_form=form)


def _solver(pkg, tables, be, form, seed=5):
    w, tg, info = box_state(pkg, tables, SHAPE, seed=seed, temperature=1500.0, perturb_y=0.4)
    s = pkg.solver.Solver(be, tables, *SHAPE, NG, [info["length"]] * 3, numerics=_num(pkg, form),
                          rho0=info["rho0"], u0=info["u_rms"])
    s.set_state(w, tg)
    s.compute_aux()
    return s


def _state(s, ns):
    s.be.sync()
    w = interior(s.be.to_numpy(s.w), NG).copy()
    t = interior(s.be.to_numpy(s.w_aux), NG)[ns + 4].copy()
    return w, t


def test_batch_interval_matches_field_rosenbrock_oracle(pkg, h2_tables, hip, cpu):
    ns = h2_tables.n_species
    g = _solver(pkg, h2_tables, hip, 0)
    c = _solver(pkg, h2_tables, cpu, -1)
    print("[chem_batch] one interval of 2e-8 s, batch form 0 vs the oracle's field kernel", flush=True)
    g.chemistry_batch(2e-8)
    c.rosenbrock(2e-8)
    for s in (g, c):
        s.compute_aux(central=1, ghost=0)       # T from the new energy split, whatever the guess
    (gw, gt), (cw, ct) = _state(g, ns), _state(c, ns)
    err = rel_maxnorm(gw, cw)
    terr = np.abs(gt - ct).max() / ct.max()
    print(f"[chem_batch] species/energy {err.max():.2e}, T {terr:.2e}", flush=True)
    assert err.max() < 1e-9, err
    assert terr < 1e-9


def test_batch_strang_steps_match_field_rosenbrock_oracle(pkg, h2_tables, hip, cpu):
    ns = h2_tables.n_species
    g = _solver(pkg, h2_tables, hip, 0, seed=17)
    c = _solver(pkg, h2_tables, cpu, -1, seed=17)
    dt = c.compute_dt(0.5)
    out = []
    for s in (g, c):
        s.dt = dt
        for k in range(3):
            print(f"[chem_batch] {s.be.name} Strang step {k}", flush=True)
            s.advance()
        s.compute_aux(central=1, ghost=0)
        out.append(_state(s, ns))
    (gw, gt), (cw, ct) = out
    err = rel_maxnorm(gw, cw)
    terr = np.abs(gt - ct).max() / ct.max()
    print(f"[chem_batch] 3 Strang steps, form 0 vs oracle: species {err.max():.2e}, T {terr:.2e}", flush=True)
    assert err.max() < 1e-9, err
    assert terr < 1e-9


@pytest.mark.parametrize("mech", ["h2", "keromnes"])
def test_mixed_batch_strang_steps_match_fp64_batch(pkg, h2_tables, hip, mech):
    tables = h2_tables if mech == "h2" else pkg.cases.keromnes_mechanism().tables()
    ns = tables.n_species
    out = []
    dt = None
    for form in (0, 1):
        s = _solver(pkg, tables, hip, form, seed=23)
        if dt is None:
            dt = s.compute_dt(0.5)
        s.dt = dt
        for k in range(3):
            print(f"[chem_batch] {mech} form {form} Strang step {k}", flush=True)
            s.advance()
        s.compute_aux(central=1, ghost=0)
        out.append(_state(s, ns))
    (w0, t0), (w1, t1) = out
    assert np.isfinite(w1).all() and np.isfinite(t1).all()
    rho_max = w0[:ns].sum(axis=0).max()
    yerr = np.abs(w1[:ns] - w0[:ns]).max(axis=(1, 2, 3)) / rho_max
    terr = np.abs(t1 - t0).max() / t0.max()
    print(f"[chem_batch] {mech}: mixed vs FP64 batch after 3 Strang steps: species (mass-fraction scale) "
          f"{yerr.max():.2e}, per-species relative {rel_maxnorm(w1[:ns], w0[:ns]).max():.2e}, T {terr:.2e}",
          flush=True)
    assert yerr.max() < 1e-6, yerr
    assert terr < 1e-6
    assert rel_maxnorm(w1[ns:], w0[ns:]).max() < 1e-8       # momentum and energy: only through the transport
