"""Pins of the oracle's LES and PaSR routines (oracle/oracle.c) against mathematics that does
not come from the oracle:

* the WALE eddy viscosity of les_wale_mut (kernels_gpu.F90:9921-10047) on linear velocity
  fields, which the central differences differentiate exactly: zero in pure shear (the
  model's defining property, Nicoud and Ducros 1999), rho C_w^2 Delta^2 (2/3)^(1/4) Omega
  in solid-body rotation (S = 0, |S^d| = sqrt(2/3) Omega^2), and the closed form of a
  planar strain; the sensor and div/3 as eval_aux2 gives them;
* eval_velaux + eval_aux_les (:9879-9917, 10049-10161) against eval_aux on the same state:
  the same thermodynamic state, and transport coefficients that differ by exactly
  mu_sgs, mu_sgs c_p / Pr_sgs (c_p from the NASA polynomials on the host) and
  mu_sgs / Sc_sgs;
* eval_chem_aux (:9676-9877): the production rates conserve mass (sum_s wdot_s = 0) and
  equal compute_chemistry's source at the same state; the heat release rate is
  -sum_s h298_s wdot_s / W_s; the PaSR factor is one common factor
  gamma = tau_c / (tau_c + tau_k) in (0, 1] (the partially stirred reactor closure,
  tau_c the geometric mean of the slowest and fastest species times), equal to 1 with
  c_mix = 0;
* the PaSR closure inside the integrators (rosenbrock and rosenbrock_krylov): with
  c_mix = 0 they reproduce the integrators without it bit for bit, and over a step short
  against the chemistry the species change is gamma times the change without it.
"""
import numpy as np
import pytest

import oracle_backend
from parity_util import box_state

NG = 3
LES = dict(enable_les=1, les_model=1, les_pr=0.9, les_sc=0.7, les_c_wale=0.5)


@pytest.fixture(scope="module")
def be():
    return oracle_backend.OracleBackend()


@pytest.fixture(scope="module")
def tb(pkg):
    return pkg.cases.h2_air_mechanism().tables()


def _solver(pkg, be, tb, shape, lengths, **kw):
    num = pkg.solver.Numerics(**kw)
    return pkg.solver.Solver(be, tb, *shape, NG, lengths, numerics=num)


def _coords(s):
    x, y, z = (a.coord for a in s.axes)
    return np.meshgrid(z, y, x, indexing="ij")[::-1]


def test_wale_eddy_viscosity_closed_forms(pkg, be, tb):
    ns = tb.n_species
    shape, lengths = (10, 8, 6), [2e-3, 1.2e-3, 0.9e-3]
    s = _solver(pkg, be, tb, shape, lengths, enable_chemistry=0, ep_order=6, visc_order=6, sensor_type=0, **LES)
    s.eps_sensor = 3.0e6
    X, Y, Z = _coords(s)
    cw = LES["les_c_wale"]
    delta2 = np.prod([L / n for L, n in zip(lengths, shape)]) ** (2.0 / 3.0)
    a, om = 4.0e3, 2.5e3
    sd_strain, s_strain = a * a * np.sqrt(6.0) / 3.0, a * np.sqrt(2.0)
    cases = {
        "shear": ((a * Y, 0 * X, 0 * Z), 0.0),
        "rotation": ((-om * Y, om * X, 0 * Z), cw * cw * delta2 * (2.0 / 3.0) ** 0.25 * om),
        "strain": ((a * X, -a * Y, 0 * Z), cw * cw * delta2 * sd_strain ** 3 / (s_strain ** 5 + sd_strain ** 2.5)),
    }
    rho = 0.4 + 0.1 * np.cos(X / lengths[0] * 2 * np.pi)
    jles = 2 * ns + 13
    for name, (vel, nu_want) in cases.items():
        wa = np.zeros_like(s.w_aux)
        for d in range(3):
            wa[ns + d] = vel[d]
        wa[ns + 10] = rho
        s.w_aux[:] = wa
        be.call("les_wale_mut", ns, *shape, NG, s.nv_aux, 6, s.w_aux, s.cd1, *s.dxi, cw, s.eps_sensor, 0, None)
        mu = s.w_aux[jles, NG:-NG, NG:-NG, NG:-NG]
        want = rho[NG:-NG, NG:-NG, NG:-NG] * nu_want
        np.testing.assert_allclose(mu, want, rtol=1e-11, atol=1e-300, err_msg=name)
        les_duc = s.w_aux[ns + 8].copy()
        les_div = s.w_aux[ns + 9].copy()
        s.w_aux[:] = wa
        s.eval_aux2()
        np.testing.assert_array_equal(les_duc[NG:-NG, NG:-NG, NG:-NG], s.w_aux[ns + 8, NG:-NG, NG:-NG, NG:-NG])
        np.testing.assert_array_equal(les_div[NG:-NG, NG:-NG, NG:-NG], s.w_aux[ns + 9, NG:-NG, NG:-NG, NG:-NG])


def test_les_auxiliaries_add_the_eddy_viscosity(pkg, be, tb):
    ns = tb.n_species
    shape = (12, 10, 8)
    w, tg, info = box_state(pkg, tb, shape, seed=31, temperature=1300.0, perturb_y=0.3)
    lengths = [info["length"]] * 3
    ref = _solver(pkg, be, tb, shape, lengths, enable_chemistry=0)
    ref.set_state(w, tg)
    ref.compute_aux()
    les = _solver(pkg, be, tb, shape, lengths, enable_chemistry=0, **LES)
    les.set_state(w, tg)
    les.compute_aux()            # compute_aux_les
    a, b = ref.w_aux, les.w_aux
    # rho, velocity, Y, and the thermodynamic state: the same arithmetic
    for m in list(range(ns + 3)) + [ns + 10]:
        np.testing.assert_array_equal(b[m], a[m], err_msg=f"aux {m + 1}")
    for m in (ns + 3, ns + 4, ns + 5, ns + 6):
        np.testing.assert_allclose(b[m], a[m], rtol=1e-14, err_msg=f"aux {m + 1}")
    mu_sgs = b[2 * ns + 13]
    inner = (slice(NG, -NG),) * 3
    assert (mu_sgs[inner] > 0).all() and np.isfinite(mu_sgs).all()
    np.testing.assert_allclose(b[ns + 7] - a[ns + 7], mu_sgs, rtol=1e-9, atol=1e-12 * a[ns + 7].max())
    T = a[ns + 4].ravel()
    Y = a[:ns].reshape(ns, -1).T
    cp = tb.cp_from_T_many(T, Y).reshape(mu_sgs.shape)
    np.testing.assert_allclose(b[ns + 12] - a[ns + 12], mu_sgs * cp / LES["les_pr"], rtol=1e-8,
                               atol=1e-12 * a[ns + 12].max())
    for sp in range(ns):
        m = ns + 13 + sp
        np.testing.assert_allclose(b[m] - a[m], mu_sgs / LES["les_sc"], rtol=1e-8, atol=1e-12 * a[m].max())
    # the ghost layers of mu_sgs come from the periodic exchange
    np.testing.assert_array_equal(mu_sgs[NG:-NG, NG:-NG, :NG], mu_sgs[NG:-NG, NG:-NG, -2 * NG:-NG])


def _reacting(pkg, be, tb, shape=(6, 5, 4), seed=41, **kw):
    w, tg, info = box_state(pkg, tb, shape, seed=seed, temperature=1500.0, perturb_y=0.5)
    num = dict(enable_chemistry=2, rosenbrock_version=1, **LES)
    num.update(les_model=2, les_c_yoshi=0.0886, les_c_eps=1.048, les_c_mix=0.1)
    num.update(kw)
    s = _solver(pkg, be, tb, shape, [info["length"]] * 3, **num)
    s.set_state(w, tg)
    s.compute_aux()
    return s, w, tg


def _gamma(tb, s, c_mix):
    """The PaSR factor of every interior cell from the model's definitions, on the rates
    without the closure (wdot, the species times Y_s / |wdot_s / rho|)."""
    ns = tb.n_species
    inner = (slice(NG, -NG),) * 3
    rho = s.w_aux[ns + 10][inner]
    y = s.w_aux[:ns][(slice(None),) + inner]
    wdot = s.w_aux[2 * ns + 14:3 * ns + 14][(slice(None),) + inner]
    rate = np.abs(wdot / rho)
    tau = np.where(rate > 1e-8, y / np.where(rate > 0, rate, 1.0), np.nan)
    tau_f = np.nanmax(np.where(np.isnan(tau), 0.0, tau), axis=0)
    tau_s = np.nanmin(np.where((tau > 0), tau, np.inf), axis=0)
    tau_c = np.sqrt(tau_s * tau_f)
    n = s.num
    dx = np.prod([a.dxi[0] for a in s.axes])
    delta = dx ** (-1.0 / 3.0)
    nu_sgs = s.w_aux[2 * ns + 13][inner] / rho
    eps_sgs = n.les_c_eps * n.les_c_yoshi ** 1.5 * nu_sgs ** 3 / delta ** 4
    tau_k = c_mix * np.sqrt(s.w_aux[ns + 7][inner] / rho / eps_sgs)
    return tau_c / (tau_c + tau_k)


def test_chemistry_diagnostics_and_the_pasr_factor(pkg, be, tb):
    ns = tb.n_species
    s, w, tg = _reacting(pkg, be, tb)
    s.num.les_model = 1                       # rates without the closure first
    s.compute_chem_aux()
    inner = (slice(NG, -NG),) * 3
    wdot = s.w_aux[2 * ns + 14:3 * ns + 14][(slice(None),) + inner].copy()
    scale = np.abs(wdot).max(axis=0)
    assert (scale > 0).all()
    assert (np.abs(wdot.sum(axis=0)) <= 1e-12 * ns * scale).all()          # mass conservation
    hrr = s.w_aux[3 * ns + 14][inner]
    want = -(tb.h298[:, None, None, None] * wdot / tb.mw[:, None, None, None]).sum(axis=0)
    np.testing.assert_allclose(hrr, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    gam_want = _gamma(tb, s, s.num.les_c_mix)
    s.num.les_model = 2
    s.compute_chem_aux()
    w2 = s.w_aux[2 * ns + 14:3 * ns + 14][(slice(None),) + inner]
    live = wdot != 0                           # the inert species have no rate at all
    gam = np.where(live, w2 / np.where(live, wdot, 1.0), np.nan)
    g = np.nanmean(gam, axis=0)
    assert np.all((g > 0) & (g <= 1.0))
    np.testing.assert_allclose(gam[live], np.broadcast_to(g, gam.shape)[live], rtol=1e-13)   # one factor
    np.testing.assert_allclose(g, gam_want, rtol=1e-12)
    # the explicit source at the same state (compute_chemistry subtracts wdot from fl)
    c, _, _ = _reacting(pkg, be, tb)
    c.fl[:] = 0.0
    c.compute_chemistry()
    np.testing.assert_allclose(-c.fl[:ns], wdot, rtol=1e-9, atol=1e-12 * scale.max())
    assert g.min() < 0.999                     # the closure acts on this state
    s.num.les_c_mix = 0.0
    s.compute_chem_aux()
    np.testing.assert_array_equal(s.w_aux[2 * ns + 14:3 * ns + 14][(slice(None),) + inner], wdot)


@pytest.mark.parametrize("version", [1, 2])
def test_pasr_inside_the_integrators(pkg, be, tb, version):
    ns = tb.n_species
    inner = (slice(NG, -NG),) * 3

    def run(c_mix, les_model, dt):
        s, w, tg = _reacting(pkg, be, tb, rosenbrock_version=version, les_c_mix=c_mix, les_model=les_model)
        if version == 2:
            s.w_aux[ns + 11] = 0.5          # Krylov integrates cells with 5e-4 < Z < 1 - 5e-4
        w0 = s.w.copy()
        (s.rosenbrock if version == 1 else s.rosenbrock_krylov)(dt)
        return w0, s.w.copy()

    _, plain = run(0.1, 1, 2e-7)
    _, off = run(0.0, 2, 2e-7)
    np.testing.assert_array_equal(off, plain)           # c_mix = 0: tau_k = 0, factor 1
    # one step short against the fastest chemical time: dw = dt f + O(dt^2 J f)
    dt = 1e-13
    w0, base = run(0.1, 1, dt)
    _, pasr = run(0.1, 2, dt)
    s, _, _ = _reacting(pkg, be, tb, les_model=1)
    s.compute_chem_aux()
    gam = _gamma(tb, s, 0.1)
    d_base = (base - w0)[(slice(0, ns),) + inner]
    d_pasr = (pasr - w0)[(slice(0, ns),) + inner]
    big = np.abs(d_base) > 1e-3 * np.abs(d_base).max(axis=0)
    ratio = d_pasr[big] / d_base[big]
    assert gam.min() < 0.99
    np.testing.assert_allclose(ratio, np.broadcast_to(gam, d_base.shape)[big], rtol=1e-3)
