"""More pins of the CPU oracle (oracle/oracle.c), for the routines the reactor and flux pins
do not reach, each against mathematics that does not come from the oracle:

* the Ducros sensor and div/3 of eval_aux2 (kernels_gpu.F90:6310-6382): closed forms on
  linear velocity fields (pure dilatation, pure rotation, their combination), which the
  central differences differentiate exactly;
* the non-Laplacian viscous terms of visflx_reduced (:3539-3656): convergence at the order
  of the derivative stencil to the divergence of the full Newtonian stress minus its
  Laplacian part, both formed from their definitions with exact derivatives of plane waves;
* the adiabatic reflecting wall of bcwall_ad (:5749-5852): a tube with walls at both ends
  is the mirror-symmetric half of a periodic tube of twice the length, and its mass flux
  through the walls is zero;
* the SRI fall-off form of the rate constant (:9600-9609): one reaction against the
  textbook formula k = k_inf Pr/(1+Pr) d [a exp(-b/T) + exp(-T/c)]^X T^e,
  X = 1/(1 + log10(Pr)^2) (Stewart, Larson and Golden 1989).
"""
import numpy as np
import pytest

import oracle_backend


@pytest.fixture(scope="module")
def be():
    return oracle_backend.OracleBackend()


def _box(pkg, be, tb, shape, lengths, **kw):
    num = pkg.solver.Numerics(enable_chemistry=0, **kw)
    return pkg.solver.Solver(be, tb, *shape, 3, lengths, numerics=num)


def _coords(s):
    """Cell coordinates of the padded box, each (pz, py, px) broadcast."""
    x, y, z = (a.coord for a in s.axes)
    return np.meshgrid(z, y, x, indexing="ij")[::-1]


# ------------------------------------------------------------------------ Ducros sensor
@pytest.mark.parametrize("sensor_type", [0, 1])
def test_ducros_sensor_closed_forms(pkg, be, sensor_type):
    tb = pkg.cases.h2_air_mechanism().tables()
    ns = tb.n_species
    shape, lengths = (10, 8, 6), [2e-3, 1.2e-3, 0.9e-3]        # different spacings per axis
    s = _box(pkg, be, tb, shape, lengths, sensor_type=sensor_type, visc_order=6)
    eps = s.eps_sensor = 3.0e6
    X, Y, Z = _coords(s)
    a, om = 4.0e3, 2.5e3                                         # strain and rotation rates, 1/s
    cases = {
        # u = a x: div = 3a, no vorticity -> (-div)^2/(div^2 + eps)
        "compression": ((-a * X, -a * Y, -a * Z), 9 * a * a / (9 * a * a + eps), -a),
        "expansion": ((a * X, a * Y, a * Z), 9 * a * a / (9 * a * a + eps), a),
        # solid-body rotation: div = 0, |omega| = 2 om
        "rotation": ((-om * Y, om * X, 0 * Z), 0.0, 0.0),
        # planar: div = -2a, omega_z = 2 om -> 4a^2 / (4 om^2 + 4 a^2 + eps)
        "mixed": ((-a * X - om * Y, om * X - a * Y, 0 * Z), 4 * a * a / (4 * om * om + 4 * a * a + eps),
                  -2 * a / 3),
    }
    for name, (vel, duc, div3) in cases.items():
        wa = np.zeros_like(s.w_aux)
        for d in range(3):
            wa[ns + d] = vel[d]
        s.w_aux[:] = wa
        s.eval_aux2()
        got = s.w_aux[ns + 8, 3:-3, 3:-3, 3:-3]
        gdiv = s.w_aux[ns + 9, 3:-3, 3:-3, 3:-3]
        want = duc
        if sensor_type == 1 and name == "expansion":
            want = 0.0                                             # max(-div/..., 0) clips expansion
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-14, err_msg=name)
        want_div = {"compression": -a, "expansion": a, "rotation": 0.0, "mixed": -2 * a / 3}[name]
        np.testing.assert_allclose(gdiv, want_div, rtol=1e-11, atol=1e-9, err_msg=name)


# -------------------------------------------------------------------- visflx_reduced
class PlaneWaves:
    """f(x) = c0 + sum_q c_q sin(k_q . x + phi_q) with exact first and second derivatives."""

    def __init__(self, c0, terms):
        self.c0, self.terms = c0, terms

    def f(self, X):
        return self.c0 + sum(c * np.sin(sum(k[d] * X[d] for d in range(3)) + ph) for c, k, ph in self.terms)

    def d(self, X, a):
        return sum(c * k[a] * np.cos(sum(k[d] * X[d] for d in range(3)) + ph) for c, k, ph in self.terms)

    def dd(self, X, a, b):
        return sum(-c * k[a] * k[b] * np.sin(sum(k[d] * X[d] for d in range(3)) + ph) for c, k, ph in self.terms)


def test_visflx_reduced_converges_to_the_stress_divergence(pkg, be):
    """fl = -(div tau - div(mu grad u)) in the momentum rows and
    -(u . (div tau - div(mu grad u)) + tau : grad u - mu grad u : grad u) in the energy row,
    tau = mu (grad u + grad u^T - 2/3 div u I), on a periodic box with varying mu, at the
    order of the visc_order = 6 central differences."""
    tb = pkg.cases.h2_air_mechanism().tables()
    ns = tb.n_species
    L = 1e-3
    kk = 2 * np.pi / L
    rng = np.random.default_rng(8)

    def wave(c0, amp):
        return PlaneWaves(c0, [(amp * rng.uniform(0.5, 1.0), kk * rng.integers(-2, 3, 3), rng.uniform(0, 6.28))
                               for _ in range(3)])

    vel = [wave(0.0, 30.0) for _ in range(3)]
    mu = wave(4e-5, 1e-5)

    def exact(X):
        g = [[vel[i].d(X, j) for j in range(3)] for i in range(3)]          # g[i][j] = du_i/dx_j
        div = g[0][0] + g[1][1] + g[2][2]
        ddiv = [sum(vel[i].dd(X, i, j) for i in range(3)) for j in range(3)]
        m, dm = mu.f(X), [mu.d(X, j) for j in range(3)]
        tau = [[m * (g[i][j] + g[j][i] - (2.0 / 3.0 if i == j else 0.0) * div) for j in range(3)] for i in range(3)]
        red = []
        for i in range(3):
            full = sum(dm[j] * (g[i][j] + g[j][i] - (2.0 / 3.0 if i == j else 0.0) * div)
                       + m * (vel[i].dd(X, j, j) + vel[j].dd(X, i, j) - (2.0 / 3.0 if i == j else 0.0) * ddiv[j])
                       for j in range(3))
            lap = sum(dm[j] * g[i][j] + m * vel[i].dd(X, j, j) for j in range(3))
            red.append(full - lap)
        work = sum(tau[i][j] * g[i][j] - m * g[i][j] * g[i][j] for i in range(3) for j in range(3))
        energy = sum(vel[i].f(X) * red[i] for i in range(3)) + work
        return red + [energy], div

    errs = []
    for n in (8, 16, 32):
        s = _box(pkg, be, tb, (n, n, n), [L] * 3, visc_order=6)
        X = _coords(s)
        want, div = exact(X)
        wa = np.zeros_like(s.w_aux)
        for d in range(3):
            wa[ns + d] = vel[d].f(X)
        wa[ns + 7] = mu.f(X)
        wa[ns + 9] = div / 3.0
        s.w_aux[:] = wa
        s.fl[:] = 0.0
        s.visflx(2)
        got = -s.fl[ns:ns + 4]
        e = [np.abs(got[q] - want[q][3:-3, 3:-3, 3:-3]).max() / np.abs(want[q]).max() for q in range(4)]
        errs.append(e)
    errs = np.array(errs)
    rate = np.log2(errs[1] / errs[2])
    assert (rate > 5.5).all(), (errs, rate)
    assert (errs[2] < 1e-4).all(), errs


# ------------------------------------------------------------------------- bcwall_ad
def test_adiabatic_walls_are_mirror_planes(pkg, be):
    """Walls at both ends of an n-cell tube (bcwall_ad on faces 1 and 2) give the same
    convective and viscous divergence as the periodic 2n-cell tube holding the tube's
    state followed by its mirror image (velocity reversed), and no mass crosses the walls:
    the cell sums of the mass rows vanish."""
    tb = pkg.cases.h2_air_mechanism().tables()
    ns = tb.n_species
    n, length = 24, 2.4e-3
    rng = np.random.default_rng(12)
    x = (np.arange(n) + 0.5) / n
    names = [nm.upper() for nm in tb.names]
    y = np.zeros((ns, n))
    y[names.index("O2")], y[names.index("N2")] = 0.233, 0.767
    y[names.index("H2")] = 0.01 * (1 + np.sin(3 * x))
    y = y / y.sum(axis=0)
    temp = 900.0 + 300.0 * np.exp(-((x - 0.3) / 0.15) ** 2)
    pres = 101325.0 * (1 + 0.2 * np.cos(2 * x))
    u = 40.0 * np.sin(np.pi * x) + rng.uniform(-5, 5, n)
    v = 10.0 * np.cos(5 * x)
    w1 = np.zeros((ns + 4, n))
    for i in range(n):
        rho = pres[i] / (tb.rmix(y[:, i]) * temp[i])
        w1[:ns, i] = rho * y[:, i]
        w1[ns, i], w1[ns + 1, i] = rho * u[i], rho * v[i]
        w1[ns + 3, i] = rho * (tb.e_from_T(temp[i], y[:, i]) + 0.5 * (u[i] ** 2 + v[i] ** 2))
    mirror = w1[:, ::-1].copy()
    mirror[ns:ns + 3] = -mirror[ns:ns + 3]
    num = pkg.solver.Numerics(enable_chemistry=0, sensor_threshold=2.0)

    def divergence(wline, bctags, nn, ln):
        s = pkg.solver.Solver(be, tb, nn, 1, 1, 3, [ln, 1.0, 1.0], bctags=bctags, numerics=num)
        wp = np.pad(wline, [(0, 0), (3, 3)], mode="wrap" if bctags[0] == 0 else "constant")
        s.set_state(np.broadcast_to(wp[:, None, None, :], (ns + 4, 7, 7, nn + 6)).copy(), np.full((7, 7, nn + 6), 1000.0))
        s.update_ghost()
        s.bcswap()
        s.compute_aux()
        s.fl[:] = 0.0
        s.euler(0)
        s.visflx(1)
        return s.fl[:, 0, 0, :].copy(), s.w[:, 3, 3, :3].copy()

    # a non-periodic axis of length L has n - 1 intervals (generate_grid_uniform_x), so the
    # walled tube gets L (n - 1) / n to keep the spacing of the periodic one
    half, ghosts = divergence(w1, (8, 8, 0, 0, 0, 0), n, length * (n - 1) / n)
    full, _ = divergence(np.concatenate([w1, mirror], axis=1), (0, 0, 0, 0, 0, 0), 2 * n, 2 * length)
    # bcwall_ad's ghosts: mirror of the first ng cells, velocity reversed
    want = w1[:, 2::-1].copy()
    want[ns:ns + 3] = -want[ns:ns + 3]
    np.testing.assert_array_equal(ghosts, want)
    scale = np.abs(full[:, :n]).max(axis=1, keepdims=True)
    assert (np.abs(half - full[:, :n]) <= 1e-12 * scale).all()
    mass = np.abs(half[:ns]).sum(axis=1)
    assert (np.abs(half[:ns].sum(axis=1)) <= 1e-12 * mass + 1e-300).all()


# ------------------------------------------------------------------------------ SRI
def test_sri_falloff_against_formula(pkg, be):
    """A mechanism whose only active reaction is H + O2 (+M) <=> HO2 (+M) with SRI fall-off
    (irreversible, so the source is the forward rate): fl_H / W_H = k(T, [M]) [H][O2]
    (fl holds -d(rho Y)/dt)."""
    tb = pkg.cases.h2_air_mechanism().tables()
    ns = tb.n_species
    names = [nm.upper() for nm in tb.names]
    ih, io2, iho2 = names.index("H"), names.index("O2"), names.index("HO2")
    nr = tb.nreactions
    lr = [r for r in range(nr) if tb.r_coeffs[r, ih] == 1 and tb.r_coeffs[r, io2] == 1
          and tb.p_coeffs[r, iho2] == 1][0]
    tb.arr_a = tb.arr_a * 0.0                  # every other reaction: plain Arrhenius with A = 0
    tb.reac_ty = np.zeros_like(tb.reac_ty)
    tb.reac_ty[lr] = 4
    tb.isRev = tb.isRev.copy()
    tb.isRev[lr] = 0
    a0, b0, e0 = 6.366e14, -1.72, 2.2e6          # low-pressure limit (m^6/kmol^2/s, -, J/kmol)
    ai, bi, ei = 4.65e9, 0.44, 0.0               # high-pressure limit
    sa, sb, sc, sd, se = 0.45, 797.0, 979.0, 1.2, 0.1
    tb.arr_a[lr] = (a0, ai)
    tb.arr_b = tb.arr_b.copy()
    tb.arr_b[lr] = (b0, bi)
    tb.arr_ea = tb.arr_ea.copy()
    tb.arr_ea[lr] = (e0, ei)
    tb.falloff_coeffs = tb.falloff_coeffs.copy()
    tb.falloff_coeffs[lr] = (sa, sb, sc, sd, se)
    eff = tb.tb_eff[lr]
    rng = np.random.default_rng(2)
    npts = 16
    temps = rng.uniform(700.0, 2600.0, npts)
    pres = np.exp(rng.uniform(np.log(2e3), np.log(5e7), npts))   # spans low to high pressure limits
    y = rng.dirichlet(np.ones(ns), npts).T
    import aux_util
    w = aux_util.quiescent(tb, temps, pres, y)
    s = pkg.solver.Solver(be, tb, npts, 1, 1, 3, [1.0, 1.0, 1.0])
    wp = np.empty((ns + 4, 7, 7, npts + 6))
    wp[:] = np.concatenate([np.repeat(w[:, :1], 3, 1), w, np.repeat(w[:, -1:], 3, 1)], 1)[:, None, None, :]
    s.set_state(wp, np.full(wp.shape[1:], 1500.0))
    s.eval_aux()
    s.fl[:] = 0.0
    s.compute_chemistry()
    got = s.fl[ih, 0, 0, :] / tb.mw[ih]          # fl = -d(rho Y)/dt and H is consumed
    R = 8314.46261815324
    for b in range(npts):
        rho = w[:ns, b].sum()
        conc = w[:ns, b] / tb.mw
        T = s.w_aux[ns + 4, 0, 0, 3 + b]
        assert T == pytest.approx(temps[b], rel=1e-9)
        k0 = a0 * T ** b0 * np.exp(-e0 / (R * T))
        kinf = ai * T ** bi * np.exp(-ei / (R * T))
        m = float(eff @ conc)
        pr = k0 * m / kinf
        xx = 1.0 / (1.0 + np.log10(pr) ** 2)
        f = sd * (sa * np.exp(-sb / T) + np.exp(-T / sc)) ** xx * T ** se
        k = kinf * pr / (1.0 + pr) * f
        assert got[b] == pytest.approx(k * conc[ih] * conc[io2], rel=1e-12), b
        assert rho > 0


# ------------------------------------------------------------------ simpler splitting
def test_simpler_splitting_rk_stages_integrate_the_forcing_exactly(pkg, be):
    """rk_sync(simpler_splitting=1) (multideal_gpu.F90:562-612 with init_flux_simpler_cuf,
    update_simpler_flux_cuf, kernels_gpu.F90:47-66, 3316-3334): on a uniform periodic state
    the transport RHS is exactly zero, so every stage sees the flux -fl_sav.  The Wray
    coefficients sum to gamma_1 + (rho_2 + gamma_2) + (rho_3 + gamma_3) = 1, so the step
    must give w + dt*fl_sav (a sign error, a dropped stage or a swapped coefficient fails)."""
    tb = pkg.cases.h2_air_mechanism().tables()
    shape = (6, 5, 4)
    s = _box(pkg, be, tb, shape, [1e-3] * 3)
    w, tg, _ = pkg.cases.turbulence_box(tb, 6, mach_t=0.0, temperature=900.0)
    w = np.ascontiguousarray(w[:, :shape[2] + 6, :shape[1] + 6, :])
    s.set_state(w, np.ascontiguousarray(tg[:shape[2] + 6, :shape[1] + 6, :]))
    s.compute_aux()
    s.dt = 3.0e-8
    rng = np.random.default_rng(3)
    rho = w[:s.ns, 3, 3, 3].sum()
    # per-variable forcing rates: species gain mass (a clipped negative density would break
    # the identity), momenta and energy either sign
    rate = np.concatenate([rng.uniform(0.1, 1.0, s.ns) * rho, rng.uniform(-1.0, 1.0, 3) * rho * 300.0,
                           rng.uniform(-1.0, 1.0, 1) * w[s.ns + 3, 3, 3, 3]]) * 1e3
    fs = np.ones(s.fl.shape) * rate[:, None, None, None]
    s.fl_sav = be.asarray(fs)
    w0 = s.w.copy()
    s.rk_sync(fused=False, simpler=True)
    got = s.w[:, 3:-3, 3:-3, 3:-3]
    want = w0[:, 3:-3, 3:-3, 3:-3] + s.dt * fs
    scale = np.maximum(np.abs(want).max(axis=(1, 2, 3)), np.abs(s.dt * rate))
    err = np.abs(got - want).max(axis=(1, 2, 3)) / scale
    assert err.max() < 1e-13, err
    # sav_flx keeps the RHS it just computed (zero for the uniform state) and leaves w alone
    w1 = s.w.copy()
    s.fl[...] = 1.0
    s.sav_flx()
    assert np.abs(s.fl_sav).max() < 1e-9 * np.abs(rate).max()
    np.testing.assert_array_equal(s.w[:, 3:-3, 3:-3, 3:-3], w1[:, 3:-3, 3:-3, 3:-3])


def test_krylov_forcing_against_an_independent_ode_solution(pkg, be):
    """The transport forcing of rosenbrock_krylov_kernel (kernels_gpu.F90:10542-10543,
    10657-10658, 10724-10725): with simpler_splitting = 1 the integrator solves
    dw/dt = omega(w) W - fl_sav.  One reacting cell with a forcing that moves each variable
    by a few percent over the step is compared with scipy's Radau solution of that ODE at
    tight tolerance, whose right-hand side is the oracle's explicit chemistry source
    (pinned against Cantera) minus the forcing.  The Krylov method's own error control is
    relative 1e-4, so the bound is 1e-3 of each variable's value (measured: <2e-4); a dropped
    forcing or a wrong sign moves the result by the forcing itself (measured: >1.5e-2)."""
    from scipy.integrate import solve_ivp
    tb = pkg.cases.h2_air_mechanism().tables()
    num = pkg.solver.Numerics(enable_chemistry=2, rosenbrock_version=2)
    s = pkg.solver.Solver(be, tb, 1, 1, 1, 3, [1e-3] * 3, numerics=num)
    y = pkg.cases.mole_to_mass(tb, {"H2": 2.0, "O2": 1.0, "N2": 3.76, "OH": 0.02, "H": 0.01})
    t0, p0 = 1300.0, 2.0e5
    rho = p0 / (tb.rmix(y) * t0)
    wc = np.zeros(s.nv)
    wc[:s.ns] = rho * y
    wc[s.ns + 3] = rho * tb.e_from_T(t0, y)
    dt = 2.0e-6
    forcing = -0.05 * wc / dt * np.linspace(-1.0, 1.0, s.nv)    # dw/dt contribution -fl_sav
    forcing[s.ns: s.ns + 3] = 0.0
    fs = np.ascontiguousarray(-forcing[:, None, None, None])

    def load(wv):
        s.set_state(np.broadcast_to(wv[:, None, None, None], s.w.shape).copy(), np.full(s.w.shape[1:], t0))

    def rhs(_, wv):
        load(wv)
        s.compute_aux()
        s.fl[...] = 0.0
        s.compute_chemistry()
        return -s.fl[:, 0, 0, 0] - fs[:, 0, 0, 0]

    ref = solve_ivp(rhs, (0.0, dt), wc, method="Radau", rtol=1e-10, atol=1e-14 * np.abs(wc).max()).y[:, -1]
    load(wc)
    s.compute_aux()
    s.rosenbrock_krylov(dt, fl_sav=fs)
    got = s.w[:, 3, 3, 3]
    scale = np.maximum(np.abs(ref), 1e-30)
    err = np.abs(got - ref) / scale
    assert err.max() < 1e-3, err
    # the forcing matters at this size: dropping it moves the state well beyond the bound
    load(wc)
    s.compute_aux()
    s.rosenbrock_krylov(dt)
    assert (np.abs(s.w[:, 3, 3, 3] - ref) / scale).max() > 1e-2


# ------------------------------------------------------------- Laplacian viscous form
def test_visflx_laplacian_form_converges_on_a_stretched_grid(pkg, be):
    """visflx_cuf (conservative_viscous = 0, kernels_gpu.F90:3357-3537): fl = -div tau in the
    momentum rows and -(div(tau u) + div(k grad T)) in the energy row, tau = mu (grad u +
    grad u^T - 2/3 div u I), on a periodic box stretched in x and y (x = xi + a sin(2 pi
    xi / L)), so that the second-derivative metric term dcsidx2 counts; k = Pr' mu makes the
    form's conduction (grad mu . grad T + mu lap T) k / mu exact.  Converges at the order of
    the sixth-order stencils."""
    tb = pkg.cases.h2_air_mechanism().tables()
    ns = tb.n_species
    L = 1e-3
    kk = 2 * np.pi / L
    rng = np.random.default_rng(12)

    def wave(c0, amp):
        return PlaneWaves(c0, [(amp * rng.uniform(0.5, 1.0), kk * rng.integers(-2, 3, 3), rng.uniform(0, 6.28))
                               for _ in range(3)])

    vel = [wave(0.0, 30.0) for _ in range(3)]
    mu = wave(4e-5, 1e-5)
    temp = wave(1200.0, 80.0)
    ratio = 1.7                                          # k / mu
    amp = 0.25 * L / (2 * np.pi)

    def stretch(xi):
        return xi + amp * np.sin(2 * np.pi * xi / L)

    def exact(X):
        g = [[vel[i].d(X, j) for j in range(3)] for i in range(3)]
        div = g[0][0] + g[1][1] + g[2][2]
        ddiv = [sum(vel[i].dd(X, i, j) for i in range(3)) for j in range(3)]
        m, dm = mu.f(X), [mu.d(X, j) for j in range(3)]
        tau = [[m * (g[i][j] + g[j][i] - (2.0 / 3.0 if i == j else 0.0) * div) for j in range(3)] for i in range(3)]
        dtau = []
        for i in range(3):
            dtau.append(sum(dm[j] * (g[i][j] + g[j][i] - (2.0 / 3.0 if i == j else 0.0) * div)
                            + m * (vel[i].dd(X, j, j) + vel[j].dd(X, i, j) - (2.0 / 3.0 if i == j else 0.0) * ddiv[j])
                            for j in range(3)))
        work = sum(tau[i][j] * g[i][j] for i in range(3) for j in range(3))
        cond = ratio * sum(dm[j] * temp.d(X, j) + m * temp.dd(X, j, j) for j in range(3))
        energy = sum(vel[i].f(X) * dtau[i] for i in range(3)) + work + cond
        return dtau + [energy], div

    errs = []
    for n in (8, 16, 32):
        s = pkg.solver.Solver(be, tb, n, n, n, 3, [L] * 3, numerics=pkg.solver.Numerics(
            enable_chemistry=0, visc_order=6, conservative_viscous=0), mappings=[stretch, stretch, None])
        X = _coords(s)
        want, div = exact(X)
        wa = np.zeros_like(s.w_aux)
        for d in range(3):
            wa[ns + d] = vel[d].f(X)
        wa[ns + 4] = temp.f(X)
        wa[ns + 7] = mu.f(X)
        wa[ns + 9] = div / 3.0
        wa[ns + 12] = ratio * mu.f(X)
        s.w_aux[:] = wa
        s.fl[:] = 0.0
        s.visflx(0)
        got = -s.fl[ns:ns + 4]
        assert np.all(s.fl[:ns] == 0.0)                  # no species diffusion in this form
        e = [np.abs(got[q] - want[q][3:-3, 3:-3, 3:-3]).max() / np.abs(want[q]).max() for q in range(4)]
        errs.append(e)
    errs = np.array(errs)
    rate = np.log2(errs[1] / errs[2])
    assert (rate > 5.0).all(), (errs, rate)
    assert (errs[2] < 1e-3).all(), errs


# --------------------------------------------------- characteristic boundary treatment
def _argon_tube(pkg, be, n, bctags_nr, amp_sign, u0, eps=1e-5):
    """Argon tube (N_S = 1: c_p = 5R/2, gamma = 5/3 exactly) at rest or moving with u0,
    carrying a small acoustic simple wave p' = eps p0 sin(k x), u' = amp_sign p'/(rho c),
    rho' = p'/c^2 (amp_sign +1 right-running, -1 left-running); returns the solver and the
    exact conserved field and its x derivative as callables of x."""
    mech = pkg.mechanism.Mechanism(str(pkg.cases.mechanism_path("h2_premix_11sp.yaml")), species=["Ar"],
                                   reactions=False)
    tb = mech.tables()
    L, t0, p0 = 0.1, 300.0, 1.0e5
    R = tb.rgas[0]
    rho0 = p0 / (R * t0)
    c0 = np.sqrt(5.0 / 3.0 * R * t0)
    k = 2 * np.pi * 1.5 / L

    def prim(x):
        a = eps * p0 * np.sin(k * x + 0.3)
        p = p0 + a
        rho = rho0 + a / c0 ** 2
        u = u0 + amp_sign * a / (rho0 * c0)
        return rho, u, p

    def cons(x):
        rho, u, p = prim(x)
        tt = np.atleast_1d(p / (rho * R))
        e = tb.e_from_T_many(tt, np.ones((tt.size, 1))).reshape(np.shape(rho))   # NASA e(T), incl. its offset
        return np.stack([rho, rho * u, 0 * rho, 0 * rho, rho * (e + 0.5 * u * u)])

    def flux(x):
        rho, u, p = prim(x)
        wv = cons(x)
        return np.stack([rho * u, rho * u * u + p, 0 * rho, 0 * rho, (wv[4] + p) * u])

    num = pkg.solver.Numerics(enable_chemistry=0)
    s = pkg.solver.Solver(be, tb, n, 1, 1, 3, [L, 1.0, 1.0], bctags=(2, 2, 0, 0, 0, 0), numerics=num,
                          rho0=rho0, u0=c0, bctags_nr=bctags_nr)
    x = s.axes[0].coord
    wv = cons(x)
    wp = np.broadcast_to(wv[:, None, None, :], (5, 7, 7, x.size)).copy()
    s.set_state(wp, np.full(wp.shape[1:], t0))
    s.compute_aux()
    s.fl[:] = 0.0
    s.bc_nr()
    h = 1e-7 * L
    dfdx = lambda xx: (flux(xx + h) - flux(xx - h)) / (2 * h)   # noqa: E731
    return s, x, dfdx


def test_nonreflecting_face_passes_outgoing_and_drops_incoming_waves(pkg, be):
    """bc_nr_lat (kernels_gpu.F90:4330-4479), nr_type 1 at xmax of a tube carrying a small
    acoustic wave: a right-running (outgoing) wave gives the convective term dF/dx of the
    face cell (second-order one-sided difference: the error falls by 4 per doubling); a
    left-running (incoming) wave is removed: the face cell's term is O(eps) of dF/dx."""
    errs = []
    for n in (40, 80, 160):
        s, x, dfdx = _argon_tube(pkg, be, n, (0, 1, 0, 0, 0, 0), +1, 20.0)
        want = dfdx(x[3 + n - 1])
        got = s.fl[:, 0, 0, n - 1]
        errs.append(np.abs(got[[0, 1, 4]] - want[[0, 1, 4]]).max() / np.abs(want[[0, 1, 4]]).max())
        assert np.all(s.fl[:, 0, 0, :n - 1] == 0.0)            # only the face cells
    rate = np.log2(errs[0] / errs[1]), np.log2(errs[1] / errs[2])
    assert min(rate) > 1.8 and errs[2] < 5e-3, (errs, rate)
    s, x, dfdx = _argon_tube(pkg, be, 80, (0, 1, 0, 0, 0, 0), -1, 20.0)
    want = dfdx(x[3 + 79])
    got = s.fl[:, 0, 0, 79]
    assert np.abs(got).max() < 1e-3 * np.abs(want).max()
    # the same wave through a face without the treatment is untouched by bc_nr
    s, _, _ = _argon_tube(pkg, be, 80, (0, 0, 0, 0, 0, 0), -1, 20.0)
    assert np.all(s.fl == 0.0)


def test_reflective_wall_reflects_the_outgoing_acoustic_wave(pkg, be):
    """nr_type 6 at xmin of a tube at rest with a left-running wave reaching the wall: the
    incoming acoustic amplitude is set to the outgoing one, so the two acoustic fields add
    with equal amplitude: the momentum term of the face cell vanishes (u = 0 at rest) while
    the mass term is twice the outgoing wave's."""
    s, x, dfdx = _argon_tube(pkg, be, 80, (6, 0, 0, 0, 0, 0), -1, 0.0)
    want = dfdx(x[3])
    got = s.fl[:, 0, 0, 0]
    assert abs(got[1]) < 1e-3 * abs(want[1])
    assert abs(got[0] - 2.0 * want[0]) < 2e-2 * abs(want[0])
    # nr_type 1 at the same face: the outgoing wave alone
    s, x, dfdx = _argon_tube(pkg, be, 80, (1, 0, 0, 0, 0, 0), -1, 0.0)
    got1 = s.fl[:, 0, 0, 0]
    assert abs(got1[0] - want[0]) < 2e-2 * abs(want[0]) and abs(got1[1] - want[1]) < 2e-2 * abs(want[1])
