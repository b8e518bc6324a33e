"""0-D reactors on the oracle (test infrastructure).

A reactor is one cell of conserved variables w = (rho*Y_1..rho*Y_NS, rho*u, rho*v, rho*w,
rho*E) integrated by the oracle's Rosenbrock restatement (orc_rosenbrock_batch, which
follows rosenbrock_kernel, kernels_gpu.F90:8928-9193).  With no transport the reference's
Strang-split step (advance_solution, multideal_gpu.F90:1881-1903) is two rosenbrock calls of
dt/2 each starting with dttry = dt/2, which is what `advance` reproduces.  A
constant-pressure reactor (no counterpart in the reference) has the state (Y_1..Y_NS, 0, 0,
0, h) and integrates at a given pressure (orc_reactor_batch).
"""
import ctypes

import numpy as np

R_UNIV = 8314.46261815324
TOL_ITER_NR = 1.0e-12


def fortran(a, dtype=np.float64):
    """C-contiguous copy whose memory is the Fortran layout of `a`."""
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).T)


class ReactorBatch:
    def __init__(self, lib, tables, tol=1e-8, maxsteps=10000, maxtry=40):
        self.lib = lib.lib if hasattr(lib, "lib") else lib
        self.tb = tables
        self.ns = tables.n_species
        self.nv = self.ns + 4
        self.tol, self.maxsteps, self.maxtry = tol, maxsteps, maxtry
        t = tables
        self._keep = [fortran(t.mw), fortran(t.arr_a), fortran(t.arr_b), fortran(t.arr_ea), fortran(t.tb_eff),
                      fortran(t.falloff_coeffs), fortran(t.reac_ty, np.int32), fortran(t.isRev, np.int32),
                      fortran(t.r_coeffs), fortran(t.p_coeffs), fortran(t.kc_tab), fortran(t.cv_coeff),
                      fortran(t.trange)]
        self._keep += [fortran(t.rgas), fortran(t.cp_coeff)]
        fn = self.lib.orc_reactor_batch
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        fn.restype = I
        fn.argtypes = [I, I, P, P, P, I, P, P, P, P, P, P, P, P, P, P, I, D, D, D, D, P, I, P, I, I, P, P, I, I, D,
                       P, P, P, P, P]
        self.fn = fn

    def state(self, temperature, pressure, y):
        """Conserved variables of quiescent reactors at (T, p, Y); y is (NS, nb)."""
        y = np.asarray(y, dtype=np.float64)
        if y.ndim == 1:
            y = y[:, None]
        temperature = np.broadcast_to(np.asarray(temperature, dtype=np.float64), y.shape[1:]).copy()
        pressure = np.broadcast_to(np.asarray(pressure, dtype=np.float64), y.shape[1:]).copy()
        nb = y.shape[1]
        w = np.zeros((self.nv, nb))
        for b in range(nb):
            rm = self.tb.rmix(y[:, b])
            rho = pressure[b] / (rm * temperature[b])
            w[:self.ns, b] = rho * y[:, b]
            w[self.ns + 3, b] = rho * self.tb.e_from_T(temperature[b], y[:, b])
        return w, temperature

    def state_cp(self, temperature, y):
        """Constant-pressure reactor states (Y, 0, 0, 0, h) at (T, Y); y is (NS, nb)."""
        y = np.asarray(y, dtype=np.float64)
        if y.ndim == 1:
            y = y[:, None]
        temperature = np.broadcast_to(np.asarray(temperature, dtype=np.float64), y.shape[1:]).copy()
        w = np.zeros((self.nv, y.shape[1]))
        w[:self.ns] = y
        for b in range(y.shape[1]):
            w[self.ns + 3, b] = self.tb.e_from_T(temperature[b], y[:, b]) + self.tb.rmix(y[:, b]) * temperature[b]
        return w, temperature

    def integrate(self, w, tg, dttry, time_end, pressure=None):
        """Integrate every reactor over [0, time_end]; w (nv, nb) and tg (nb) updated in place.
        pressure (nb) given: constant-pressure reactors with w = (Y, 0, 0, 0, h).
        Returns (accepted, rejected) step counts per reactor."""
        t = self.tb
        nb = w.shape[1]
        assert w.flags.c_contiguous and w.dtype == np.float64 and w.shape[0] == self.nv
        dttry = np.broadcast_to(np.asarray(dttry, dtype=np.float64), (nb,)).copy()
        tend = np.broadcast_to(np.asarray(time_end, dtype=np.float64), (nb,)).copy()
        nacc = np.zeros(nb, dtype=np.int32)
        nrej = np.zeros(nb, dtype=np.int32)
        k = self._keep
        pres = None if pressure is None else np.broadcast_to(np.asarray(pressure, dtype=np.float64), (nb,)).copy()
        rc = self.fn(self.ns, nb, w.ctypes.data, tg.ctypes.data, k[0].ctypes.data, t.nreactions,
                     *(a.ctypes.data for a in k[1:11]), t.num_t_tab, t.t_min_tab, t.dt_tab, R_UNIV, TOL_ITER_NR,
                     k[11].ctypes.data, t.nsetcv, k[12].ctypes.data, t.indx_cp_l, t.indx_cp_r, dttry.ctypes.data,
                     tend.ctypes.data, self.maxsteps, self.maxtry, self.tol, nacc.ctypes.data, nrej.ctypes.data,
                     None if pres is None else pres.ctypes.data, k[13].ctypes.data, k[14].ctypes.data)
        assert rc == 0
        return nacc, nrej

    def advance(self, w, tg, dt):
        """One reference time step without transport: rosenbrock(dt/2) twice."""
        self.integrate(w, tg, 0.5 * dt, 0.5 * dt)
        self.integrate(w, tg, 0.5 * dt, 0.5 * dt)

    def temperature(self, w, tg):
        """Temperature of each reactor by the host NASA inversion (independent of the oracle)."""
        out = np.empty(w.shape[1])
        for b in range(w.shape[1]):
            rho = w[:self.ns, b].sum()
            y = w[:self.ns, b] / rho
            e = w[self.ns + 3, b] / rho
            out[b] = t_from_e(self.tb, e, y)
        return out


def t_from_h(tables, h, y, lo=200.0, hi=6000.0):
    """Root of h(T) = e(T) + R_mix T = h by bracketing, independent of the oracle's Newton."""
    from scipy.optimize import brentq
    r = tables.rmix(y)
    return brentq(lambda t: tables.e_from_T(t, y) + r * t - h, lo, hi, xtol=1e-12, rtol=1e-15, maxiter=200)


def t_from_e(tables, e, y, lo=200.0, hi=6000.0):
    """Root of e(T) = e by bracketing (scipy brentq on the host NASA polynomials): an
    inversion independent of the Newton iteration under test."""
    from scipy.optimize import brentq
    return brentq(lambda t: tables.e_from_T(t, y) - e, lo, hi, xtol=1e-12, rtol=1e-15, maxiter=200)


class KrylovBatch:
    """The oracle's Rosenbrock-Krylov restatement (orc_rosenbrock_krylov_batch, following
    rosenbrock_krylov_kernel, kernels_gpu.F90:10360-10611) over a flat batch of reactors."""

    def __init__(self, lib, tables, maxsteps=10000):
        self.rb = ReactorBatch(lib, tables)
        self.lib, self.tb, self.maxsteps = self.rb.lib, tables, maxsteps
        fn = self.lib.orc_rosenbrock_krylov_batch
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        fn.restype = I
        fn.argtypes = [I, I, P, P, P, I] + [P] * 10 + [I, D, D, D, D, P, I, P, I, I, P, P, I, P]
        self.fn = fn

    def integrate(self, w, tg, dttry, time_end):
        """w (nv, nb) in place from t = 0 to time_end with first step dttry; tg is the
        temperature guess (not updated).  Returns the attempted steps per reactor."""
        t = self.tb
        nb = w.shape[1]
        assert w.flags.c_contiguous and w.dtype == np.float64
        tg = np.ascontiguousarray(np.broadcast_to(np.asarray(tg, dtype=np.float64), (nb,)))
        dttry = np.broadcast_to(np.asarray(dttry, dtype=np.float64), (nb,)).copy()
        tend = np.broadcast_to(np.asarray(time_end, dtype=np.float64), (nb,)).copy()
        steps = np.zeros(nb, dtype=np.int32)
        k = self.rb._keep
        rc = self.fn(t.n_species, nb, w.ctypes.data, tg.ctypes.data, k[0].ctypes.data, t.nreactions,
                     *(a.ctypes.data for a in k[1:11]), t.num_t_tab, t.t_min_tab, t.dt_tab, R_UNIV, TOL_ITER_NR,
                     k[11].ctypes.data, t.nsetcv, k[12].ctypes.data, t.indx_cp_l, t.indx_cp_r, dttry.ctypes.data,
                     tend.ctypes.data, self.maxsteps, steps.ctypes.data)
        assert rc == 0
        return steps

