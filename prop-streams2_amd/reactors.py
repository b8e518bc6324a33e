"""Batched 0-D reactors on the GPU (device-resident state), constant volume or constant
pressure.

Each reactor is one cell of the reference's rosenbrock_kernel (kernels_gpu.F90:8928-9193)
integrated from t = 0 to its own horizon; streams_reactor_batch runs any number of them in
one persistent launch with a device-wide work counter, so reactors of very different
stiffness keep every lane busy.  Three forms of the integrator (``mixed``):

* 0 (False): the plain FP64 integrator in the reference's arithmetic, the yardstick;
* 1 (True): mixed precision on the species block (one-sided species Jacobian, FP32 LU
  refined against the FP64 residual);
* 2: form 1 with an FP64 LU, its like-for-like FP64 counterpart.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import capi
from .mechanism import R_UNIV, Tables

TOL_ITER_NR = 1.0e-12   # src/parameters.F90:49


class ReactorBatchGPU:
    """Tables of one mechanism on the device plus a work counter; `integrate` advances a
    batch of reactors whose state is in device memory (SoA: w[m, b])."""

    def __init__(self, backend, tables: Tables, tol: float = 1e-8, maxsteps: int = 10000, maxtry: int = 40):
        self.be = backend
        self.tb = tables
        self.ns = tables.n_species
        self.nv = self.ns + 4
        self.tol, self.maxsteps, self.maxtry = tol, maxsteps, maxtry
        f = lambda a: backend.asarray(np.asarray(a, dtype=np.float64).T)  # noqa: E731
        self.t_mw = backend.asarray(tables.mw)
        self.t_rgas = backend.asarray(tables.rgas)
        self.t_cp = f(tables.cp_coeff)
        self.t_kin = backend.asarray(capi.kinetics_pack(tables, backend.lib))
        self.t_kc = f(tables.kc_tab)
        self.t_cv = f(tables.cv_coeff)
        self.t_trange = f(tables.trange)
        self.work = backend.zeros((2,), "i4")
        self.wg_per_cu = 0           # 0: the library's default (4); set by tune()
        self.tuning = {}

    def tune(self, w, tg, dttry, time_end, mixed=1, candidates=(1, 2, 4, 8), pressure=None) -> int:
        """Pick the persistent grid's workgroups per CU for batches like this one: one timed
        call per candidate on device copies of the inputs, the fastest kept (and used by
        later calls that pass wg_per_cu = 0).  The results do not depend on the choice:
        lanes pull reactors from the shared counter."""
        torch = self.be.torch
        times = {}
        for per in candidates:
            wc, tc = w.clone(), tg.clone()
            st = torch.cuda.current_stream(self.be.device)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            self.integrate(wc, tc, dttry, time_end, wg_per_cu=per, mixed=mixed, pressure=pressure)
            e1.record(st)
            e1.synchronize()
            times[per] = e0.elapsed_time(e1) / 1e3
        self.wg_per_cu = min(times, key=times.get)
        self.tuning = {"wg_per_cu": self.wg_per_cu, "seconds": times, "reactors": int(w.shape[1])}
        return self.wg_per_cu

    def integrate(self, w, tg, dttry, time_end, nacc=None, nrej=None, wg_per_cu: int = 0, stream=None,
                  mixed=False, pressure=None, work=None, t_done=None, dt_next=None,
                  maxsteps: Optional[int] = None):
        """w (nv, nb) float64, tg (nb) float64, dttry/time_end (nb) float64, all on the device;
        nacc/nrej optional int32 (nb).  Enqueued on `stream` (default: the backend's).
        mixed: 0/False the plain FP64 integrator, 1/True the species-block integrator
        (FP32 factor, FP64 refinement), 2 the species-block integrator in FP64.
        pressure (nb) on the device: constant-pressure reactors with w = (Y, 0, 0, 0, h).
        work: a private counter for launches that may overlap on other streams.
        t_done/dt_next (nb) float64 on the device: resumable call, at most `maxsteps` steps
        per reactor from t_done with the step dt_next, both updated."""
        tb = self.tb
        nb = int(w.shape[1])
        self.be.call("reactor_batch", self.ns, nb, w, tg, dttry, time_end, pressure, self.t_mw, self.t_rgas,
                     self.t_cp, tb.nreactions, self.t_kin, self.t_kc, tb.num_t_tab, tb.t_min_tab, tb.dt_tab, R_UNIV,
                     TOL_ITER_NR, self.t_cv, tb.nsetcv, self.t_trange, tb.indx_cp_l, tb.indx_cp_r,
                     self.maxsteps if maxsteps is None else maxsteps, self.maxtry, self.tol, nacc, nrej, t_done,
                     dt_next, self.work if work is None else work, wg_per_cu or self.wg_per_cu, int(mixed),
                     self.be.stream if stream is None else stream)


ROS_DTTRY = 1e-9   # ros_dttry of the reference's inputs (e.g. examples/batch_reactor/multideal.ini)


def reactor_inputs(tables: Tables, n: int, seed: int = 7, dt: float = 1e-6, dttry: float = ROS_DTTRY):
    """The batched-chemistry workload: n reactors with T log-uniform on [800, 2500] K, p
    log-uniform on [0.5, 20] atm, Y from a seeded Dirichlet(1) over the species, at rest;
    horizon dt.  The first trial step is the reference's ros_dttry input: the reference's
    current code starts from dt_chem instead, and its step control accepts the step that
    ends the interval unconditionally, so a first trial as long as the whole interval
    leaves about a tenth of these very stiff reactors in a wrong state."""
    from .cases import reactor_batch
    w, tg, info = reactor_batch(tables, n, seed=seed)
    return w, tg, np.full(n, dttry), np.full(n, dt), info


def mixture_r(tables: Tables, Y: np.ndarray) -> np.ndarray:
    """R_mix of many states (n, N_S), summed species by species: the same bits for a state
    whatever else is in the array (a BLAS product's blocking depends on the batch)."""
    r = np.zeros(Y.shape[0])
    for s in range(Y.shape[1]):
        r = r + Y[:, s] * tables.rgas[s]
    return r


def temperature_from_e(tables: Tables, e: np.ndarray, Y: np.ndarray, t_guess: np.ndarray) -> np.ndarray:
    """Newton inversion of the mixture internal energy on the host for many reactor states
    (the integrator's own temperature output belongs to its last right-hand side)."""
    T = np.array(t_guess, dtype=np.float64)
    r = mixture_r(tables, Y)
    act = np.arange(T.size)            # each state stops on its own convergence test
    for _ in range(50):
        if act.size == 0:
            break
        Ta, Ya = T[act], Y[act]
        f = tables.e_from_T_many(Ta, Ya) - e[act]
        dT = f / (tables.cp_from_T_many(Ta, Ya) - r[act])
        T[act] = Ta - dT
        act = act[~(np.abs(dT) <= 1e-11 * T[act])]
    return T


def final_temperatures(tables: Tables, w: np.ndarray, t_guess: np.ndarray) -> np.ndarray:
    """T of constant-volume reactor states w (N_S + 4, n), NaN where a state is not finite."""
    ns = tables.n_species
    ok = np.isfinite(w).all(axis=0)
    out = np.full(w.shape[1], np.nan)
    rho = w[:ns, ok].sum(axis=0)
    Y = (w[:ns, ok] / rho).T
    out[ok] = temperature_from_e(tables, w[ns + 3, ok] / rho, Y, t_guess[ok])
    return out
