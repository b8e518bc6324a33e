"""Continuous-batching server of 0-D reactor requests on one GPU.

Requests (a batch of 1..N constant-pressure adiabatic reactors of one mechanism, each with
its own initial state and a common horizon) arrive over time.  The server keeps every
admitted reactor's state resident in device memory, in one slot pool per mechanism, and
advances all resident reactors of a pool together, one slice of their own horizon per
tick, with one persistent launch of the batched integrator per pool and tick.  The pools
run on their own streams, so the two mechanisms' launches overlap on the device.  A request
that arrives while others are in flight is admitted at the next tick as soon as its
reactors fit in the pool's free device memory (admission by HBM footprint, first come
first served within each pool), and its slots are freed when it completes, so the batch composition changes
from tick to tick (continuous batching, not a queue of whole batches).

Every reactor's arithmetic is independent of which other reactors share its launch: a
request gives the same result alone and in any mix (each reactor integrates the same
slices of its own horizon).  The mixed-precision integrator is the default; a reactor whose
slice fails (non-finite state or the step limit) is recomputed from the slice's initial
state with the FP64 integrator (fallback), and a sample of requests is shadowed by the
FP64 integrator to report the largest difference.

Ignition delay: the first slice end with T > T0 + 400 K, linearly interpolated between slice
ends (the resolution is horizon / slices).
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import Dict, List, Optional

import numpy as np

from .mechanism import Tables
from .reactors import ROS_DTTRY, ReactorBatchGPU, mixture_r


@dataclasses.dataclass
class Request:
    rid: int
    mech: str
    temperature: np.ndarray      # (n,) initial T
    pressure: np.ndarray         # (n,) Pa, constant
    y: np.ndarray                # (n, N_S) initial mass fractions
    horizon: float               # s
    arrival: float = 0.0         # s after the start of the stream
    # results
    done_at: Optional[float] = None
    t_final: Optional[np.ndarray] = None
    y_final: Optional[np.ndarray] = None
    ignition: Optional[np.ndarray] = None
    fallbacks: int = 0

    @property
    def n(self) -> int:
        return int(self.temperature.size)


def enthalpy_many(tables: Tables, T: np.ndarray, Y: np.ndarray) -> np.ndarray:
    """Mixture enthalpy h = e + R T for many states."""
    return tables.e_from_T_many(T, Y) + mixture_r(tables, Y) * T


def temperature_from_h(tables: Tables, h: np.ndarray, Y: np.ndarray, t_guess: np.ndarray) -> np.ndarray:
    """Newton inversion of the mixture enthalpy on the host for many states (results read
    back from the device: the integrator's own temperature output belongs to its last
    right-hand side evaluation)."""
    T = np.array(t_guess, dtype=np.float64)
    r = mixture_r(tables, Y)
    act = np.arange(T.size)            # states still iterating: each stops on its own test, so
    for _ in range(50):                # a state's result does not depend on the others in the call
        if act.size == 0:
            break
        Ta, Ya = T[act], Y[act]
        f = tables.e_from_T_many(Ta, Ya) + r[act] * Ta - h[act]
        dT = f / tables.cp_from_T_many(Ta, Ya)
        T[act] = Ta - dT
        act = act[~(np.abs(dT) <= 1e-11 * T[act])]
    return T


class Pool:
    """Device-resident slots of one mechanism."""

    BYTES_PER_SLOT_EXTRA = 7 * 8 + 2 * 4   # tg, first step, slice, pressure, time done, next step, copy; counters

    def __init__(self, backend, tables: Tables, capacity: int, slices: int, mixed: bool, budget: int = 0):
        import torch
        self.torch = torch
        self.be = backend
        self.tb = tables
        self.ns = tables.n_species
        self.nv = self.ns + 4
        self.cap = capacity
        self.slices = slices
        self.mixed = mixed
        self.budget = budget                 # integrator steps per reactor and tick (0: whole slices)
        self.gpu = ReactorBatchGPU(backend, tables, tol=1e-8)
        self.gpu_fp64 = self.gpu
        self.stream = torch.cuda.Stream(device=backend.device)          # default priority
        self.work = backend.zeros((2,), "i4")
        dev = backend.device
        f8 = torch.float64
        self.w = torch.zeros((self.nv, capacity), dtype=f8, device=dev)
        self.tg = torch.zeros(capacity, dtype=f8, device=dev)
        self.first = torch.zeros(capacity, dtype=f8, device=dev)
        self.step = torch.zeros(capacity, dtype=f8, device=dev)
        self.pres = torch.zeros(capacity, dtype=f8, device=dev)
        self.tdone = torch.zeros(capacity, dtype=f8, device=dev)      # time reached in the current slice
        self.dtn = torch.zeros(capacity, dtype=f8, device=dev)        # next step size (0: start with first)
        self.nacc = torch.zeros(capacity, dtype=torch.int32, device=dev)
        # host bookkeeping
        self.free = list(range(capacity - 1, -1, -1))
        self.owner: Dict[int, tuple] = {}      # slot -> (request, index in request)
        self.slice_no = np.zeros(capacity, dtype=np.int64)
        self.t0 = np.zeros(capacity)
        self.hist_prev = np.zeros(capacity)
        # (request id, reactor index) pairs whose mixed-precision result is treated as failed
        # on every tick, so that they take the FP64 fallback (a test hook)
        self.force_fallback = set()

    def bytes_per_reactor(self) -> int:
        return 8 * self.nv + self.BYTES_PER_SLOT_EXTRA

    def fits(self, req: Request) -> bool:
        return len(self.free) >= req.n

    def admit(self, req: Request):
        torch = self.torch
        slots = [self.free.pop() for _ in range(req.n)]
        h = enthalpy_many(self.tb, req.temperature, req.y)
        wv = np.zeros((self.nv, req.n))
        wv[:self.ns] = req.y.T
        wv[self.ns + 3] = h
        with torch.cuda.stream(self.stream):
            idx = torch.tensor(slots, dtype=torch.int64, device=self.be.device)
            self.w[:, idx] = torch.from_numpy(wv).to(self.be.device, non_blocking=False)
            self.tg[idx] = torch.from_numpy(req.temperature.copy()).to(self.be.device)
            self.pres[idx] = torch.from_numpy(req.pressure.copy()).to(self.be.device)
            self.first[idx] = ROS_DTTRY
            self.step[idx] = req.horizon / self.slices
            self.tdone[idx] = 0.0
            self.dtn[idx] = 0.0
        req.ignition = np.full(req.n, np.nan)
        req._slots = slots
        req._remaining = req.n
        for k, s in enumerate(slots):
            self.owner[s] = (req, k)
            self.slice_no[s] = 0
            self.t0[s] = req.temperature[k]
            self.hist_prev[s] = req.temperature[k]

    def active(self) -> List[int]:
        return sorted(self.owner)

    def launch(self, slots: List[int]):
        """Advance the given slots (asynchronously on the pool's stream): by one slice, or with
        a step budget by at most `budget` integrator steps within their current slice."""
        torch = self.torch
        if not slots:
            return None
        # the slot arrays were last written on the caller's stream (finish)
        self.stream.wait_stream(torch.cuda.current_stream(self.be.device))
        with torch.cuda.stream(self.stream):
            idx = torch.tensor(slots, dtype=torch.int64, device=self.be.device)
            w = self.w.index_select(1, idx).contiguous()
            tg = self.tg.index_select(0, idx).contiguous()
            first = self.first.index_select(0, idx).contiguous()
            step = self.step.index_select(0, idx).contiguous()
            pres = self.pres.index_select(0, idx).contiguous()
            nacc = torch.zeros(len(slots), dtype=torch.int32, device=self.be.device)
            job = dict(idx=idx, slots=slots, w=w, tg=tg, w0=w.clone(), t0=tg.clone(), first=first, step=step,
                       pres=pres, nacc=nacc)
            if self.budget > 0:
                job["td"] = self.tdone.index_select(0, idx).contiguous()
                job["dn"] = self.dtn.index_select(0, idx).contiguous()
                job["td0"], job["dn0"] = job["td"].clone(), job["dn"].clone()
                self.gpu.integrate(w, tg, first, step, nacc, None, stream=self.stream.cuda_stream, mixed=self.mixed,
                                   pressure=pres, work=self.work, t_done=job["td"], dt_next=job["dn"],
                                   maxsteps=self.budget)
            else:
                self.gpu.integrate(w, tg, first, step, nacc, None, stream=self.stream.cuda_stream, mixed=self.mixed,
                                   pressure=pres, work=self.work)
        return job

    def finish(self, job, now: float, completed: List[Request]):
        """Collect one launch: fallback for failed reactors, write back, advance slices and
        retire finished requests."""
        torch = self.torch
        if job is None:
            return
        self.stream.synchronize()
        w = job["w"]
        budgeted = "td" in job
        bad = ~torch.isfinite(w).all(dim=0)
        if not budgeted:
            bad = bad | (job["nacc"] >= self.gpu.maxsteps)
        if self.force_fallback:
            forced = [c for c, s in enumerate(job["slots"])
                      if (self.owner[s][0].rid, self.owner[s][1]) in self.force_fallback]
            if forced:
                bad[torch.tensor(forced, dtype=torch.int64, device=bad.device)] = True
        nbad = int(bad.sum().item())
        if nbad and self.mixed:
            bi = torch.nonzero(bad).flatten()
            sel = lambda a: a.index_select(0, bi).contiguous()           # noqa: E731
            # gathers and the fallback launch on the pool's stream, after the caller's
            # stream has produced `bad` / `bi`
            self.stream.wait_stream(torch.cuda.current_stream(self.be.device))
            with torch.cuda.stream(self.stream):
                wb = job["w0"].index_select(1, bi).contiguous()
                tb_ = sel(job["t0"])
                if budgeted:
                    td, dn = sel(job["td0"]), sel(job["dn0"])
                    self.gpu_fp64.integrate(wb, tb_, sel(job["first"]), sel(job["step"]), None, None,
                                            stream=self.stream.cuda_stream, mixed=False, pressure=sel(job["pres"]),
                                            work=self.work, t_done=td, dt_next=dn, maxsteps=self.budget)
                else:
                    self.gpu_fp64.integrate(wb, tb_, sel(job["first"]), sel(job["step"]), None, None,
                                            stream=self.stream.cuda_stream, mixed=False, pressure=sel(job["pres"]),
                                            work=self.work)
                w[:, bi] = wb
                job["tg"][bi] = tb_
                if budgeted:
                    job["td"][bi] = td
                    job["dn"][bi] = dn
            self.stream.synchronize()
        idx = job["idx"]
        self.w[:, idx] = w
        self.tg[idx] = job["tg"]
        if budgeted:
            slice_done = (job["td"] >= job["step"])
            # a finished slice starts the next one afresh (time 0, first step), as whole-slice
            # ticks do; an unfinished one resumes where it stopped
            self.tdone[idx] = torch.where(slice_done, torch.zeros_like(job["td"]), job["td"])
            self.dtn[idx] = torch.where(slice_done, torch.zeros_like(job["dn"]), job["dn"])
            done_h = slice_done.cpu().numpy()
        else:
            done_h = np.ones(len(job["slots"]), dtype=bool)
        wh = w.cpu().numpy()
        tgh = job["tg"].cpu().numpy()
        badh = bad.cpu().numpy()
        ys = np.maximum(wh[:self.ns], 0.0).T
        ys = ys / ys.sum(axis=1, keepdims=True)
        cols = np.nonzero(done_h)[0]
        temps = np.full(len(job["slots"]), np.nan)
        if cols.size:
            temps[cols] = temperature_from_h(self.tb, wh[self.ns + 3, cols], ys[cols], tgh[cols])
        for c, s in enumerate(job["slots"]):
            req, k = self.owner[s]
            if badh[c]:
                req.fallbacks += 1
            if not done_h[c]:
                continue
            self.slice_no[s] += 1
            tt = temps[c]
            thr = self.t0[s] + 400.0
            if np.isnan(req.ignition[k]) and tt > thr:
                dt = req.horizon / self.slices
                t_hi = self.slice_no[s] * dt
                frac = (thr - self.hist_prev[s]) / (tt - self.hist_prev[s])
                req.ignition[k] = t_hi - dt + frac * dt
            self.hist_prev[s] = tt
            if self.slice_no[s] == self.slices:
                if req.t_final is None:
                    req.t_final = np.zeros(req.n)
                    req.y_final = np.zeros((req.n, self.ns))
                req.t_final[k] = tt
                req.y_final[k] = ys[c]
                req._remaining -= 1
                del self.owner[s]
                self.free.append(s)
                if req._remaining == 0:
                    req.done_at = now
                    completed.append(req)


class ReactorServer:
    def __init__(self, backend, mechanisms: Dict[str, Tables], hbm_bytes: int = 1 << 30, slices: int = 20,
                 budget: int = 0,
                 mixed: bool = True):
        self.be = backend
        self.pools = {}
        for name, tb in mechanisms.items():
            per = 8 * (tb.n_species + 4) + Pool.BYTES_PER_SLOT_EXTRA
            cap = max(1, int(hbm_bytes // len(mechanisms) // per))
            self.pools[name] = Pool(backend, tb, cap, slices, mixed, budget)

    def run(self, requests: List[Request], realtime: bool = True, max_seconds: float = 600.0):
        """Serve `requests` (sorted by arrival).  realtime=True admits a request only once
        the wall clock passes its arrival time; False admits everything at once."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: List[Request] = []
        completed: List[Request] = []
        start = time.perf_counter()
        ticks = 0
        while (pending or waiting or any(p.owner for p in self.pools.values())):
            now = time.perf_counter() - start
            if now > max_seconds:
                break
            while pending and (not realtime or pending[0].arrival <= now):
                waiting.append(pending.pop(0))
            waiting = admit_fifo_per_pool(waiting, self.pools)
            jobs = {name: p.launch(p.active()) for name, p in self.pools.items()}
            if all(j is None for j in jobs.values()):
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                continue
            for name, p in self.pools.items():
                p.finish(jobs[name], time.perf_counter() - start, completed)
            ticks += 1
        wall = time.perf_counter() - start
        lat = np.array([r.done_at - r.arrival for r in completed]) if completed else np.zeros(1)
        n_reactors = sum(r.n for r in completed)
        return {"requests": len(completed), "reactors": n_reactors, "wall_s": wall, "ticks": ticks,
                "reactor_solves_per_s": n_reactors / wall if wall > 0 else 0.0,
                "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
                "fallback_reactors": int(sum(r.fallbacks for r in completed)), "completed": completed}


def admit_fifo_per_pool(waiting: List[Request], pools) -> List[Request]:
    """Admit waiting requests first come, first served within each mechanism's pool: a
    request that does not fit its pool blocks the later requests of that pool only, never
    those of another pool.  Returns the requests still waiting, in arrival order."""
    blocked = set()
    still = []
    for r in waiting:
        pool = pools[r.mech]
        if r.mech not in blocked and pool.fits(r):
            pool.admit(r)
        else:
            blocked.add(r.mech)
            still.append(r)
    return still


def poisson_requests(mechanisms: Dict[str, Tables], rate: float, duration: float, seed: int = 2024,
                     max_reactors: int = 4096, horizon=(1e-5, 1e-2), t0=(900.0, 1800.0), phi=(0.3, 3.0),
                     pressure: float = 101325.0, fuels: Optional[Dict[str, dict]] = None) -> List[Request]:
    """Seeded Poisson stream: inter-arrival times exponential at `rate`; each request picks
    a mechanism uniformly, 1..max_reactors reactors (log-uniform), per-reactor T0 uniform
    and equivalence ratio log-uniform, and a horizon log-uniform on `horizon`."""
    rng = np.random.default_rng(seed)
    names = sorted(mechanisms)
    fuels = fuels or {}
    out, t, rid = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        mech = names[rng.integers(len(names))]
        tb = mechanisms[mech]
        n = int(round(math.exp(rng.uniform(0.0, math.log(max_reactors)))))
        T = rng.uniform(t0[0], t0[1], n)
        ph = np.exp(rng.uniform(math.log(phi[0]), math.log(phi[1]), n))
        fuel = fuels.get(mech, {"H2": 1.0})
        y = premixed(tb, ph, fuel)
        hz = math.exp(rng.uniform(math.log(horizon[0]), math.log(horizon[1])))
        out.append(Request(rid, mech, T, np.full(n, pressure), y, hz, arrival=t))
        rid += 1
    return out


def premixed(tables: Tables, phi: np.ndarray, fuel: dict) -> np.ndarray:
    """Mass fractions of fuel/air mixtures at equivalence ratios phi (air O2:N2 = 1:3.76)."""
    names = [n.upper() for n in tables.names]
    comp = {"H2": {"H": 2}, "CO": {"C": 1, "O": 1}, "CH4": {"C": 1, "H": 4}}
    o2_per_fuel = sum(v * (comp[k].get("C", 0) + comp[k].get("H", 0) / 4.0 - comp[k].get("O", 0) / 2.0)
                      for k, v in fuel.items())
    mw = tables.mw
    out = np.zeros((phi.size, tables.n_species))
    for q, p in enumerate(phi):
        x = np.zeros(tables.n_species)
        for k, v in fuel.items():
            x[names.index(k)] += v
        x[names.index("O2")] += o2_per_fuel / p
        x[names.index("N2")] += 3.76 * o2_per_fuel / p
        y = x * mw
        out[q] = y / y.sum()
    return out
