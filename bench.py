"""Benchmark: the 256^3 9-species FP64 reacting RHS + explicit chemistry on one MI355X.

Workload: periodic reacting decaying-turbulence cube, lean H2-air (phi = 0.5, 1 atm,
T = 1200 K), Passot-Pouquet solenoidal velocity at M_t = 0.3, KEEP central scheme (WENO
branch off: sensor threshold above the sensor's range), Wray RK3, explicit finite-rate
chemistry.  One step = one full RK3 time step = 3 RHS+chemistry substeps; the metric counts
cell-updates (one cell through one substep).  Inputs are resident in HBM before timing.

Prints one JSON line with the roofline of the dominant entry point (timed live with HIP
events on the launch stream) and the CPU restatement's rate on a bounded sample of the
same workload (rank 0, N = 1).

Besides the 256^3 line, the default run measures the other half of the metric: the
batched stiff chemistry (2^21 reactors, mixed precision, with the FP64 integrators on a
2^18 sample) and the reactor server's throughput and p50/p99 latency on a Poisson stream.

    python bench.py [--gpus 1] [--steps 4] [--warmup 1] [--n 256] [--no-cpu]
                    [--reactors 2097152] [--serving 0.25]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
METRIC = "cell-updates/s (RHS+chem) at {n}^3; reactor-solves/s + p99 latency at 1e-6"


def _pkg():
    if "prop_streams2_amd" in sys.modules:
        return sys.modules["prop_streams2_amd"]
    d = ROOT / "prop-streams2_amd"
    spec = importlib.util.spec_from_file_location("prop_streams2_amd", d / "__init__.py",
                                                  submodule_search_locations=[str(d)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["prop_streams2_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class TimedBackend:
    """Wraps the HIP backend: records HIP events around selected entry points on the
    stream they are launched on (torch's current stream)."""

    def __init__(self, be, names):
        self.be = be
        self.names = set(names)
        self.records = {}

    def __getattr__(self, item):
        return getattr(self.be, item)

    @property
    def stream(self):
        return self.be.stream

    def call(self, name, *args):
        # one key per set of GPU kernels; the tag names the call's mode for its byte count:
        # rhs_finish_next launches the rhs_finish kernels, and an x sweep with init = 2 (fln
        # already set) the same sweep kernel as with init = 1
        if name == "rhs_sweep":
            key = f"rhs_sweep_{args[0]}"
            tag = key + ("_set" if args[-5] == 2 else "")
        elif name == "rhs_finish_next":
            key, tag = "rhs_finish", name
        elif name == "eval_aux2":
            key, tag = name, ("eval_aux2_div" if args[-2] < 0 else name)     # sensor_type < 0: div/3 only
        else:
            key = tag = name
        if key not in self.names:
            return self.be.call(name, *args)
        torch = self.be.torch
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = self.be.call(name, *args)
        e1.record()
        self.records.setdefault(key, []).append((e0, e1, tag))
        return rc

    def summary(self):
        """key -> (total ms, calls, the calls' tags)"""
        out = {}
        for name, evs in self.records.items():
            ms = [a.elapsed_time(b) for a, b, _ in evs]
            out[name] = (sum(ms), len(ms), [t for _, _, t in evs])
        return out


# GPU kernels each timed entry point launches (the fused path, KEEP central, sixth order)
ENTRY_KERNELS = {
    "rhs_finish": ["finish_species_kernel<NS>", "finish_mom_kernel<NS, 3>"],
    "rhs_sweep_1": ["sweep_kernel<NS, 1, 0, 3>"],
    "rhs_sweep_2": ["sweep_kernel<NS, 2, 0, 3>"],
    "rhs_sweep_3": ["sweep_kernel<NS, 3, 0, 3>"],
    "eval_aux": ["eval_aux_kernel<NS, 0>", "eval_transport_kernel<NS>"],
    "eval_aux2": ["eval_aux2_kernel<NS, 3>"],
}
TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_256cube_h2air9.md"


def measured_traffic(entry: str, n: int):
    """HBM bytes per launch of `entry` from the FETCH_SIZE / WRITE_SIZE passes recorded in
    the traffic summary (its ```traffic block); None off the 256^3 workload or when absent."""
    if n != 256 or not TRAFFIC_FILE.exists():
        return None
    text = TRAFFIC_FILE.read_text()
    if "```traffic" not in text:
        return None
    block = text.split("```traffic", 1)[1].split("```", 1)[0]
    for line in block.strip().splitlines():
        name, val = line.split()
        if name == entry:
            return int(val)
    return None


def cpu_baseline(pkg, tables, n_sample: int, steps: int):
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_backend
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    be = oracle_backend.OracleBackend(oracle_backend.load_oracle("liboracle_omp.so"))
    w, tg, info = pkg.cases.turbulence_box(tables, n_sample, seed=2024)
    s = pkg.solver.Solver(be, tables, n_sample, n_sample, n_sample, 3, [info["length"]] * 3,
                          rho0=info["rho0"], u0=info["u_rms"])
    s.set_state(w, tg)
    s.compute_aux()
    s.compute_dt(0.5)
    s.rk_substep(0)                       # warm-up
    times = []
    for q in range(steps):
        t0 = time.perf_counter()
        s.rk_sync()
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    rate = 3 * n_sample ** 3 / med
    return {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{n_sample}^3 box of the same workload, {steps} RK3 steps (3 substeps each), "
                      f"median step; oracle/oracle.c FP64 restatement, OpenMP over z-slabs"}


def copy_probe(gib: float):
    import torch
    n = int(gib * (1 << 30)) // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(json.dumps({"copy_bytes_each_way": 8 * n, "ms": ms, "GBps_read_plus_write": 16 * n / ms / 1e6}))


def reactor_bench(pkg, n: int, n_fp64: int = 262144, reps: int = 1):
    """Batched constant-volume reactors (T log-uniform 800-2500 K, p 0.5-20 atm, Dirichlet
    Y, horizon 1e-6 s, tolerance 1e-8): reactor-solves/s of the mixed-precision species-block
    integrator on all n reactors, and of the plain FP64 integrator (the reference's
    algorithm and arithmetic) and of the FP64 species-block integrator on the first n_fp64
    of them (the per-reactor cost distribution is the same: the reactors are independent
    draws); final-T differences of the mixed and the FP64 species-block forms against the
    plain FP64 one on those n_fp64 reactors (the final states inverted for T on the host);
    the reactors whose state leaves the tables' temperature range are counted and listed."""
    import torch
    tables = pkg.cases.h2_air_mechanism().tables()
    w, tg, dttry, tend, info = pkg.reactors.reactor_inputs(tables, n)
    be = pkg.solver.HipBackend("cuda:0")
    G = pkg.reactors.ReactorBatchGPU(be, tables, tol=1e-8)
    n_fp64 = min(n, n_fp64)
    out = {"reactors": n, "reactors_fp64_sample": n_fp64, "horizon_s": float(tend[0]), "tol": 1e-8,
           "workload": "2^21 independent constant-volume reactors, 9-species H2-air, T log-uniform 800-2500 K, "
                       "p log-uniform 0.5-20 atm, Dirichlet(1) mass fractions, dt = 1e-6 s"
                       if n == 1 << 21 else f"{n} reactors of the batched-chemistry workload"}
    finals = {}
    # the persistent grid's workgroups per CU, tuned on a 2^17-reactor sample of the workload
    # (larger than the GPU holds at one wave per SIMD, so it runs the instance the full batch
    # runs)
    ns_ = min(n, 1 << 17)
    G.tune(be.asarray(np.ascontiguousarray(w[:, :ns_])), be.asarray(tg[:ns_]), be.asarray(dttry[:ns_]),
           be.asarray(tend[:ns_]), mixed=1)
    out["grid_tuning"] = G.tuning
    for key, mode, nb in (("mixed", 1, n), ("fp64", 0, n_fp64), ("fp64_species_block", 2, n_fp64)):
        print(f"[bench] reactors: {key} on {nb}", file=sys.stderr, flush=True)
        d_dt, d_end = be.asarray(dttry[:nb]), be.asarray(tend[:nb])
        na, nr = be.zeros((nb,), "i4"), be.zeros((nb,), "i4")
        times = []
        for rep in range(reps + 1):
            wd, td = be.asarray(np.ascontiguousarray(w[:, :nb])), be.asarray(tg[:nb])
            torch.cuda.synchronize()
            st = torch.cuda.current_stream()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            G.integrate(wd, td, d_dt, d_end, na, nr, mixed=mode)
            e1.record(st)
            torch.cuda.synchronize()
            if rep > 0 or reps == 0:
                times.append(e0.elapsed_time(e1) / 1e3)
        t = sorted(times)[len(times) // 2]
        nacc = be.to_numpy(na)
        wf = be.to_numpy(wd)
        finals[key] = wf[:, :n_fp64]
        bad = np.nonzero(~np.isfinite(wf).all(axis=0))[0]
        out[key] = {"seconds": t, "reactor_solves_per_s": nb / t, "steps_accepted_mean": float(nacc.mean()),
                    "steps_accepted_max": int(nacc.max()), "steps_rejected_total": int(be.to_numpy(nr).sum()),
                    "nonfinite_reactors": int(bad.size), "nonfinite_first": bad[:8].tolist()}
    out["mixed_speedup_vs_fp64"] = out["mixed"]["reactor_solves_per_s"] / out["fp64"]["reactor_solves_per_s"]
    out["mixed_speedup_vs_fp64_species_block"] = (out["mixed"]["reactor_solves_per_s"]
                                                  / out["fp64_species_block"]["reactor_solves_per_s"])
    t0 = pkg.reactors.final_temperatures(tables, finals["fp64"], tg[:n_fp64])
    for key in ("mixed", "fp64_species_block"):
        tk = pkg.reactors.final_temperatures(tables, finals[key], tg[:n_fp64])
        ok = np.isfinite(tk) & np.isfinite(t0)
        dT = np.abs(tk[ok] - t0[ok]) / t0[ok]
        out[f"{key}_vs_fp64_final_T_rel_diff"] = {
            "p50": float(np.median(dT)), "p99": float(np.quantile(dT, 0.99)), "max": float(dT.max()),
            "fraction_above_1e-6": float((dT > 1e-6).mean())}
    return out


def cpu_reactor_baseline(pkg, n: int = 2048):
    """The oracle's plain FP64 Rosenbrock (OpenMP over reactors) on the first n reactors of
    the batched-chemistry workload: reactor-solves/s on the host cores."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_backend
    import reactor_util
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    tables = pkg.cases.h2_air_mechanism().tables()
    w, tg, dttry, tend, info = pkg.reactors.reactor_inputs(tables, n)
    R = reactor_util.ReactorBatch(oracle_backend.load_oracle("liboracle_omp.so"), tables, tol=1e-8)
    wc = np.ascontiguousarray(w)
    t0 = time.perf_counter()
    R.integrate(wc, tg.copy(), dttry, tend)
    el = time.perf_counter() - t0
    return {"value": n / el, "unit": "reactor-solves/s", "cores": threads, "kind": "port",
            "sample": f"first {n} reactors of the batched-chemistry workload; oracle/oracle.c plain FP64 "
                      f"Rosenbrock, OpenMP over reactors"}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def serving_bench(pkg, duration: float, rate: float = 200.0, budget: int = 200):
    """Poisson stream (seeded, `rate` req/s for `duration` s) of constant-pressure reactor
    requests, 1-4096 reactors each (log-uniform), H2-air (9 species) and H2/CO syngas
    (Keromnes 14 species), T0 900-1800 K, phi 0.3-3 (log-uniform), horizon 1e-5..1e-2 s
    (log-uniform), served with continuous batching; throughput and latency percentiles."""
    mechs = {"h2": pkg.cases.h2_air_mechanism().tables(),
             "syngas": pkg.cases.keromnes_mechanism().tables()}
    fuels = {"h2": {"H2": 1.0}, "syngas": {"H2": 0.5, "CO": 0.5}}
    reqs = pkg.serving.poisson_requests(mechs, rate=rate, duration=duration, seed=2024, fuels=fuels)
    be = pkg.solver.HipBackend("cuda:0")
    srv = pkg.serving.ReactorServer(be, mechs, hbm_bytes=8 << 30, slices=10, mixed=True, budget=budget)
    st = srv.run(reqs, realtime=True, max_seconds=240)
    st.pop("completed")
    st.update({"arrival_rate_per_s": rate, "stream_s": duration, "step_budget_per_tick": budget,
               "submitted": len(reqs),
               "submitted_reactors": int(sum(r.n for r in reqs))})
    return st


def c4_bench(pkg, shape=(512, 256, 128), steps: int = 2, warmup: int = 1, explicit: bool = True):
    """The 14-species Keromnes-2013 syngas mixing layer on one rank's 512 x 256 x 128 box, with
    the sensor threshold set so that about 10 % of the x faces take the WENO flux (WENO5, Roe
    splitting): Strang steps (chemistry dt/2, RK3 transport, chemistry dt/2) with the
    interior cells' implicit chemistry run as one mixed-precision reactor batch, in
    cell-updates/s; with explicit, also RK3 steps with the explicit chemistry source."""
    import torch
    tb = pkg.cases.keromnes_mechanism().tables()
    nx, ny, nz = shape
    print(f"[bench] building the {nx}x{ny}x{nz} syngas mixing layer", file=sys.stderr, flush=True)
    w, tg, info = pkg.cases.mixing_layer(tb, nx, ny, nz, seed=2024)
    be = pkg.solver.HipBackend("cuda:0")
    num = pkg.solver.Numerics(weno_scheme=3, enable_chemistry=2, operator_splitting=2, chem_batch_form=1)
    s = pkg.solver.Solver(be, tb, nx, ny, nz, 3, info["length"], numerics=num, rho0=info["rho0"], u0=info["u0"])
    s.set_state(w, tg)
    w0 = s.w.clone()
    del w, tg
    s.compute_aux()
    duc = be.to_numpy(s.w_aux[tb.n_species + 8])[3:-3, 3:-3, 3:-3]
    num.sensor_threshold = pkg.cases.weno_threshold_for_fraction(duc, num.weno_scheme, 0.10)
    del duc
    dt = s.compute_dt(0.5)

    def timed(step):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    print("[bench] Strang steps, batched mixed-precision chemistry", file=sys.stderr, flush=True)
    el = timed(s.advance)
    out = {"workload": f"{nx}x{ny}x{nz} H2/CO-air temporal mixing layers, Keromnes-2013 (14 species, "
                       f"37 reactions), M_c 0.8, WENO5 on ~10 % of the faces, KEEP elsewhere; step = Strang "
                       f"step: implicit chemistry of every cell over dt/2 as one mixed-precision reactor batch, "
                       f"Wray RK3 transport, chemistry over dt/2",
           "cells": nx * ny * nz, "steps": steps, "ms_per_step": 1e3 * el / steps,
           "cell_updates_per_s": 3 * nx * ny * nz * steps / el, "sensor_threshold": num.sensor_threshold,
           "dt_s": dt}
    # one chemistry half-step of the last state in each batch form: the precision speed-up on
    # field cells (form 1 mixed, form 2 FP64 species block, form 0 FP64 reference arithmetic)
    jt = tb.n_species + 4
    w_s, t_s = s.w.clone(), s.w_aux[jt].clone()
    half = {}
    for form in (1, 2, 0):
        s.w.copy_(w_s)
        s.w_aux[jt].copy_(t_s)
        num.chem_batch_form = form
        print(f"[bench] one chemistry half-step, batch form {form}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.chemistry_batch(0.5 * dt)
        torch.cuda.synchronize()
        half[form] = time.perf_counter() - t0
    num.chem_batch_form = 1
    out["chemistry_half_step_ms"] = {"mixed": 1e3 * half[1], "fp64_species_block": 1e3 * half[2],
                                     "fp64_reference": 1e3 * half[0]}
    if explicit:
        print("[bench] RK3 steps, explicit chemistry source", file=sys.stderr, flush=True)
        s.w.copy_(w0)
        num.enable_chemistry = 1
        s.compute_aux()
        s.dt = dt
        el = timed(s.rk_sync)
        out["explicit_chemistry"] = {"ms_per_step": 1e3 * el / steps,
                                     "cell_updates_per_s": 3 * nx * ny * nz * steps / el}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--reactors", type=int, default=1 << 21,
                    help="batched chemistry on this many reactors (0: skip)")
    ap.add_argument("--reactors-fp64", type=int, default=1 << 18,
                    help="reactors of the same workload timed with the FP64 integrators")
    ap.add_argument("--serving", type=float, default=0.25,
                    help="reactor server on a Poisson stream of this many seconds (0: skip)")
    ap.add_argument("--c4", action="store_true",
                    help="also time the 512x256x128 14-species syngas mixing layer with ~10 %% WENO faces")
    ap.add_argument("--c4-only", action="store_true",
                    help="only the 512x256x128 mixing layer (for profiling the WENO pre-pass)")
    ap.add_argument("--reactors-only", action="store_true",
                    help="only the batched-chemistry part (for profiling the reactor kernel)")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a device-to-device copy of this many GiB (5 times): the known-size "
                         "kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU's subdomain; --gpus {args.gpus} is not supported "
              f"(one process drives one GPU; the multi-GPU decomposition is not built)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return

    pkg = _pkg()
    import torch
    if args.c4_only:
        print(json.dumps({"c4_mixing_layer": c4_bench(pkg, steps=args.steps, warmup=args.warmup)}))
        return
    if args.reactors_only:
        print(json.dumps({"batched_chemistry": reactor_bench(pkg, args.reactors, args.reactors_fp64)}))
        return
    tables = pkg.cases.h2_air_mechanism().tables()
    n = args.n
    print(f"[bench] building {n}^3 input", file=sys.stderr, flush=True)
    w, tg, info = pkg.cases.turbulence_box(tables, n, seed=2024)
    be = pkg.solver.HipBackend("cuda:0")
    entries = ["eval_aux", "eval_aux2", "euler_x", "euler_y", "euler_z", "visflx_x", "visflx_y", "visflx_z",
               "visflx_reduced", "compute_chemistry", "init_flux", "update_flux", "update_field", "bcswap",
               "bcextr_var", "rhs_sweep_1", "rhs_sweep_2", "rhs_sweep_3", "rhs_finish"]
    tbe = TimedBackend(be, entries)
    s = pkg.solver.Solver(tbe, tables, n, n, n, 3, [info["length"]] * 3, rho0=info["rho0"], u0=info["u_rms"])
    s.set_state(w, tg)
    del w, tg
    s.compute_aux()
    s.compute_dt(0.5)
    torch.cuda.synchronize()
    print(f"[bench] dt = {s.dt:.3e} s; warmup {args.warmup} step(s)", file=sys.stderr, flush=True)
    tbe.records = {}                       # profile the warm-up steps only, not the set-up
    for _ in range(args.warmup):
        s.rk_sync()
    torch.cuda.synchronize()
    prof = tbe.summary()
    dominant = max(prof, key=lambda k: prof[k][0])
    # timed region: only the dominant entry point carries events
    tbe.names = {dominant}
    tbe.records = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.rk_sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    cells = n ** 3
    value = 3 * cells * args.steps / elapsed
    dom_ms, dom_calls, dom_tags = tbe.summary()[dominant]
    dom_avg_s = dom_ms / dom_calls / 1e3
    rf = pkg.roofline
    # algorithmic bytes per launch, averaged over the timed launches' modes
    dom_bytes = sum(rf.launch_bytes(t, tables.n_species, n, n, n, 3) for t in dom_tags) / dom_calls
    achieved = dom_bytes / dom_avg_s / 1e9
    traffic = [measured_traffic(t, n) for t in dom_tags]
    traffic = None if None in traffic else sum(traffic) / dom_calls
    survey_bytes = rf.substep_bytes_survey(tables.n_species, cells, n)
    breakdown = {k: round(v[0] / v[1], 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}
    modes = {k: sorted(set(v[2])) for k, v in prof.items() if len(set(v[2])) > 1}
    out = {
        "metric": METRIC.format(n=n),
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{n}^3 periodic reacting decaying turbulence, 9-species H2-air (21 reactions), KEEP "
                        f"central, explicit chemistry; step = Wray RK3 step = 3 RHS+chem substeps",
            "cells": cells, "n_species": tables.n_species, "reactions": tables.nreactions, "ng": 3,
            "ep_order": 6, "dt_s": s.dt,
            "substep_effective_GBps": survey_bytes / (elapsed / (3 * args.steps)) / 1e9,
            "entry_ms_per_call": breakdown,
            "entry_modes": modes,
        },
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": rf.HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / rf.HBM_PEAK_GBS,
                     "traffic": traffic,
                     "gpu_kernels": [k.replace("NS", str(tables.n_species)) for k in ENTRY_KERNELS.get(dominant, [])],
                     "avg_launch_ms": dom_avg_s * 1e3, "algorithmic_bytes": int(dom_bytes)},
    }
    if args.reactors > 0:
        print(f"[bench] batched chemistry, {args.reactors} reactors", file=sys.stderr, flush=True)
        out["batched_chemistry"] = reactor_bench(pkg, args.reactors, args.reactors_fp64)
    if args.serving > 0:
        print(f"[bench] reactor server, {args.serving} s stream", file=sys.stderr, flush=True)
        out["serving"] = serving_bench(pkg, args.serving)
    if args.c4:
        out["c4_mixing_layer"] = c4_bench(pkg)
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(pkg, tables, args.cpu_n, args.cpu_steps)
        if args.reactors > 0:
            out["cpu_baseline_reactors"] = cpu_reactor_baseline(pkg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
