"""The continuous-batching reactor server on the GPU: a request's result is the same alone
and in a mix of requests of both mechanisms arriving over time, and the mixed-precision
results agree with the FP64 integrator within 1e-6 in final T and ignition
delay."""
import copy

import numpy as np
import pytest

import oracle_backend

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(pkg):
    hip = pkg.solver.HipBackend()
    mechs = {"h2": pkg.cases.h2_air_mechanism().tables(),
             "syngas": pkg.mechanism.Mechanism(
                 str(pkg.cases.mechanism_path("keromnes2013_14sp.yaml"))).tables()}
    fuels = {"h2": {"H2": 1.0}, "syngas": {"H2": 0.5, "CO": 0.5}}
    reqs = pkg.serving.poisson_requests(mechs, rate=400.0, duration=0.05, seed=9, max_reactors=48,
                                        horizon=(2e-5, 2e-4), t0=(1100.0, 1600.0), fuels=fuels)
    assert len({r.mech for r in reqs}) == 2 and len(reqs) >= 8
    return hip, mechs, reqs


def _serve(pkg, hip, mechs, reqs, mixed=True, budget=0):
    srv = pkg.serving.ReactorServer(hip, mechs, hbm_bytes=1 << 24, slices=8, mixed=mixed, budget=budget)
    stats = srv.run(copy.deepcopy(reqs), realtime=True, max_seconds=120)
    return stats


def test_request_same_alone_and_in_the_mix(pkg, setup):
    hip, mechs, reqs = setup
    print(f"[serving] {len(reqs)} requests, {sum(r.n for r in reqs)} reactors", flush=True)
    mix = _serve(pkg, hip, mechs, reqs)
    assert mix["requests"] == len(reqs)
    done = {r.rid: r for r in mix["completed"]}
    for probe in (reqs[0], reqs[len(reqs) // 2], reqs[-1]):
        alone = _serve(pkg, hip, mechs, [probe])["completed"][0]
        got = done[probe.rid]
        np.testing.assert_array_equal(got.t_final, alone.t_final)
        np.testing.assert_array_equal(got.y_final, alone.y_final)
        np.testing.assert_array_equal(np.isnan(got.ignition), np.isnan(alone.ignition))
        np.testing.assert_array_equal(got.ignition[~np.isnan(got.ignition)],
                                      alone.ignition[~np.isnan(alone.ignition)])
    assert mix["latency_p99_s"] >= mix["latency_p50_s"] > 0


def test_mixed_precision_server_against_fp64(pkg, setup):
    hip, mechs, reqs = setup
    a = {r.rid: r for r in _serve(pkg, hip, mechs, reqs, mixed=True)["completed"]}
    b = {r.rid: r for r in _serve(pkg, hip, mechs, reqs, mixed=False)["completed"]}
    assert set(a) == set(b)
    worst_t, worst_ign = 0.0, 0.0
    for rid in a:
        worst_t = max(worst_t, float(np.max(np.abs(a[rid].t_final - b[rid].t_final) / b[rid].t_final)))
        ia, ib = a[rid].ignition, b[rid].ignition
        assert (np.isnan(ia) == np.isnan(ib)).all()
        ok = ~np.isnan(ib)
        if ok.any():
            worst_ign = max(worst_ign, float(np.max(np.abs(ia[ok] - ib[ok]) / ib[ok])))
    print(f"[serving] mixed vs FP64: final T {worst_t:.2e}, ignition delay {worst_ign:.2e}", flush=True)
    assert worst_t <= 1e-6 and worst_ign <= 1e-6


def test_step_budget_gives_the_same_results(pkg, setup):
    """Ticks of at most 25 integrator steps per reactor (resumed from the time and step
    reached) give the same states as whole slices per tick: resuming continues the step
    sequence exactly."""
    hip, mechs, reqs = setup
    a = {r.rid: r for r in _serve(pkg, hip, mechs, reqs, budget=0)["completed"]}
    run_b = _serve(pkg, hip, mechs, reqs, budget=25)
    b = {r.rid: r for r in run_b["completed"]}
    assert set(a) == set(b)
    for rid in a:
        np.testing.assert_array_equal(a[rid].t_final, b[rid].t_final)
        np.testing.assert_array_equal(a[rid].y_final, b[rid].y_final)
    assert run_b["ticks"] > 8


def test_fallback_equals_fp64_run(pkg, setup):
    """A reactor whose mixed-precision result is treated as failed on every tick takes the
    FP64 fallback from the tick's initial state; its final state then equals, bit for bit,
    an FP64-only server's result for the same request."""
    hip, mechs, reqs = setup
    probe = reqs[1]
    forced = {(probe.rid, k) for k in range(0, probe.n, 2)}
    srv = pkg.serving.ReactorServer(hip, mechs, hbm_bytes=1 << 24, slices=8, mixed=True)
    for p in srv.pools.values():
        p.force_fallback = forced
    got = srv.run(copy.deepcopy([probe]), realtime=False, max_seconds=120)["completed"][0]
    ref = _serve(pkg, hip, mechs, [probe], mixed=False)["completed"][0]
    assert got.fallbacks >= len(forced) * 8      # every slice of every forced reactor
    ks = sorted(k for _, k in forced)
    np.testing.assert_array_equal(got.t_final[ks], ref.t_final[ks])
    np.testing.assert_array_equal(got.y_final[ks], ref.y_final[ks])


def test_full_size_requests_against_the_constant_pressure_oracle(pkg, setup, oracle):
    """Requests at the serving workload's full sizes (up to 4096 reactors, horizons up to
    1e-2 s) served with the mixed-precision integrator and a step budget; on 12 sampled
    reactors of every request the oracle's constant-pressure reactor, integrated slice by
    slice as the server does (each slice from the first trial step), gives the reference
    final T and ignition delay.  Over long horizons the step-control decisions of two
    arithmetics part ways (see test_gpu_reactors.py), so the comparison is statistical:
    median final-T deviation <= 1e-7, 95 % of the samples within 1e-6, all within 1e-3;
    ignition delays (interpolated between slice ends, as the server reports them) 95 % within
    1e-4, all within 1e-2; the same reactors ignite."""
    import reactor_util
    hip, mechs, _ = setup
    reqs = pkg.serving.poisson_requests(mechs, rate=80.0, duration=0.05, seed=21, max_reactors=4096,
                                        horizon=(1e-4, 1e-2), t0=(1000.0, 1600.0),
                                        fuels={"h2": {"H2": 1.0}, "syngas": {"H2": 0.5, "CO": 0.5}})
    assert max(r.n for r in reqs) >= 1000 and max(r.horizon for r in reqs) >= 1e-3
    slices = 10
    print(f"[serving] full size: {len(reqs)} requests, {sum(r.n for r in reqs)} reactors, horizons "
          f"{min(r.horizon for r in reqs):.1e}-{max(r.horizon for r in reqs):.1e} s", flush=True)
    srv = pkg.serving.ReactorServer(hip, mechs, hbm_bytes=1 << 28, slices=slices, mixed=True, budget=200)
    st = srv.run(copy.deepcopy(reqs), realtime=True, max_seconds=400)
    done = {r.rid: r for r in st["completed"]}
    assert set(done) == {r.rid for r in reqs}
    rng = np.random.default_rng(5)
    dev_t, dev_i, ign_mismatch = [], [], 0
    for r in reqs:
        tb = mechs[r.mech]
        R = reactor_util.ReactorBatch(oracle, tb, tol=1e-8)
        pick = rng.choice(r.n, size=min(12, r.n), replace=False)
        w, tg = R.state_cp(r.temperature[pick], r.y[pick].T)
        w = np.ascontiguousarray(w)
        prev = r.temperature[pick].copy()
        ign = np.full(pick.size, np.nan)
        dt = r.horizon / slices
        for k in range(slices):
            R.integrate(w, tg, pkg.reactors.ROS_DTTRY, dt, pressure=r.pressure[pick])
            ys = np.maximum(w[:tb.n_species], 0.0)
            ys = ys / ys.sum(axis=0)
            tt = np.array([reactor_util.t_from_h(tb, w[tb.n_species + 3, b], ys[:, b]) for b in range(pick.size)])
            thr = r.temperature[pick] + 400.0
            new = np.isnan(ign) & (tt > thr)
            ign[new] = (k + 1) * dt - dt + (thr[new] - prev[new]) / (tt[new] - prev[new]) * dt
            prev = tt
        got = done[r.rid]
        dev_t.append(np.abs(got.t_final[pick] - tt) / tt)
        gi = got.ignition[pick]
        ign_mismatch += int((np.isnan(gi) != np.isnan(ign)).sum())
        both = ~np.isnan(gi) & ~np.isnan(ign)
        dev_i.append(np.abs(gi[both] - ign[both]) / ign[both])
    dev_t, dev_i = np.concatenate(dev_t), np.concatenate(dev_i)
    print(f"[serving] {dev_t.size} sampled reactors: final T median {np.median(dev_t):.1e}, q95 "
          f"{np.quantile(dev_t, 0.95):.1e}, max {dev_t.max():.1e}; {dev_i.size} ignitions: q95 "
          f"{np.quantile(dev_i, 0.95) if dev_i.size else 0:.1e}, max {dev_i.max() if dev_i.size else 0:.1e}; "
          f"{st['reactor_solves_per_s']:.0f} reactor-solves/s, p99 {st['latency_p99_s']:.2f} s", flush=True)
    assert ign_mismatch == 0
    assert np.median(dev_t) <= 1e-7 and np.quantile(dev_t, 0.95) <= 1e-6 and dev_t.max() <= 1e-3
    assert dev_i.size > 0 and np.quantile(dev_i, 0.95) <= 1e-4 and dev_i.max() <= 1e-2
