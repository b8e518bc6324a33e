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
