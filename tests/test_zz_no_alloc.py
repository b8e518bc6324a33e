"""The allocation-free compute_dt (streams_compute_dt_noalloc) against streams_compute_dt:
the same nine maxima, bit for bit (the same kernels and grid, only the partial results live
in the library's fixed device array instead of stream-ordered scratch), on two calls in a
row, the second on a smaller box, so that the second reads none of the first call's
partial results."""
import numpy as np
import pytest

from parity_util import box_state

pytestmark = pytest.mark.gpu
NG = 3


def _args(s, out):
    tb = s.tb
    return (s.ns, s.nx, s.ny, s.nz, s.ng, s.nv, s.nv_aux, s.dxi[0], s.dxi[1], s.dxi[2], s.dxi2[0], s.dxi2[1],
            s.dxi2[2], s.w, s.w_aux, out, tb.indx_cp_l, tb.indx_cp_r, s.t_cp, tb.nsetcv, s.t_trange, None,
            s.be.stream)


def test_compute_dt_noalloc_matches_compute_dt(pkg, h2_tables):
    hip = pkg.solver.HipBackend()
    for shape, seed in (((64, 40, 36), 3), ((10, 6, 4), 4)):
        w, tg, info = box_state(pkg, h2_tables, shape, seed=seed, temperature=1300.0, perturb_y=0.2)
        s = pkg.solver.Solver(hip, h2_tables, *shape, NG, [info["length"]] * 3, rho0=info["rho0"], u0=info["u_rms"])
        s.set_state(w, tg)
        s.compute_aux()
        a, b = np.zeros(9), np.zeros(9)
        print(f"[no-alloc] compute_dt on {shape}", flush=True)
        hip.call("compute_dt", *_args(s, a))
        hip.call("compute_dt_noalloc", *_args(s, b))
        assert np.all(a > 0.0)
        assert np.array_equal(a, b), (a, b)


def test_compute_dt_noalloc_rejects_bad_sizes(pkg, h2_tables):
    hip = pkg.solver.HipBackend()
    shape = (8, 6, 4)
    w, tg, info = box_state(pkg, h2_tables, shape, seed=5)
    s = pkg.solver.Solver(hip, h2_tables, *shape, NG, [info["length"]] * 3)
    s.set_state(w, tg)
    out = np.zeros(9)
    args = list(_args(s, out))
    args[6] = 2 * s.ns + 12          # nv_aux too small for the auxiliary layout
    with pytest.raises(pkg.capi.StreamsError, match="-2"):
        hip.call("compute_dt_noalloc", *args)
