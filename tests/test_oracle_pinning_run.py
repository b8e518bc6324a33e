"""Pins of the oracle's run-loop monitors and field copies (oracle/oracle.c) by known answers:

* compute_residual (kernels_gpu.F90:5991-6012): with fln(I_U) = c everywhere the sum is
  n_cells (c/dt)^2, and masked cells drop out of it;
* compute_rho_t_p_minmax (:6038-6073): planted extrema come back, masked ones do not;
* find_boom (:10334-10358): the number of planted NaNs in variable 2 and the first in
  i-fastest order, with the rank offsets of the global indices;
* copy_cpu_gpu / copy_gpu_cpu (base_gpu.F90:1266-1325): the variable-fastest host layout is
  the transpose of the point-fastest device layout, and the two copies are inverse;
* field_interior: the dense interior block and back.
"""
import numpy as np
import pytest

import oracle_backend

NX, NY, NZ, NG, NS = 7, 5, 4, 2, 9
NV = NS + 4


@pytest.fixture(scope="module")
def be():
    return oracle_backend.OracleBackend()


def _mask(on):
    m = np.zeros((NZ + 2 * NG, NY + 2 * NG, NX + 2 * NG), dtype=np.int32)
    for (i, j, k) in on:
        m[k + NG - 1, j + NG - 1, i + NG - 1] = 1
    return m


def test_residual_known_answer(be):
    fln = np.zeros((NV, NZ, NY, NX))
    c, dt = 3.0e-4, 2.0e-7
    fln[NS] = c                                  # I_U
    fln[1] = 7.0                                 # other variables do not enter
    res = np.zeros(1)
    be.call("compute_residual", NX, NY, NZ, NG, NV, fln, dt, res, None, None)
    assert res[0] == pytest.approx(NX * NY * NZ * (c / dt) ** 2, rel=1e-14)
    m = _mask([(1, 1, 1), (7, 5, 4), (3, 2, 2)])
    be.call("compute_residual", NX, NY, NZ, NG, NV, fln, dt, res, m, None)
    assert res[0] == pytest.approx((NX * NY * NZ - 3) * (c / dt) ** 2, rel=1e-14)


def test_extrema_known_answer(be):
    nv_aux = 2 * NS + 13
    wa = np.zeros((nv_aux, NZ + 2 * NG, NY + 2 * NG, NX + 2 * NG))
    jr, jt, jp = NS + 10, NS + 4, NS + 5
    wa[jr], wa[jt], wa[jp] = 1.0, 1000.0, 1.0e5
    wa[jr, NG + 1, NG + 2, NG + 3] = 0.25          # (i, j, k) = (4, 3, 2)
    wa[jt, NG + 3, NG + 4, NG + 6] = 2400.0        # (7, 5, 4)
    wa[jp, NG, NG, NG] = 5.0e4                      # (1, 1, 1)
    wa[jt, 0, 0, 0] = 9999.0                        # a ghost: never looked at
    out = [np.zeros(1) for _ in range(6)]
    be.call("compute_rho_t_p_minmax", NS, NX, NY, NZ, NG, nv_aux, wa, *out, None, None)
    assert [o[0] for o in out] == [0.25, 1.0, 1000.0, 2400.0, 5.0e4, 1.0e5]
    be.call("compute_rho_t_p_minmax", NS, NX, NY, NZ, NG, nv_aux, wa, *out, _mask([(7, 5, 4)]), None)
    assert out[3][0] == 1000.0


def test_find_boom_known_answer(be):
    fln = np.ones((NV, NZ, NY, NX))
    fln[1, 3, 2, 5] = np.nan                        # (6, 3, 4)
    fln[1, 1, 4, 2] = np.nan                        # (3, 5, 2): first in i-fastest order
    fln[3, 0, 0, 0] = np.nan                        # another variable: not searched
    n, ijk = np.zeros(1, dtype=np.int32), np.zeros(3, dtype=np.int32)
    be.call("find_boom", NX, NY, NZ, NG, NV, fln, 1e-7, None, 2, 1, n, ijk, None)
    assert n[0] == 2 and ijk.tolist() == [3 + 2 * NX, 5, 2 + NZ]
    be.call("find_boom", NX, NY, NZ, NG, NV, fln, 1e-7, _mask([(3, 5, 2)]), 0, 0, n, ijk, None)
    assert n[0] == 1 and ijk.tolist() == [6, 3, 4]


def test_field_copies_are_the_layout_transpose(be):
    rng = np.random.default_rng(3)
    shape = (NZ + 2 * NG, NY + 2 * NG, NX + 2 * NG)
    host = rng.standard_normal(shape + (NV,))       # w(nv, i, j, k): variable fastest
    dev = np.zeros((NV,) + shape)
    be.call("copy_cpu_gpu", NX, NY, NZ, NG, NV, host, dev, None)
    assert np.array_equal(dev, np.moveaxis(host, -1, 0))
    back = np.zeros_like(host)
    be.call("copy_gpu_cpu", NX, NY, NZ, NG, NV, dev, back, None)
    assert np.array_equal(back, host)
    dense = np.zeros((NV, NZ, NY, NX))
    be.call("field_interior", NX, NY, NZ, NG, NV, dev, dense, 1, None)
    assert np.array_equal(dense, dev[:, NG:-NG, NG:-NG, NG:-NG])
    dev2 = np.zeros_like(dev)
    be.call("field_interior", NX, NY, NZ, NG, NV, dev2, dense * 2.0, 0, None)
    assert np.array_equal(dev2[:, NG:-NG, NG:-NG, NG:-NG], 2.0 * dense)
    assert not dev2[:, :NG].any()
