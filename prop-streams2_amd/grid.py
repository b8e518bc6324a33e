"""Uniform Cartesian grid, finite-difference coefficients and metric arrays.

Restates the parts of ``src/grid.F90`` the hot path consumes: node coordinates of a
uniform grid with ng ghost layers (``generate_grid_uniform_x/y/z``, grid.F90:569-627),
first/second-derivative coefficient sets (``get_deriv_coeffs``, grid.F90:135-182), the
metric derivatives dx/dxi by the same central differences (``compute_metrics``,
grid.F90:60-133) and the local metric arrays ``dcsidx, dcsidxs, dcsidx2`` etc.
(``set_local_grid``, field.F90:300-347).  Boundary-order reduction
(``correct_bc_order``, multideal.F90:1061-1122) fills ``ep_ord_change``.
"""
from __future__ import annotations

import dataclasses

import numpy as np

D1 = {
    1: [0.5],
    2: [2.0 / 3.0, -1.0 / 12.0],
    3: [0.75, -0.15, 1.0 / 60.0],
    4: [4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0],
}
D2 = {
    1: [-2.0, 1.0],
    2: [-2.5, 4.0 / 3.0, -1.0 / 12.0],
    3: [-245.0 / 90.0, 1.5, -0.15, 1.0 / 90.0],
    4: [-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0],
}


def deriv_coeffs(order: int):
    """(coeff_deriv1(1:4), coeff_deriv2(0:4)) for a central scheme of the given order."""
    c1 = np.zeros(4)
    c2 = np.zeros(5)
    h = order // 2
    c1[: len(D1[h])] = D1[h]
    c2[: len(D2[h])] = D2[h]
    return c1, c2


def coeff_deriv1_table() -> np.ndarray:
    """coeff_deriv1(4,4): column lmax holds the order-2*lmax coefficients (multideal.F90:688-695)."""
    t = np.zeros((4, 4), order="F")
    for lmax in range(1, 5):
        t[:, lmax - 1] = deriv_coeffs(2 * lmax)[0]
    return t


def coeff_deriv2_table() -> np.ndarray:
    """coeff_deriv2(0:4,4): column lmax holds the order-2*lmax second-derivative
    coefficients (multideal.F90:688-695)."""
    t = np.zeros((5, 4), order="F")
    for lmax in range(1, 5):
        t[:, lmax - 1] = deriv_coeffs(2 * lmax)[1]
    return t


@dataclasses.dataclass
class Axis:
    n: int
    coord: np.ndarray      # 1-ng .. n+ng
    dxi: np.ndarray        # dcsidx (1..n)
    dxi2: np.ndarray       # dcsidxs
    dxi_2: np.ndarray      # dcsidx2


def mapped_axis(n: int, length: float, ng: int, periodic: bool, mapping, metrics_order: int = 6) -> Axis:
    """Axis with node coordinates x(i) = mapping((i-1)*dxi) for i = 1-ng .. n+ng (a
    stretched grid, as a grid read from file gives), with the local metrics computed from
    the coordinates by the central differences of compute_metrics (grid.F90:60-133) and
    set_local_grid (field.F90:300-347): dcsidx = 1/x_xi, dcsidxs = dcsidx^2,
    dcsidx2 = -x_xixi dcsidxs."""
    dxi = length / n if periodic else (length / (n - 1) if n > 1 else length)
    idx = np.arange(1 - ng, n + ng + 1)
    coord = np.asarray(mapping((idx - 1) * dxi), dtype=np.float64)
    return _axis_from_coords(n, coord, ng, metrics_order)


def _axis_from_coords(n, coord, ng, metrics_order):
    c1, c2 = deriv_coeffs(metrics_order)
    mm = metrics_order // 2
    off = ng - 1                                       # coord position of index i is i + off
    dxg = np.zeros(n)
    d2xg = np.zeros(n)
    for i in range(1, n + 1):
        s = 0.0
        for l in range(1, mm + 1):
            s = s + c1[l - 1] * (coord[i + l + off] - coord[i - l + off])
        dxg[i - 1] = s
        s2 = c2[0] * coord[i + off]
        for l in range(1, mm + 1):
            s2 = s2 + c2[l] * (coord[i + l + off] + coord[i - l + off])
        d2xg[i - 1] = s2
    dxi = 1.0 / dxg
    dxi2 = dxi * dxi
    dxi_2 = -d2xg * dxi2
    return Axis(n=n, coord=coord, dxi=dxi, dxi2=dxi2, dxi_2=dxi_2)


def uniform_axis(n: int, length: float, ng: int, periodic: bool, metrics_order: int = 6) -> Axis:
    """Coordinates x(i) = (i-1)*dx for i = 1-ng .. n+ng and the local metrics."""
    dx = length / n if periodic else (length / (n - 1) if n > 1 else length)
    idx = np.arange(1 - ng, n + ng + 1)
    coord = (idx - 1) * dx
    return _axis_from_coords(n, coord, ng, metrics_order)


def ep_ord_change(nx: int, ny: int, nz: int, bctags, ep_order: int, weno_scheme: int,
                  correct_bound_ord: int) -> np.ndarray:
    """ep_ord_change(0:nx,0:ny,0:nz,1:3) as an int32 array in Fortran order (one rank)."""
    e = np.zeros((nx + 1, ny + 1, nz + 1, 3), dtype=np.int32, order="F")
    if correct_bound_ord <= 0:
        return e
    st = max(ep_order // 2, weno_scheme)
    n = (nx, ny, nz)
    for d in range(3):
        if bctags[2 * d] != 0 or bctags[2 * d + 1] != 0:
            sl = [slice(None)] * 3
            for side in (0, 1):
                base = 0 if side == 0 else n[d]
                sl[d] = base
                e[tuple(sl) + (d,)] = -st + 1
                for q in range(1, st):
                    sl[d] = q if side == 0 else n[d] - q
                    e[tuple(sl) + (d,)] = -st + q
    return e
