"""Helpers shared by the parity tests: identical inputs on both backends and the
max-norm-per-variable relative comparison."""
import numpy as np


def rel_maxnorm(a, b, axis_var=0):
    """max_v  max|a_v - b_v| / max|b_v|  per variable (first axis)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nv = a.shape[axis_var]
    out = np.zeros(nv)
    for v in range(nv):
        d = np.abs(a[v] - b[v]).max()
        s = np.abs(b[v]).max()
        out[v] = d / s if s > 0 else d
    return out


def box_state(pkg, tables, shape, seed=11, ng=3, mach_t=0.3, temperature=1200.0, perturb_y=0.0):
    """A seeded turbulent state cropped to (nx, ny, nz) with periodic ghosts."""
    nx, ny, nz = shape
    n = max(shape)
    w, tg, info = pkg.cases.turbulence_box(tables, n, seed=seed, ng=0, mach_t=mach_t, temperature=temperature)
    w = np.ascontiguousarray(w[:, :nz, :ny, :nx])
    if perturb_y > 0:
        rng = np.random.default_rng(seed + 1)
        ns = tables.n_species
        rho = w[:ns].sum(axis=0)
        y = w[:ns] / rho
        y = y * (1.0 + perturb_y * rng.standard_normal(y.shape))
        y = np.abs(y) + 1e-6
        y = y / y.sum(axis=0)
        e_int = w[ns + 3] / rho - 0.5 * ((w[ns:ns + 3] / rho) ** 2).sum(axis=0)
        w[:ns] = rho * y
    wp = np.pad(w, [(0, 0), (ng, ng), (ng, ng), (ng, ng)], mode="wrap")
    tg = np.full(wp.shape[1:], temperature)
    return wp, tg, info


def interior(a, ng):
    return a[..., ng:-ng, ng:-ng, ng:-ng]


def faces_only_mask(shape3, ng):
    """Mask of padded points that are interior or face ghosts (not edges/corners)."""
    pz, py, px = shape3
    k, j, i = np.meshgrid(np.arange(pz), np.arange(py), np.arange(px), indexing="ij")
    out_i = (i < ng) | (i >= px - ng)
    out_j = (j < ng) | (j >= py - ng)
    out_k = (k < ng) | (k >= pz - ng)
    return (out_i.astype(int) + out_j.astype(int) + out_k.astype(int)) <= 1
