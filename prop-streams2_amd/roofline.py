"""Algorithmic HBM traffic of each entry point, per launch.

Each figure counts every array element the routine must read or write exactly once
(8-byte doubles), the stencil halos of the sweeps excluded: this is the traffic a kernel
that keeps its reuse on chip moves, so ``bytes / duration`` against the 8 TB/s HBM peak is
its roofline fraction.  Counts per point, N_S species, NV = N_S + 4:

  eval_aux          padded points  read NV + 1 (w, T guess), write 2 N_S + 11
  eval_aux2         interior       read 3 (u, v, w), write 2 (Ducros, div/3); without the
                                   sensor ("eval_aux2_div") write 1
  euler_{x,y,z}     interior       read N_S + 6 (Y, rho, u, v, w, h, p), fl read+write 2 NV
  visflx_{x,y,z}    interior       read 2 N_S + 6 (u, v, w, T, mu, k, Y, rho D), fl 2 NV
  visflx_reduced    interior       read 5 (u, v, w, mu, div/3), fl 2 x 4
  compute_chemistry interior       read NV + 1, write N_S + 1 (Y, T), fl 2 N_S
  init/update_flux  interior       3 NV;   update_field 3 NV
  rhs_sweep         interior       read N_S + 6 (convective) + N_S + 3 (T, mu, k, rho D),
                                   fl read+write 2 NV (+ NV for fln when init_flux is folded;
                                   x with fln already set, "rhs_sweep_1_set": fl write NV)
  rhs_finish        interior       read 5 + N_S + 2 (u, v, w, mu, div/3, rho, T, Y),
                                   fl, fln, w read+write 6 NV
  rhs_finish_next   interior       the same without the fl store: 5 NV
"""
from __future__ import annotations

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP64_VECTOR_TFLOPS = 78.6


def per_point_doubles(name: str, ns: int) -> int:
    nv = ns + 4
    table = {
        "eval_aux": nv + 1 + 2 * ns + 11,
        "eval_aux2": 5, "eval_aux2_div": 4,
        "euler_x": ns + 6 + 2 * nv, "euler_y": ns + 6 + 2 * nv, "euler_z": ns + 6 + 2 * nv,
        "visflx_x": 2 * ns + 6 + 2 * nv, "visflx_y": 2 * ns + 6 + 2 * nv, "visflx_z": 2 * ns + 6 + 2 * nv,
        "visflx_reduced": 5 + 8,
        "compute_chemistry": nv + 1 + ns + 1 + 2 * ns,
        "init_flux": 3 * nv, "update_flux": 3 * nv, "update_field": 3 * nv,
        "rhs_sweep_1": (ns + 6) + (ns + 3) + 3 * nv,
        "rhs_sweep_2": (ns + 6) + (ns + 3) + 2 * nv,
        "rhs_sweep_3": (ns + 6) + (ns + 3) + 2 * nv,
        "rhs_sweep_1_set": (ns + 6) + (ns + 3) + nv,
        "rhs_finish": 7 + ns + 6 * nv,
        "rhs_finish_next": 7 + ns + 5 * nv,
    }
    return table.get(name, 0)


def launch_bytes(name: str, ns: int, nx: int, ny: int, nz: int, ng: int) -> int:
    """Algorithmic bytes of one call of entry point `name` on an nx*ny*nz box."""
    if name == "eval_aux":
        pts = (nx + 2 * ng) * (ny + 2 * ng) * (nz + 2 * ng)
    else:
        pts = nx * ny * nz
    return 8 * per_point_doubles(name, ns) * pts


def substep_bytes_survey(ns: int, n_cells: int, n: int, ng: int = 3) -> int:
    """Whole-substep algorithmic bytes: (30 N_S + 132) doubles per cell plus the halo
    exchange 12 ng NV / n (the accounting of the measurement plan)."""
    nv = ns + 4
    per_cell = (30 * ns + 132) + 12.0 * ng * nv / n
    return int(round(8 * per_cell * n_cells))
