// reactors.hip — batched stiff chemistry: millions of independent 0-D reactors (constant
// volume, or constant pressure) integrated in one persistent launch.
//
// Each reactor is one cell of the reference's rosenbrock_kernel (kernels_gpu.F90:8928-9193):
// adaptive 4-stage Rosenbrock (Numerical-Recipes ROS4 constants), error scaled by
// max(1,|w|).  Three forms of it (the `mixed` argument):
//  0  the plain FP64 integrator, the yardstick of the other two: central-difference
//     Jacobian over all N_S + 4 variables, FP64 Crout LU with implicit pivoting, and
//     rosenbrock_rhs (:9249-9339) in the reference's own arithmetic forms (global NASA
//     tables, reference-form kinetics from the packed records);
//  1  mixed precision on the species block: one-sided species Jacobian, FP32 LU refined
//     against the FP64 residual, fast-form kinetics and the NASA tables staged in LDS;
//  2  form 1 with the LU in FP64 (no refinement): the like-for-like FP64 counterpart of 1,
//     so that the difference between 1 and 2 is the precision of the factorisation alone.
// Reactors have uneven stiffness, so a fixed
// reactor-to-thread assignment would leave lanes idle behind the stiffest reactor of their
// wave: instead every lane pulls its next reactor from a device-wide counter as soon as it
// is done, and the grid is sized to the resident capacity of the GPU (persistent).
#include "reactors_dev.h"

namespace streams {
__global__ void zero_counter_kernel(unsigned *c) { *c = 0u; }
}  // namespace streams

using namespace streams;

// Batched integration of nbatch independent reactors (rosenbrock_kernel per reactor with
// mixed = 0, the mixed-precision species-block variant with mixed = 1, its FP64 counterpart
// with mixed = 2) from t = 0 to time_end[b]: constant volume when
// pressure is null, else constant pressure pressure[b] with the state (Y, 0, 0, 0, h).  With
// t_done / dt_next (both or neither) the call is resumable: each reactor starts at
// t_done[b] with the step dt_next[b] (dttry[b] when not positive), takes at most maxsteps
// steps, and returns the time reached and its next step there.  work points to one unsigned
// of device memory used as the shared work counter (zeroed here, on the stream);
// wg_per_cu <= 0 gives 4 workgroups per CU.
extern "C" int streams_reactor_batch(int n_s, int nbatch, double *w, double *tguess, const double *dttry,
                                     const double *time_end, const double *pressure, const double *mw,
                                     const double *rgas, const double *cp_coeff, int nreactions,
                                     const double *kin_pack, const double *kc_tab, int num_t_tab, double t_min_tab,
                                     double dt_tab, double r_univ, double tol_iter_nr, const double *cv_coeff,
                                     int nsetcv, const double *trange, int indx_cp_l, int indx_cp_r, int maxsteps,
                                     int maxtry, double tol, int *nacc, int *nrej, double *t_done, double *dt_next,
                                     unsigned *work, int wg_per_cu, int mixed, void *stream) {
  if (nbatch < 0 || nreactions < 0 || maxsteps < 1 || maxtry < 1) return STREAMS_E_DIMS;
  if (mixed < 0 || mixed > 2) return STREAMS_E_OPTION;
  if (n_s > KTB) return STREAMS_E_NSPECIES;
  if (nsetcv < 1 || nsetcv > TMAXSET) return STREAMS_E_OPTION;
  if (nbatch == 0) return 0;
  if (work == nullptr) return STREAMS_E_DIMS;
  hipStream_t st = as_stream(stream);
  if (pressure != nullptr && (rgas == nullptr || cp_coeff == nullptr)) return STREAMS_E_DIMS;
  if ((t_done == nullptr) != (dt_next == nullptr)) return STREAMS_E_DIMS;
  ReactorArgs A{w, tguess, dttry, time_end, nacc, nrej, t_done, dt_next, work, nbatch, maxsteps, maxtry, mixed, tol,
                tol_iter_nr, mw,
                rgas, pressure, KinPacked{kin_pack, kc_tab, nreactions, num_t_tab, t_min_tab, dt_tab, r_univ},
                Thermo{cv_coeff, cp_coeff, trange, nsetcv, indx_cp_l, indx_cp_r}};
  zero_counter_kernel<<<1, 1, 0, st>>>(work);
  const int ncu = device_cus();
  if ((long)nbatch > (long)ncu * 256)
    return reactor_dispatch_wide(n_s, indx_cp_l, indx_cp_r, A, ncu, wg_per_cu, st);
  return reactor_dispatch<false>(n_s, indx_cp_l, indx_cp_r, A, ncu, wg_per_cu, st);
}
