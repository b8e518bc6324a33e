// reactors.hip — batched stiff chemistry: millions of independent 0-D constant-volume
// reactors integrated in one persistent launch.
//
// Each reactor is one cell of the reference's rosenbrock_kernel (kernels_gpu.F90:8928-9193):
// adaptive 4-stage Rosenbrock (Numerical-Recipes ROS4 constants), central-difference
// Jacobian, Crout LU with implicit pivoting, error scaled by max(1,|w|).  The right-hand side
// is rosenbrock_rhs (:9249-9339) evaluated with the packed kinetics records (scalar loads)
// and the NASA tables staged in LDS.  Reactors have uneven stiffness, so a fixed
// reactor-to-thread assignment would leave lanes idle behind the stiffest reactor of their
// wave: instead every lane pulls its next reactor from a device-wide counter as soon as it
// is done, and the grid is sized to the resident capacity of the GPU (persistent).
#include "chem_common.h"
#include "kinpack.h"
#include "thermo_dev.h"

namespace streams {

// rosenbrock_rhs on the packed tables
template <int NS, int LO>
struct RosRhsPacked {
  const KinPacked &K;
  const ThermoLds &TL;
  const double *mw;
  double tol_iter_nr;
  __device__ inline void operator()(double *wv, double &tg, double *rhs) const {
    double yy[NS], conc[NS], acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) wv[s] = wv[s] > 0.0 ? wv[s] : 0.0;
    double rho = wv[0];
#pragma unroll
    for (int s = 1; s < NS; ++s) rho = rho + wv[s];
    const double ri = 1.0 / rho;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      yy[s] = wv[s] * ri;
      conc[s] = rho * yy[s] / mw[s];
      acc[s] = 0.0;
    }
    const double uu = wv[NS] * ri, vv = wv[NS + 1] * ri, ww = wv[NS + 2] * ri;
    const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
    const double ee = wv[NS + 3] * ri - qq;
    int nrange[NS];
    const double tt = temperature_from_e_lds<NS, LO>(ee, tg, TL, tol_iter_nr, yy, nrange);
    tg = tt;
    chemistry_source_packed<NS>(K, tt, conc, acc);
#pragma unroll
    for (int s = 0; s < NS; ++s) rhs[s] = -acc[s];
#pragma unroll
    for (int m = NS; m < NS + 4; ++m) rhs[m] = 0.0;
  }
};

struct ReactorArgs {
  double *w, *tg;                 // w[m * nbatch + b] (SoA), T guess / result
  const double *dttry, *time_end; // per reactor
  int *nacc, *nrej;               // per reactor (may be null)
  unsigned *next;                 // work counter (starts at 0)
  int nbatch, maxsteps, maxtry;
  double tol, tol_iter_nr;
  const double *mw;
  KinPacked K;
  Thermo th;
};

template <int NS, int LO>
__global__ void __launch_bounds__(256) reactor_kernel(ReactorArgs A) {
  constexpr int NV = NS + 4;
  __shared__ double tlds[thermo_lds_doubles<NS>(TH_CV | TH_CVH)];
  const ThermoLds TL = thermo_stage<NS, LO>(A.th, tlds, TH_CV | TH_CVH);
  __syncthreads();
  const RosRhsPacked<NS, LO> f{A.K, TL, A.mw, A.tol_iter_nr};
  const long nb = A.nbatch;
  for (;;) {
    const unsigned r = atomicAdd(A.next, 1u);
    if ((long)r >= nb) break;
    double wv[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) wv[m] = A.w[(long)m * nb + r];
    double tg = A.tg[r];
    RosStats st;
    rosenbrock_cell_f<NS>(f, wv, tg, A.dttry[r], 0.0, A.time_end[r], A.maxsteps, A.maxtry, A.tol, st);
#pragma unroll
    for (int m = 0; m < NV; ++m) A.w[(long)m * nb + r] = wv[m];
    A.tg[r] = tg;
    if (A.nacc) A.nacc[r] = st.nacc;
    if (A.nrej) A.nrej[r] = st.nrej;
  }
}

__global__ void zero_counter_kernel(unsigned *c) { *c = 0u; }

// resident workgroups of reactor_kernel<NS, LO> on this device (persistent grid size)
template <int NS, int LO>
int reactor_grid(int wg_per_cu) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reactor_kernel<NS, LO>, 256, 0) != hipSuccess) occ = 1;
  if (wg_per_cu > 0 && wg_per_cu < occ) occ = wg_per_cu;
  return ncu * (occ > 0 ? occ : 1);
}

template <int NS, int LO>
int reactor_launch(const ReactorArgs &A, int wg_per_cu, hipStream_t st) {
  const int g = reactor_grid<NS, LO>(wg_per_cu);
  if (g <= 0) return STREAMS_E_DIMS;
  reactor_kernel<NS, LO><<<g, 256, 0, st>>>(A);
  return hip_status();
}

}  // namespace streams

using namespace streams;

// Batched integration of nbatch independent constant-volume reactors (rosenbrock_kernel per
// reactor) from t = 0 to time_end[b].  work points to one unsigned of device memory used as
// the shared work counter (zeroed here, on the stream); wg_per_cu <= 0 lets the occupancy
// calculator size the persistent grid.
extern "C" int streams_reactor_batch(int n_s, int nbatch, double *w, double *tguess, const double *dttry,
                                     const double *time_end, const double *mw, int nreactions,
                                     const double *kin_pack, const double *kc_tab, int num_t_tab, double t_min_tab,
                                     double dt_tab, double r_univ, double tol_iter_nr, const double *cv_coeff,
                                     int nsetcv, const double *trange, int indx_cp_l, int indx_cp_r, int maxsteps,
                                     int maxtry, double tol, int *nacc, int *nrej, unsigned *work, int wg_per_cu,
                                     void *stream) {
  if (nbatch < 0 || nreactions < 0 || maxsteps < 1 || maxtry < 1) return STREAMS_E_DIMS;
  if (n_s > KTB) return STREAMS_E_NSPECIES;
  if (nsetcv < 1 || nsetcv > TMAXSET) return STREAMS_E_OPTION;
  if (nbatch == 0) return 0;
  if (work == nullptr) return STREAMS_E_DIMS;
  hipStream_t st = as_stream(stream);
  ReactorArgs A{w, tguess, dttry, time_end, nacc, nrej, work, nbatch, maxsteps, maxtry, tol, tol_iter_nr, mw,
                KinPacked{kin_pack, kc_tab, nreactions, num_t_tab, t_min_tab, dt_tab, r_univ},
                Thermo{cv_coeff, nullptr, trange, nsetcv, indx_cp_l, indx_cp_r}};
  zero_counter_kernel<<<1, 1, 0, st>>>(work);
  STREAMS_DISPATCH_NS(n_s, STREAMS_DISPATCH_NASA(indx_cp_l, indx_cp_r, return (reactor_launch<NS, LO>(A, wg_per_cu, st))));
  return 0;
}
