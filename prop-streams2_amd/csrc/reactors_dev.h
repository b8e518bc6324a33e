// reactors_dev.h — device code of the batched reactor integrator (streams_reactor_batch):
// right-hand sides, the species-block Rosenbrock integrator, the persistent kernel and its
// launchers.  Included by reactors.hip (the narrow instances and the entry point) and
// reactors_wide.hip (the wide instances).
#pragma once
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

// Constant-pressure adiabatic reactor (the serving workload; the reference's reactor is
// constant-volume).  State: wv = (Y_1..Y_NS, 0, 0, 0, h) with h the mixture enthalpy,
// constant in time; T from the Newton inversion of h on the cp tables, rho = p/(R T),
// dY_s/dt = omega_s W_s / rho.
template <int NS, int LO>
struct RosRhsPressure {
  const KinPacked &K;
  const ThermoLds &TL;
  const double *mw, *rgas;
  double pres, tol_iter_nr;
  __device__ inline void operator()(double *wv, double &tg, double *rhs) const {
    double yy[NS], conc[NS], acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) wv[s] = wv[s] > 0.0 ? wv[s] : 0.0;
    double ysum = wv[0];
#pragma unroll
    for (int s = 1; s < NS; ++s) ysum = ysum + wv[s];
    const double yi = 1.0 / ysum;
    double rmix = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      yy[s] = wv[s] * yi;
      rmix = rmix + rgas[s] * yy[s];
    }
    int nrange[NS];
    const double tt = temperature_newton_lds<NS, LO>(wv[NS + 3], tg, TL, TL.cp, TL.cph, tol_iter_nr, yy, nrange);
    tg = tt;
    const double rho = pres / (rmix * tt);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      conc[s] = rho * yy[s] / mw[s];
      acc[s] = 0.0;
    }
    chemistry_source_packed<NS>(K, tt, conc, acc);
    const double ri = 1.0 / rho;
#pragma unroll
    for (int s = 0; s < NS; ++s) rhs[s] = -acc[s] * ri;
#pragma unroll
    for (int m = NS; m < NS + 4; ++m) rhs[m] = 0.0;
  }
};

// rosenbrock_rhs in the reference's arithmetic: constant volume (the reference's cell)
template <int NS>
struct RosRhsPackedRef {
  const KinPacked &K;
  const Thermo &th;
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
    const double tt = temperature_from_e<NS>(ee, tg, th, tol_iter_nr, yy);
    tg = tt;
    chemistry_source_packed<NS, true>(K, tt, conc, acc, mw);
#pragma unroll
    for (int s = 0; s < NS; ++s) rhs[s] = -acc[s];
#pragma unroll
    for (int m = NS; m < NS + 4; ++m) rhs[m] = 0.0;
  }
};

// The constant-pressure reactor in the same arithmetic: T from h by the e-inversion's
// Newton iteration on the c_p tables (thh.cv_coeff holds cp_coeff), rho = p / (R T).
template <int NS>
struct RosRhsPressureRef {
  const KinPacked &K;
  const Thermo &thh;
  const double *mw, *rgas;
  double pres, tol_iter_nr;
  __device__ inline void operator()(double *wv, double &tg, double *rhs) const {
    double yy[NS], conc[NS], acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) wv[s] = wv[s] > 0.0 ? wv[s] : 0.0;
    double ysum = wv[0];
#pragma unroll
    for (int s = 1; s < NS; ++s) ysum = ysum + wv[s];
    const double yi = 1.0 / ysum;
    double rmix = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      yy[s] = wv[s] * yi;
      rmix = rmix + rgas[s] * yy[s];
    }
    const double tt = temperature_from_e<NS>(wv[NS + 3], tg, thh, tol_iter_nr, yy);
    tg = tt;
    const double rho = pres / (rmix * tt);
    const double ri = 1.0 / rho;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      conc[s] = rho * yy[s] / mw[s];
      acc[s] = 0.0;
    }
    chemistry_source_packed<NS, true>(K, tt, conc, acc, mw);
#pragma unroll
    for (int s = 0; s < NS; ++s) rhs[s] = -acc[s] * ri;
#pragma unroll
    for (int m = NS; m < NS + 4; ++m) rhs[m] = 0.0;
  }
};

// Mixed-precision variant of the same method (same ROS4 stages, step-size control and error
// norm), with three changes:
//  * The linear algebra runs on the species block.  In a reactor the momentum and energy
//    rows of the right-hand side are identically zero, so their Jacobian rows are zero, W's
//    rows there are gi * e_m, and every stage solution has exactly zero in those components;
//    the species components then solve W_ss g_s = b_s.  Solving that block alone gives the
//    same stage vectors without the 4 extra columns.
//  * The Jacobian is a one-sided difference over the N_S species columns (N_S + 1
//    right-hand sides instead of 2 (N_S + 4) + 1).
//  * W_ss is factored in FP32 and each stage solve is refined once against the FP64
//    residual b - W x, with J kept in FP64; the right-hand sides, the temperature inversion
//    and the error estimate stay FP64.  With LU = double the factorisation is FP64 and the
//    solves need no refinement (form 2).
template <int N>
__device__ inline void solve_refined(const double *jac, const float *lu, const int *indx, double gi, double *b) {
  float x32[N];
#pragma unroll
  for (int m = 0; m < N; ++m) x32[m] = (float)b[m];
  lubksb<N, float>(lu, indx, x32);
  double x[N], r[N];
#pragma unroll
  for (int m = 0; m < N; ++m) x[m] = (double)x32[m];
  // r = b - W x with W = gi*I - J (FP64)
#pragma unroll
  for (int m = 0; m < N; ++m) r[m] = b[m] - gi * x[m];
  for (int n = 0; n < N; ++n) {
    const double xn = x[n];
#pragma unroll
    for (int m = 0; m < N; ++m) r[m] = r[m] + jac[m + N * n] * xn;
  }
#pragma unroll
  for (int m = 0; m < N; ++m) x32[m] = (float)r[m];
  lubksb<N, float>(lu, indx, x32);
#pragma unroll
  for (int m = 0; m < N; ++m) b[m] = x[m] + (double)x32[m];
}

template <int N, class LU>
__device__ inline void solve_block(const double *jac, const LU *lu, const int *indx, double gi, double *b) {
  if constexpr (sizeof(LU) == sizeof(double)) lubksb<N, double>(lu, indx, b);
  else solve_refined<N>(jac, lu, indx, gi, b);
}

template <int NS, class LU, class Rhs>
__device__ void rosenbrock_cell_mp(const Rhs &rhs_fn, double *wv, double &tg, double dttry, double time_start,
                                   double time_end, int maxsteps, int maxtry, double tol, RosStats &st) {
  constexpr int NV = NS + 4;
  const double safety = 0.9, grow = 1.5, pgrow = -0.25, shrnk = 0.5, pshrnk = -1.0 / 3.0, errcon = 0.1296;
  const double gam = 1.0 / 2.0, a21 = 2.0, a31 = 48.0 / 25.0, a32 = 6.0 / 25.0;
  const double c21 = -8.0, c31 = 372.0 / 25.0, c32 = 12.0 / 5.0;
  const double c41 = -112.0 / 125.0, c42 = -54.0 / 125.0, c43 = -2.0 / 5.0;
  const double b1 = 19.0 / 9.0, b2 = 1.0 / 2.0, b3 = 25.0 / 108.0, b4 = 125.0 / 108.0;
  const double e1 = 17.0 / 54.0, e2 = 7.0 / 36.0, e3 = 0.0, e4 = 125.0 / 108.0;
  double jac[NS * NS];
  LU lu[NS * NS];
  double w_sav[NS], rhs_sav[NV], yscal[NS], rhs[NV], rp[NV];
  double g1[NS], g2[NS], g3[NS], g4[NS];
  int indx[NS];
  int finish = 0;
  double time = time_start, dt = dttry, itol = 1.0 / tol, dtnext = dt;
  if (st.resume && dt * 1.5 > time_end - time) { dt = time_end - time; finish = 1; }
  for (int iter = 1; iter <= maxsteps; ++iter) {
    const double t0 = time;
    rhs_fn(wv, tg, rhs_sav);
    for (int n = 0; n < NS; ++n) {
      const double wtmp = wv[n];
      const double h = wtmp * 1.0e-06 + 1.0e-14;
      wv[n] = wtmp + h;
      double tj = tg;
      rhs_fn(wv, tj, rp);
      wv[n] = wtmp;
      for (int m = 0; m < NS; ++m) jac[m + NS * n] = (rp[m] - rhs_sav[m]) / h;
    }
    for (int m = 0; m < NS; ++m) {
      w_sav[m] = wv[m];
      yscal[m] = fmax(1.0, fabs(w_sav[m]));
    }
    for (int jtry = 1; jtry <= maxtry; ++jtry) {
      const double gi = 1.0 / (gam * dt);
      for (int n = 0; n < NS; ++n) {
        for (int m = 0; m < NS; ++m) lu[m + NS * n] = (LU)(-jac[m + NS * n]);
        lu[n + NS * n] = (LU)(gi - jac[n + NS * n]);
      }
      ludcmp<NS, LU>(lu, indx);
      for (int m = 0; m < NS; ++m) g1[m] = rhs_sav[m];
      solve_block<NS>(jac, lu, indx, gi, g1);
      for (int m = 0; m < NS; ++m) wv[m] = w_sav[m] + a21 * g1[m];
      time = t0 + dt;
      rhs_fn(wv, tg, rhs);
      for (int m = 0; m < NS; ++m) g2[m] = rhs[m] + c21 * g1[m] / dt;
      solve_block<NS>(jac, lu, indx, gi, g2);
      for (int m = 0; m < NS; ++m) wv[m] = w_sav[m] + a31 * g1[m] + a32 * g2[m];
      time = t0 + 3.0 / 5.0 * dt;
      rhs_fn(wv, tg, rhs);
      for (int m = 0; m < NS; ++m) g3[m] = rhs[m] + (c31 * g1[m] + c32 * g2[m]) / dt;
      solve_block<NS>(jac, lu, indx, gi, g3);
      for (int m = 0; m < NS; ++m) g4[m] = rhs[m] + (c41 * g1[m] + c42 * g2[m] + c43 * g3[m]) / dt;
      solve_block<NS>(jac, lu, indx, gi, g4);
      double errmax = 0.0;
      for (int m = 0; m < NS; ++m) {
        wv[m] = w_sav[m] + b1 * g1[m] + b2 * g2[m] + b3 * g3[m] + b4 * g4[m];
        const double err = e1 * g1[m] + e2 * g2[m] + e3 * g3[m] + e4 * g4[m];
        errmax = fmax(errmax, fabs(err / yscal[m]));
      }
      time = t0 + dt;
      errmax = errmax * itol;
      if (errmax < 1.0 || finish == 1) {
        if (errmax > errcon) dtnext = safety * dt * pow(errmax, pgrow);
        else dtnext = grow * dt;
        break;
      } else {
        st.nrej += 1;
        dtnext = safety * dt * pow(errmax, pshrnk);
        dt = copysign(fmax(fabs(dtnext), shrnk * fabs(dt)), dt);
      }
    }
    st.nacc += 1;
    const double remaining = time_end - time;
    if (finish == 1 || remaining < 1.0e-12 * time_end) { st.done = 1; break; }
    if (dtnext * 1.5 > remaining) { dt = remaining; finish = 1; }
    else { dt = dtnext; finish = 0; }
  }
  st.t_end = time;
  st.dt_next = st.done ? dtnext : dt;
}

struct ReactorArgs {
  double *w, *tg;                 // w[m * nbatch + b] (SoA), T guess / result
  const double *dttry, *time_end; // per reactor
  int *nacc, *nrej;               // per reactor (may be null)
  double *t_done, *dt_next;       // per reactor, resumable mode (both null: from t = 0)
  unsigned *next;                 // work counter (starts at 0)
  int nbatch, maxsteps, maxtry, mode;   // mode: 0 plain FP64, 1 mixed, 2 FP64 species block
  double tol, tol_iter_nr;
  const double *mw, *rgas;
  const double *pressure;         // per reactor; null: constant volume
  KinPacked K;
  Thermo th;
};

template <int NS, int MODE, class Rhs>
__device__ inline void integrate_one(const Rhs &f, const ReactorArgs &A, unsigned r) {
  constexpr int NV = NS + 4;
  const long nb = A.nbatch;
  double wv[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) wv[m] = A.w[(long)m * nb + r];
  double tg = A.tg[r];
  RosStats st;
  double t0 = 0.0, dt0 = A.dttry[r];
  if (A.t_done) {
    st.resume = true;
    t0 = A.t_done[r];
    if (A.dt_next[r] > 0.0) dt0 = A.dt_next[r];
    if (t0 >= A.time_end[r]) {                        // finished earlier: nothing to do
      if (A.nacc) A.nacc[r] = 0;
      if (A.nrej) A.nrej[r] = 0;
      return;
    }
  }
  if constexpr (MODE == 1)
    rosenbrock_cell_mp<NS, float>(f, wv, tg, dt0, t0, A.time_end[r], A.maxsteps, A.maxtry, A.tol, st);
  else if constexpr (MODE == 2)
    rosenbrock_cell_mp<NS, double>(f, wv, tg, dt0, t0, A.time_end[r], A.maxsteps, A.maxtry, A.tol, st);
  else
    rosenbrock_cell_f<NS>(f, wv, tg, dt0, t0, A.time_end[r], A.maxsteps, A.maxtry, A.tol, st);
#pragma unroll
  for (int m = 0; m < NV; ++m) A.w[(long)m * nb + r] = wv[m];
  A.tg[r] = tg;
  if (A.nacc) A.nacc[r] = st.nacc;
  if (A.nrej) A.nrej[r] = st.nrej;
  if (A.t_done) {
    A.t_done[r] = st.done ? A.time_end[r] : st.t_end;
    A.dt_next[r] = st.dt_next;
  }
}

// Waves per SIMD the compiler budgets registers for.  Every form spills (the per-lane
// Jacobian, its factors and the four stage vectors do not fit).  A batch larger than the
// GPU holds at one wave per SIMD (CUs x 256 lanes) runs the WIDE instance: more resident
// waves then hide the spill and table latencies better than fewer, larger register files.
// Measured on the 2^19-reactor H2-air workload (mixed / FP64 species block, reactor-solves/s
// at 1, 2, 3, 4 waves): 155k/171k, 237k/257k, 253k/283k, 221k/234k; on the 14-species
// mixing-layer cells best at 2 waves in both forms; the plain FP64 form at 2.  A smaller
// batch (the serving loop's) cannot fill the extra waves, and each lane is faster at one
// wave: the reactor server ran 15 % slower with the wide instances.  The two instances are
// compiled in separate translation units (reactors.hip, reactors_wide.hip): device code
// that both called would get one register budget for both.
template <int NS, int MODE, bool WIDE>
constexpr int reactor_waves() { return !WIDE ? 1 : MODE == 0 ? 2 : (NS <= 9 ? 3 : 2); }

template <int NS, int LO, int MODE, bool CP, bool WIDE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(reactor_waves<NS, MODE, WIDE>())))
reactor_kernel(ReactorArgs A) {
  constexpr int WHICH = CP ? (TH_CP | TH_CPH) : (TH_CV | TH_CVH);
  // the plain FP64 form reads the NASA tables from global memory in the reference's forms
  __shared__ double tlds[MODE == 0 ? 1 : thermo_lds_doubles<NS>(WHICH)];
  ThermoLds TL{};
  if constexpr (MODE != 0) {
    TL = thermo_stage<NS, LO>(A.th, tlds, WHICH);
    __syncthreads();
  }
  Thermo thh = A.th;
  thh.cv_coeff = A.th.cp_coeff;
  const long nb = A.nbatch;
  for (;;) {
    const unsigned r = atomicAdd(A.next, 1u);
    if ((long)r >= nb) break;
    if constexpr (MODE == 0) {
      if (CP) {
        const RosRhsPressureRef<NS> f{A.K, thh, A.mw, A.rgas, A.pressure[r], A.tol_iter_nr};
        integrate_one<NS, MODE>(f, A, r);
      } else {
        const RosRhsPackedRef<NS> f{A.K, A.th, A.mw, A.tol_iter_nr};
        integrate_one<NS, MODE>(f, A, r);
      }
    } else if (CP) {
      const RosRhsPressure<NS, LO> f{A.K, TL, A.mw, A.rgas, A.pressure[r], A.tol_iter_nr};
      integrate_one<NS, MODE>(f, A, r);
    } else {
      const RosRhsPacked<NS, LO> f{A.K, TL, A.mw, A.tol_iter_nr};
      integrate_one<NS, MODE>(f, A, r);
    }
  }
}

// Persistent grid: wg_per_cu workgroups per CU (default 4, at least the resident capacity
// of these register-heavy kernels), never more than the batch needs.  Any size is correct:
// lanes pull reactors from the shared counter until it runs out, so workgroups queued
// behind the resident ones start after it has run out and exit at once.  The occupancy
// API is not used: its query has returned error 999 after many other kernels had been
// launched in the process.
inline int device_cus() {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      ncu <= 0) {
    (void)hipGetLastError();
    ncu = 256;
  }
  return ncu;
}

inline int reactor_grid(int ncu, int wg_per_cu, int min_per_cu, long nbatch) {
  long per = wg_per_cu > 0 ? wg_per_cu : 4;
  per = per < min_per_cu ? min_per_cu : per;
  const long need = (nbatch + 255) / 256;
  const long g = ncu * per;
  return (int)(g < need ? g : need);
}

template <int NS, int LO, int MODE, bool CP, bool WIDE>
int reactor_launch_t(const ReactorArgs &A, int ncu, int wg_per_cu, hipStream_t st) {
  // a workgroup is one wave per SIMD: an instance needs as many per CU as its waves
  const int g = reactor_grid(ncu, wg_per_cu, reactor_waves<NS, MODE, WIDE>(), A.nbatch);
  if (g <= 0) return STREAMS_E_DIMS;
  reactor_kernel<NS, LO, MODE, CP, WIDE><<<g, 256, 0, st>>>(A);
  return hip_status();
}

template <int NS, int LO, bool WIDE>
int reactor_launch(const ReactorArgs &A, int ncu, int wg_per_cu, hipStream_t st) {
  const bool cp = A.pressure != nullptr;
  switch (A.mode) {
    case 0: return cp ? reactor_launch_t<NS, LO, 0, true, WIDE>(A, ncu, wg_per_cu, st)
                      : reactor_launch_t<NS, LO, 0, false, WIDE>(A, ncu, wg_per_cu, st);
    case 1: return cp ? reactor_launch_t<NS, LO, 1, true, WIDE>(A, ncu, wg_per_cu, st)
                      : reactor_launch_t<NS, LO, 1, false, WIDE>(A, ncu, wg_per_cu, st);
    case 2: return cp ? reactor_launch_t<NS, LO, 2, true, WIDE>(A, ncu, wg_per_cu, st)
                      : reactor_launch_t<NS, LO, 2, false, WIDE>(A, ncu, wg_per_cu, st);
    default: return STREAMS_E_OPTION;
  }
}

// Launch of one instance for any N_S and NASA range (instantiated in one translation unit
// per instance).
template <bool WIDE>
int reactor_dispatch(int n_s, int indx_cp_l, int indx_cp_r, const ReactorArgs &A, int ncu, int wg_per_cu,
                     hipStream_t st) {
  STREAMS_DISPATCH_NS(n_s, STREAMS_DISPATCH_NASA(indx_cp_l, indx_cp_r,
                                                 return (reactor_launch<NS, LO, WIDE>(A, ncu, wg_per_cu, st))));
  return 0;
}

// the wide instances, compiled in reactors_wide.hip
int reactor_dispatch_wide(int n_s, int indx_cp_l, int indx_cp_r, const ReactorArgs &A, int ncu, int wg_per_cu,
                          hipStream_t st);

}  // namespace streams
