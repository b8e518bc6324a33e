// chem.hip — finite-rate chemistry source and the per-cell Rosenbrock integrator (FP64).
//
// compute_chemistry: one thread per interior cell, reaction loop over the tables in
// their Fortran layout (arr_a(nr,2), tb_eff(nr,N_S), kc_tab(num_t_tab+1,nr), ...).
// rosenbrock: one thread per cell restating rosenbrock_kernel with its central-difference
// Jacobian and Crout LU; it is the FP64 yardstick for the batched integrators of reactors.hip.
#include "chem_common.h"

namespace streams {

template <int NS>
__global__ void __launch_bounds__(256) chemistry_kernel(Box b, double *__restrict__ w, double *__restrict__ w_aux,
                                                        double *__restrict__ fl, Kin K, Thermo th,
                                                        double tol_iter_nr) {
  using X = Idx<NS>;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.fstride()) return;
  int i = 1 + (int)(t % b.nx);
  long r = t / b.nx;
  int j = 1 + (int)(r % b.ny);
  int k = 1 + (int)(r / b.ny);
  const long pt = b.p(i, j, k), ps = b.pstride();
  double yy[NS], conc[NS];
  double rho = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double v = w[pt + ps * s];
    if (!(v > 0.0)) { v = 0.0; w[pt + ps * s] = 0.0; }
    yy[s] = v;
    rho = (s == 0) ? v : rho + v;
  }
  const double ri = 1.0 / rho;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    yy[s] = yy[s] * ri;
    w_aux[pt + ps * s] = yy[s];
    conc[s] = rho * yy[s] / K.mw[s];
  }
  const double uu = w[pt + ps * (X::I_U - 1)] * ri, vv = w[pt + ps * (X::I_V - 1)] * ri;
  const double ww = w[pt + ps * (X::I_W - 1)] * ri, rhoe = w[pt + ps * (X::I_E - 1)];
  const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
  const double ee = rhoe * ri - qq;
  const double tt = temperature_from_e<NS>(ee, w_aux[pt + ps * (X::J_T - 1)], th, tol_iter_nr, yy);
  w_aux[pt + ps * (X::J_T - 1)] = tt;
  const double Rtti = 1.0 / (K.r_univ * tt);
  int itt = (int)((tt - K.t_min_tab) / K.dt_tab) + 1;
  itt = itt < 1 ? 1 : itt;
  itt = itt > K.num_t_tab ? K.num_t_tab : itt;
  const double ttleft = (itt - 1) * K.dt_tab + K.t_min_tab;
  const double dtt = (tt - ttleft) / K.dt_tab;
  double acc[NS];
  const long fs_ = b.fstride();
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = fl[t + fs_ * s];
  for (int lr = 0; lr < K.nr; ++lr) {
    const double q = progress_rate<NS>(K, lr, tt, Rtti, itt, dtt, conc);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double wdt = (K.p_co[lr + K.nr * s] - K.r_co[lr + K.nr * s]) * q * K.mw[s];
      acc[s] = acc[s] - wdt;
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) fl[t + fs_ * s] = acc[s];
}

// ------------------------------------------------------------------ Rosenbrock
// 2 waves per SIMD (the compiler's own choice is 1, with 2.8-3.7 KB of per-lane spills):
// the extra resident wave hides the spill latency (one integration of a 128^3 box at 1500 K,
// 87.3 -> 67.2 ms; 3 waves: 92.9 ms; bit-identical)
template <int NS>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) rosenbrock_kernel(Box b, double *__restrict__ w, double *__restrict__ w_aux,
                                                         Kin K, Thermo th, double tol_iter_nr, double dttry,
                                                         double time_start, double time_end, int maxsteps,
                                                         int maxtry, double tol, Pasr pasr) {
  using X = Idx<NS>;
  constexpr int NV = NS + 4;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.fstride()) return;
  int i = 1 + (int)(t % b.nx);
  long r = t / b.nx;
  int j = 1 + (int)(r % b.ny);
  int k = 1 + (int)(r / b.ny);
  const long pt = b.p(i, j, k), ps = b.pstride();
  double wv[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) wv[m] = w[pt + ps * m];
  double tg = w_aux[pt + ps * (X::J_T - 1)];
  const double tau_k = pasr.on ? pasr.tau_k<NS>(w_aux, pt, ps, i, j, k) : 0.0;
  RosStats st;
  rosenbrock_cell<NS>(K, th, tol_iter_nr, wv, tg, dttry, time_start, time_end, maxsteps, maxtry, tol, st, tau_k);
#pragma unroll
  for (int m = 0; m < NV; ++m) w[pt + ps * m] = wv[m];
  w_aux[pt + ps * (X::J_T - 1)] = tg;
}

}  // namespace streams

using namespace streams;

extern "C" int streams_compute_chemistry(int n_s, int nx, int ny, int nz, int nv, int nv_aux, int ng,
                                         int nreactions, double *w, double *w_aux, double *fl, const double *mw,
                                         const double *arr_a, const double *arr_b, const double *arr_ea,
                                         const double *tb_eff, const double *falloff_coeffs, const int *reac_ty,
                                         const int *isrev, const double *r_coeffs, const double *p_coeffs,
                                         const double *kc_tab, int num_t_tab, double t_min_tab, double dt_tab,
                                         double r_univ, int indx_cp_l, int indx_cp_r, const double *cv_coeff,
                                         int nsetcv, const double *trange, double tol_iter_nr, void *stream) {
  if (nv != n_s + 4 || nv_aux < 2 * n_s + 13 || nreactions < 0) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  long n = b.fstride();
  if (n == 0) return 0;
  Kin K{mw, arr_a, arr_b, arr_ea, tb_eff, falloff_coeffs, r_coeffs, p_coeffs, kc_tab, reac_ty, isrev,
        nreactions, num_t_tab, t_min_tab, dt_tab, r_univ};
  Thermo th{cv_coeff, nullptr, trange, nsetcv, indx_cp_l, indx_cp_r};
  STREAMS_DISPATCH_NS(n_s, (chemistry_kernel<NS><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
                                b, w, w_aux, fl, K, th, tol_iter_nr)));
  return hip_status();
}

extern "C" int streams_rosenbrock(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng, double *w,
                                  double *w_aux, const double *fl, const double *mw, int nreactions,
                                  const double *arr_a, const double *arr_b, const double *arr_ea,
                                  const double *tb_eff, const double *falloff_coeffs, const int *reac_ty,
                                  const int *isrev, const double *r_coeffs, const double *p_coeffs,
                                  const double *kc_tab, int num_t_tab, double t_min_tab, double dt_tab,
                                  double r_univ, double tol_iter_nr, const double *cv_coeff, int nsetcv,
                                  const double *trange, int indx_cp_l, int indx_cp_r, double dttry,
                                  double time_start, double time_end, int maxsteps, int maxtry, double tol,
                                  int simpler_splitting, int enable_pasr, double les_c_yoshi, double les_c_mix,
                                  double les_c_eps, const double *dcsidx, const double *detady,
                                  const double *dzitdz, void *stream) {
  (void)fl; (void)simpler_splitting;
  if (nv != n_s + 4 || nv_aux < 2 * n_s + 13 || maxsteps < 1 || maxtry < 1) return STREAMS_E_DIMS;
  if (enable_pasr > 0 && (nv_aux < 2 * n_s + 14 || !dcsidx || !detady || !dzitdz)) return STREAMS_E_DIMS;
  const Pasr pasr{enable_pasr > 0, les_c_yoshi, les_c_mix, les_c_eps, dcsidx, detady, dzitdz};
  Box b{nx, ny, nz, ng};
  long n = b.fstride();
  if (n == 0) return 0;
  Kin K{mw, arr_a, arr_b, arr_ea, tb_eff, falloff_coeffs, r_coeffs, p_coeffs, kc_tab, reac_ty, isrev,
        nreactions, num_t_tab, t_min_tab, dt_tab, r_univ};
  Thermo th{cv_coeff, nullptr, trange, nsetcv, indx_cp_l, indx_cp_r};
  STREAMS_DISPATCH_NS(n_s, (rosenbrock_kernel<NS><<<blocks_for(n, 128), 128, 0, as_stream(stream)>>>(
                                b, w, w_aux, K, th, tol_iter_nr, dttry, time_start, time_end, maxsteps, maxtry,
                                tol, pasr)));
  return hip_status();
}
