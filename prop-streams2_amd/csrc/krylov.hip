// krylov.hip — the Rosenbrock–Krylov integrator of the stiff chemistry (FP64).
//
// One thread per interior cell restates rosenbrock_krylov_kernel
// (kernels_gpu.F90:10360-10611): a 4-stage ROK method whose linear algebra lives in a
// Krylov space of at most KM = 4 vectors (M_max, index_define.h:27), rebuilt by Arnoldi
// (arnoldi, :10664-10792) on one-sided finite-difference Jacobian-vector products
// (rosenbrock_krylov_jacobian, :10613-10662) for every attempted step.  Only cells with a
// Bilger mixture fraction strictly inside (5e-4, 1 - 5e-4) are integrated; the right-hand
// side is rosenbrock_krylov_rhs (:10794-10976), which is rosenbrock_rhs's arithmetic on a
// copy of the state with the temperature guess passed by value.
#include "chem_common.h"

namespace streams {

constexpr int KM = 4;       // Krylov space size
constexpr int KSTAGE = 4;   // stages

// rosenbrock_krylov_rhs: the state is not modified and the temperature guess is not updated
template <int NS>
__device__ inline void kry_rhs(const Kin &K, const Thermo &th, double tol_iter_nr, const double *w, double t_iter,
                               double *rhs, double tau_k) {
  constexpr int NV = NS + 4;
  double tmp[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) tmp[m] = w[m];
  double tg = t_iter;
  ros_rhs<NS>(K, th, tol_iter_nr, tmp, tg, rhs, tau_k);
}

template <int N>
__device__ inline double norm2v(const double *v) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) s = s + v[i] * v[i];
  return sqrt(s);
}

// inv_matrix_krylov (:11087-11119): inverse of the leading d x d block of a (KM x KM,
// column-major) in place, through the Crout LU of the a8 integrator on that block
template <int D>
__device__ inline void kry_inv_block(double *a) {
  double blk[D * D], y[D * D];
  int indx[D];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int r = 0; r < D; ++r) blk[r + D * c] = a[r + KM * c];
  ludcmp<D>(blk, indx);
#pragma unroll
  for (int c = 0; c < D; ++c) {
#pragma unroll
    for (int r = 0; r < D; ++r) y[r + D * c] = r == c ? 1.0 : 0.0;
    lubksb<D>(blk, indx, y + D * c);
  }
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int r = 0; r < D; ++r) a[r + KM * c] = y[r + D * c];
}

__device__ inline void kry_inv(double *a, int d) {
  if (d == 1) a[0] = 1.0 / a[0];
  else if (d == 2) kry_inv_block<2>(a);
  else if (d == 3) kry_inv_block<3>(a);
  else kry_inv_block<4>(a);
}

// arnoldi: basis Vb (NV x KM) and inv = (I - dt gam H)^-1 of the first Krylov size kk with
// |dt gam H(kk+1,kk)| |lambda(kk)| < 1e-6; Vb and inv stay zero when no size passes.
template <int NS>
__device__ void kry_arnoldi(const Kin &K, const Thermo &th, double tol_iter_nr, const double *w_sav, double dt,
                            double t_iter, const double *fls, double *Vb, double *inv, double *rhs_sav,
                            double tau_k) {
  constexpr int NV = NS + 4;
  const double gam = 0.572816062482135, TOL = 1.0e-6, eps = 1.0e-08;
  double HH[(KM + 1) * KM], V[NV * (KM + 1)], HHinv[KM * KM], rpr[KM], jac[NV], wl[NV];
  for (int q = 0; q < (KM + 1) * KM; ++q) HH[q] = 0.0;
  for (int q = 0; q < NV * (KM + 1); ++q) V[q] = 0.0;
  for (int q = 0; q < KM * KM; ++q) { HHinv[q] = 0.0; inv[q] = 0.0; }
  for (int q = 0; q < NV * KM; ++q) Vb[q] = 0.0;
  kry_rhs<NS>(K, th, tol_iter_nr, w_sav, t_iter, rhs_sav, tau_k);
  if (fls)
    for (int m = 0; m < NV; ++m) rhs_sav[m] = rhs_sav[m] - fls[m];
  const double nrm = norm2v<NV>(rhs_sav);
  for (int m = 0; m < NV; ++m) V[m] = rhs_sav[m] / nrm;
  for (int kk = 1; kk <= KM; ++kk) {
    for (int m = 0; m < NV; ++m) wl[m] = w_sav[m] + eps * V[m + NV * (kk - 1)];
    kry_rhs<NS>(K, th, tol_iter_nr, wl, t_iter, jac, tau_k);
    if (fls)
      for (int m = 0; m < NV; ++m) jac[m] = jac[m] - fls[m];
    for (int m = 0; m < NV; ++m) jac[m] = (jac[m] - rhs_sav[m]) / eps;
    for (int jj = 1; jj <= kk; ++jj) {
      double d = 0.0;
      for (int m = 0; m < NV; ++m) d = d + V[m + NV * (jj - 1)] * jac[m];
      HH[(jj - 1) + (KM + 1) * (kk - 1)] = d;
      for (int m = 0; m < NV; ++m) jac[m] = jac[m] - d * V[m + NV * (jj - 1)];
    }
    const double hn = norm2v<NV>(jac);
    HH[kk + (KM + 1) * (kk - 1)] = hn;
    for (int m = 0; m < NV; ++m) V[m + NV * kk] = jac[m] / hn;
    // V^T rhs_sav over the first kk basis vectors (the rows beyond are zero)
    for (int ll = 1; ll <= KM; ++ll) {
      double r = 0.0;
      if (ll <= kk)
        for (int m = 0; m < NV; ++m) r = r + V[m + NV * (ll - 1)] * rhs_sav[m];
      rpr[ll - 1] = r;
    }
    for (int n = 1; n <= kk; ++n)
      for (int nn = 1; nn <= kk; ++nn) {
        double v = -(dt * gam * HH[(n - 1) + (KM + 1) * (nn - 1)]);
        if (n == nn) v = v + 1.0;
        HHinv[(n - 1) + KM * (nn - 1)] = v;
      }
    for (int n = kk + 1; n <= KM; ++n) HHinv[(n - 1) + KM * (n - 1)] = 1.0;
    kry_inv(HHinv, kk);
    double lam = 0.0;
    for (int jj = 1; jj <= KM; ++jj) lam = lam + HHinv[(kk - 1) + KM * (jj - 1)] * rpr[jj - 1];
    if (fabs(dt * gam * hn) * fabs(lam) < TOL) {
      for (int c = 0; c < kk; ++c)
        for (int m = 0; m < NV; ++m) Vb[m + NV * c] = V[m + NV * c];
      for (int q = 0; q < KM * KM; ++q) inv[q] = HHinv[q];
      break;
    }
  }
}

// One cell from time_start to time_end; returns the number of attempted steps.
template <int NS>
__device__ int krylov_cell(const Kin &K, const Thermo &th, double tol_iter_nr, double *vec, double t_iter,
                           double dttry, double time_start, double time_end, int maxsteps, const double *fls,
                           double tau_k = 0.0) {
  constexpr int NV = NS + 4;
  using X = Idx<NS>;
  const double gam = 0.572816062482135;
  const double a21 = 0.432364435748567, a31 = -0.514211316876170, a32 = 1.382271144617360;
  const double a41 = -0.514211316876170, a42 = 1.382271144617360, a43 = 0.0;
  const double g21 = -0.602765307997356, g31 = -1.389195789724843, g32 = 1.072950969011413;
  const double g41 = 0.992356412977094, g42 = -1.390032613873701, g43 = -0.440875890223325;
  const double bs[KSTAGE] = {0.194335256262729, 0.483167813989227, 0.0, 0.322496929748044};
  const double bt[KSTAGE] = {-0.217819895945721, 1.03130847478467, 0.186511421161047, 0.0};
  const double alpha[KSTAGE][KSTAGE] = {{0, 0, 0, 0}, {a21, 0, 0, 0}, {a31, a32, 0, 0}, {a41, a42, a43, 0}};
  const double gamc[KSTAGE][KSTAGE] = {{0, 0, 0, 0}, {g21, 0, 0, 0}, {g31, g32, 0, 0}, {g41, g42, g43, 0}};
  double Vb[NV * KM], inv[KM * KM], rhs_sav[NV], A[NV * KM], Ks[NV * KSTAGE];
  double ys[NV], F[NV], GK[NV], Bm[NV * NV];
  double time = time_start, dt = dttry, errn = 1.0;
  int it = 0;
  while (time < (time_end - 1.0e-12)) {
    dt = dt < time_end - time ? dt : time_end - time;
    it = it + 1;
    kry_arnoldi<NS>(K, th, tol_iter_nr, vec, dt, t_iter, fls, Vb, inv, rhs_sav, tau_k);
    // A = Vb (I - inv)
    for (int kk = 0; kk < KM; ++kk)
      for (int ll = 0; ll < NV; ++ll) {
        double a = Vb[ll] * ((kk == 0 ? 1.0 : 0.0) - inv[0 + KM * kk]);
        for (int jj = 1; jj < KM; ++jj) a = a + Vb[ll + NV * jj] * ((jj == kk ? 1.0 : 0.0) - inv[jj + KM * kk]);
        A[ll + NV * kk] = a;
      }
    for (int kk = 0; kk < NV; ++kk)
      for (int ll = 0; ll < NV; ++ll) {
        double bv = A[ll] * Vb[kk];
        for (int jj = 1; jj < KM; ++jj) bv = bv + A[ll + NV * jj] * Vb[kk + NV * jj];
        Bm[ll + NV * kk] = bv;
      }
    for (int ii = 0; ii < KSTAGE; ++ii) {
      for (int m = 0; m < NV; ++m) ys[m] = vec[m];
      for (int jj = 0; jj < ii; ++jj)
        for (int m = 0; m < NV; ++m) ys[m] = ys[m] + dt * alpha[ii][jj] * Ks[m + NV * jj];
      kry_rhs<NS>(K, th, tol_iter_nr, ys, t_iter, F, tau_k);
      for (int m = 0; m < NV; ++m) GK[m] = 0.0;
      for (int jj = 0; jj < ii; ++jj)
        for (int m = 0; m < NV; ++m) GK[m] = GK[m] + gamc[ii][jj] / gam * Ks[m + NV * jj];
      for (int m = 0; m < NV; ++m) F[m] = F[m] + GK[m];
      if (fls)
        for (int m = 0; m < NV; ++m) F[m] = F[m] - fls[m];
      // C = B F with B = A Vb^T, in the reference's order
      for (int m = 0; m < NV; ++m) {
        double c = Bm[m] * F[0];
        for (int jj = 1; jj < NV; ++jj) c = c + Bm[m + NV * jj] * F[jj];
        Ks[m + NV * ii] = F[m] - c - GK[m];
      }
    }
    double err2 = 0.0;
    for (int m = 0; m < NV; ++m) {
      double s = vec[m], t = vec[m];
      for (int ii = 0; ii < KSTAGE; ++ii) {
        s = s + dt * bs[ii] * Ks[m + NV * ii];
        t = t + dt * bt[ii] * Ks[m + NV * ii];
      }
      ys[m] = s;
      const double e = (t - s) / ((1.0e-4) * s + 1.0e-8);
      err2 = err2 + e * e;
    }
    const double errmax = sqrt(err2);
    double fac = 0.625 * pow(errn, 0.1) / pow(errmax, 0.175);
    fac = fac > 0.2 ? fac : 0.2;
    fac = fac < 5.0 ? fac : 5.0;
    const double dtnext = dt * fac;
    if (errmax < 1.0) {
      time = time + dt;
      for (int m = 0; m < NV; ++m) vec[m] = ys[m];
      double rho = vec[0];
      for (int s = 1; s < NS; ++s) rho = rho + vec[s];
      const double ri = 1.0 / rho;
      double yy[NS];
      for (int s = 0; s < NS; ++s) yy[s] = vec[s] * ri;
      const double uu = vec[X::I_U - 1] * ri, vv = vec[X::I_V - 1] * ri, ww = vec[X::I_W - 1] * ri;
      const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
      const double ee = vec[X::I_E - 1] * ri - qq;
      t_iter = temperature_from_e<NS>(ee, t_iter, th, tol_iter_nr, yy);
    }
    dt = dtnext;
    errn = errmax;
    if (it > maxsteps) break;
  }
  return it;
}

// 2 waves per SIMD (the compiler's own choice is 1, with 2.8-3.7 KB of per-lane spills):
// the extra resident wave hides the spill latency (one integration of a 128^3 box at 1500 K,
// 25.7 -> 22.2 ms; 3 waves: 26.1 ms; bit-identical)
template <int NS>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) krylov_kernel(Box b, double *__restrict__ w, const double *__restrict__ w_aux,
                                                     Kin K, Thermo th, double tol_iter_nr, double dttry,
                                                     double time_start, double time_end, int maxsteps,
                                                     const double *__restrict__ fl_sav, Pasr pasr) {
  using X = Idx<NS>;
  constexpr int NV = NS + 4;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.fstride()) return;
  const int i = 1 + (int)(t % b.nx);
  const long r = t / b.nx;
  const int j = 1 + (int)(r % b.ny);
  const int k = 1 + (int)(r / b.ny);
  const long pt = b.p(i, j, k), ps = b.pstride(), fs_ = b.fstride();
  const double zb = w_aux[pt + ps * (X::J_Z - 1)];
  if (!(zb > 5.0e-4 && zb < 1.0 - 5.0e-4)) return;
  double vec[NV], fls[NV];
  for (int m = 0; m < NV; ++m) vec[m] = w[pt + ps * m];
  if (fl_sav)
    for (int m = 0; m < NV; ++m) fls[m] = fl_sav[t + fs_ * m];
  const double tau_k = pasr.on ? pasr.tau_k<NS>(w_aux, pt, ps, i, j, k) : 0.0;
  krylov_cell<NS>(K, th, tol_iter_nr, vec, w_aux[pt + ps * (X::J_T - 1)], dttry, time_start, time_end, maxsteps,
                  fl_sav ? fls : nullptr, tau_k);
  for (int m = 0; m < NV; ++m) w[pt + ps * m] = vec[m];
}

}  // namespace streams

using namespace streams;

extern "C" int streams_rosenbrock_krylov(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng, double *w,
                                         double *w_aux, const double *fl, const double *mw, int nreactions,
                                         const double *arr_a, const double *arr_b, const double *arr_ea,
                                         const double *tb_eff, const double *falloff_coeffs, const int *reac_ty,
                                         const int *isrev, const double *r_coeffs, const double *p_coeffs,
                                         const double *kc_tab, int num_t_tab, double t_min_tab, double dt_tab,
                                         double r_univ, double tol_iter_nr, const double *cv_coeff, int nsetcv,
                                         const double *trange, int indx_cp_l, int indx_cp_r, double dttry,
                                         double time_start, double time_end, int simpler_splitting,
                                         const double *fl_sav, int maxsteps, int enable_pasr, double les_c_yoshi,
                                         double les_c_mix, double les_c_eps, const double *dcsidx,
                                         const double *detady, const double *dzitdz, void *stream) {
  (void)fl;
  if (nv != n_s + 4 || nv_aux < 2 * n_s + 13 || maxsteps < 0) return STREAMS_E_DIMS;
  if (enable_pasr > 0 && (nv_aux < 2 * n_s + 14 || !dcsidx || !detady || !dzitdz)) return STREAMS_E_DIMS;
  const Pasr pasr{enable_pasr > 0, les_c_yoshi, les_c_mix, les_c_eps, dcsidx, detady, dzitdz};
  if (simpler_splitting == 1 && fl_sav == nullptr) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  const long n = b.fstride();
  if (n == 0) return 0;
  Kin K{mw, arr_a, arr_b, arr_ea, tb_eff, falloff_coeffs, r_coeffs, p_coeffs, kc_tab, reac_ty, isrev,
        nreactions, num_t_tab, t_min_tab, dt_tab, r_univ};
  Thermo th{cv_coeff, nullptr, trange, nsetcv, indx_cp_l, indx_cp_r};
  const double *fls = simpler_splitting == 1 ? fl_sav : nullptr;
  STREAMS_DISPATCH_NS(n_s, (krylov_kernel<NS><<<blocks_for(n, 128), 128, 0, as_stream(stream)>>>(
                                b, w, w_aux, K, th, tol_iter_nr, dttry, time_start, time_end, maxsteps, fls,
                                pasr)));
  return hip_status();
}
