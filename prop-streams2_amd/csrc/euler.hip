// euler.hip — hybrid KEEP-central / WENO convective flux along x, y and z.
//
// Structure (per direction):
//   1. shock pre-pass (only when the sensor threshold is below 1, the Ducros sensor's
//      upper bound): one thread per face evaluates the sensor window and, on shocked
//      faces, the characteristic (Roe) or component-wise (Rusanov) WENO flux, which it
//      parks in fhat;
//   2. fused central pass: every face's KEEP split-form flux (or the parked WENO flux) and
//      the flux-difference update of fl in the same kernel, so face fluxes never round-
//      trip through HBM:
//        x: faces of all (j,k) lines are flattened and tiled 256 to a workgroup with a
//           one-face overlap; fluxes meet in LDS and each thread updates its cell;
//        y, z: each thread marches a segment of one (i, .) column, keeping the previous
//           face flux in registers; lanes run along i so every load is coalesced.
// Arithmetic follows euler_x_fluxes_hybrid_kernel (kernels_gpu.F90:88-425),
// euler_y_hybrid_kernel (:1427-1759), euler_z_hybrid_kernel (:795-1123), the Rusanov
// variants (:427-730, :1125-1425, :1761-2063), compute_roe_average (:2986-3295),
// eigenvectors_x/y/z (:2703-2984), wenorec_1d[_rusanov] (:2065-2637).
#include "euler_dev.h"

namespace streams {

// ---------------------------------------------------------------- shock pre-pass
// one thread per face of every line; shocked faces get their WENO flux in fhat
template <int NS, int DIR, bool RUS>
__global__ void __launch_bounds__(128) weno_prepass_kernel(Box b, EulerArgs A, int t1n, long nfaces) {
  using X = Idx<NS>;
  constexpr int NV = NS + 4;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nf = A.fe - A.fs + 1;
  if (t >= nfaces) return;
  int f, t1, t2;
  if (DIR == 1) {  // faces fastest (contiguous along x)
    f = A.fs + (int)(t % nf);
    long r = t / nf;
    t1 = 1 + (int)(r % t1n);
    t2 = 1 + (int)(r / t1n);
  } else {  // i fastest
    t1 = 1 + (int)(t % t1n);
    long r = t / t1n;
    f = A.fs + (int)(r % nf);
    t2 = 1 + (int)(r / nf);
  }
  Line L{A.w_aux, 0, 0, b.pstride()};
  line_base<DIR>(b, t1, t2, L.base, L.st);
  if (!shocked<NS>(L, f, A.weno_scheme, A.sensor_threshold)) return;
  int ord = A.weno_scheme + epo_at<DIR>(b, A.epo, t1, t2, f);
  ord = ord < 1 ? 1 : ord;
  double fh[NV];
  weno_flux<NS, DIR, RUS>(L, f, A, ord, fh);
  long fb, fst;
  line_base<DIR>(b, t1, t2, fb, fst);
#pragma unroll
  for (int m = 0; m < NV; ++m) A.fhat[fb + (long)f * fst + b.pstride() * m] = fh[m];
  (void)X::J_R;
}

// ---------------------------------------------------------------- fused x sweep
constexpr int XT = 256;  // faces per workgroup (XT-1 cells updated)

template <int NS>
__global__ void __launch_bounds__(XT) euler_x_fused_kernel(Box b, EulerArgs A, long nfaces, bool check_shock) {
  constexpr int NV = NS + 4;
  __shared__ double sf[NV][XT];
  const int nf = A.fe - A.fs + 1;
  const long q = (long)blockIdx.x * (XT - 1) + threadIdx.x;
  const bool valid = q < nfaces;
  int f = 0, j = 1, k = 1;
  if (valid) {
    f = A.fs + (int)(q % nf);
    long r = q / nf;
    j = 1 + (int)(r % b.ny);
    k = 1 + (int)(r / b.ny);
    Line L{A.w_aux, 0, 0, b.pstride()};
    line_base<1>(b, j, k, L.base, L.st);
    double fh[NV];
    face_flux<NS, 1>(b, A, L, j, k, f, check_shock, fh);
#pragma unroll
    for (int m = 0; m < NV; ++m) sf[m][threadIdx.x] = fh[m];
  }
  __syncthreads();
  if (!valid || threadIdx.x == 0 || f == A.fs) return;
  const int c = f;  // cell between faces c-1 and c
  const double met = A.metric[c - 1];
  const long fc = b.f(c, j, k), fs_ = b.fstride();
#pragma unroll
  for (int m = 0; m < NV; ++m)
    A.fl[fc + fs_ * m] = A.fl[fc + fs_ * m] + (sf[m][threadIdx.x] - sf[m][threadIdx.x - 1]) * met;
}

// ------------------------------------------------------------- fused y / z sweeps
template <int NS, int DIR>
__global__ void __launch_bounds__(128) euler_yz_fused_kernel(Box b, EulerArgs A, int seg, int nseg, long nthreads,
                                                            bool check_shock) {
  constexpr int NV = NS + 4;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nthreads) return;
  const int t1n = b.nx;                 // i fastest
  const int t2n = DIR == 2 ? b.nz : b.ny;
  int t1 = 1 + (int)(t % t1n);
  long r = t / t1n;
  int t2 = 1 + (int)(r % t2n);
  int sg = (int)(r / t2n);
  (void)nseg;
  const int us = A.fs + 1, ue = A.fe;   // updated cells
  int c0 = us + sg * seg;
  int c1 = c0 + seg - 1;
  if (c1 > ue) c1 = ue;
  if (c0 > c1) return;
  Line L{A.w_aux, 0, 0, b.pstride()};
  line_base<DIR>(b, t1, t2, L.base, L.st);
  double prev[NV], cur[NV];
  face_flux<NS, DIR>(b, A, L, t1, t2, c0 - 1, check_shock, prev);
  const long fs_ = b.fstride();
  for (int c = c0; c <= c1; ++c) {
    face_flux<NS, DIR>(b, A, L, t1, t2, c, check_shock, cur);
    const double met = A.metric[c - 1];
    long fc = DIR == 2 ? b.f(t1, c, t2) : b.f(t1, t2, c);
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      A.fl[fc + fs_ * m] = A.fl[fc + fs_ * m] + (cur[m] - prev[m]) * met;
      prev[m] = cur[m];
    }
  }
}

template <int NS, int DIR>
inline void prepass_launch(const Box &b, const EulerArgs &A, int t1n, long nfaces, hipStream_t st) {
  if (A.flux_splitting == 1)
    weno_prepass_kernel<NS, DIR, true><<<blocks_for(nfaces, 128), 128, 0, st>>>(b, A, t1n, nfaces);
  else
    weno_prepass_kernel<NS, DIR, false><<<blocks_for(nfaces, 128), 128, 0, st>>>(b, A, t1n, nfaces);
}

int weno_prepass_launch(int dir, int ns, const Box &b, const EulerArgs &A, hipStream_t st) {
  const int nf = A.fe - A.fs + 1;
  const int t1n = dir == 1 ? b.ny : b.nx;
  const int t2n = dir == 3 ? b.ny : b.nz;
  const long nfaces = (long)nf * t1n * t2n;
  if (nfaces <= 0) return 0;
  if (dir == 1) { STREAMS_DISPATCH_NS(ns, (prepass_launch<NS, 1>(b, A, t1n, nfaces, st))); }
  else if (dir == 2) { STREAMS_DISPATCH_NS(ns, (prepass_launch<NS, 2>(b, A, t1n, nfaces, st))); }
  else { STREAMS_DISPATCH_NS(ns, (prepass_launch<NS, 3>(b, A, t1n, nfaces, st))); }
  return hip_status();
}

template <int NS, int DIR>
int euler_launch(const Box &b, EulerArgs A, hipStream_t st) {
  const int nf = A.fe - A.fs + 1;
  if (nf < 2) return 0;
  const bool check = A.sensor_threshold < 1.0;
  const int t1n = DIR == 1 ? b.ny : b.nx;
  const int t2n = DIR == 3 ? b.ny : b.nz;
  const long nfaces = (long)nf * t1n * t2n;
  if (check) {
    prepass_launch<NS, DIR>(b, A, t1n, nfaces, st);
    int rc = hip_status();
    if (rc) return rc;
  }
  if (DIR == 1) {
    long nblk = (nfaces - 1 + (XT - 2)) / (XT - 1);
    euler_x_fused_kernel<NS><<<(unsigned)nblk, XT, 0, st>>>(b, A, nfaces, check);
  } else {
    const int ncell = nf - 1;
    const int seg = 16;
    const int nseg = (ncell + seg - 1) / seg;
    long nth = (long)b.nx * t2n * nseg;
    euler_yz_fused_kernel<NS, DIR><<<blocks_for(nth, 128), 128, 0, st>>>(b, A, seg, nseg, nth, check);
  }
  return hip_status();
}

}  // namespace streams

using namespace streams;

static int euler_common(int dir, int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng, int fs, int fe,
                        int us, int ue, int lmax_base, int nkeep, const double *rgas, const double *w_aux,
                        const double *cd1, const double *metric, double *fhat, int zmin, int zmax,
                        int weno_scheme, int weno_version, double sensor_threshold, const double *cv_coeff,
                        const double *cp_coeff, int nsetcv, const double *trange, int lo, int hi,
                        const int *epo, double tol_iter_nr, double rho0, double u0, int flux_splitting,
                        double *fl, void *stream) {
  if (nv != n_s + 4 || nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  if (lmax_base < 1 || lmax_base > 4 || weno_scheme < 1 || weno_scheme > 4) return STREAMS_E_OPTION;
  int n = dir == 1 ? nx : dir == 2 ? ny : nz;
  if (ue < us) return 0;
  // fused path: faces us-1..ue feed the update of cells us..ue
  if (fs != us - 1 || fe != ue) return STREAMS_E_DIMS;
  if (us < 1 || ue > n) return STREAMS_E_DIMS;
  if (ng < lmax_base || ng < weno_scheme) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  EulerArgs A{w_aux, cd1, metric, rgas, fhat, fl, epo, Thermo{cv_coeff, cp_coeff, trange, nsetcv, lo, hi},
              lmax_base, nkeep, zmin, zmax, weno_scheme, weno_version, flux_splitting, sensor_threshold,
              tol_iter_nr, rho0, u0, fs, fe, n};
  hipStream_t st = as_stream(stream);
  if (dir == 1) { STREAMS_DISPATCH_NS(n_s, return (euler_launch<NS, 1>(b, A, st))); }
  if (dir == 2) { STREAMS_DISPATCH_NS(n_s, return (euler_launch<NS, 2>(b, A, st))); }
  STREAMS_DISPATCH_NS(n_s, return (euler_launch<NS, 3>(b, A, st)));
  return 0;
}

extern "C" int streams_euler_x(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng, int istart_face,
                               int iend_face, int lmax_base, int nkeep, const double *rgas, const double *w_aux,
                               const double *coeff_deriv1, const double *dcsidx, double *fhat,
                               int force_zero_flux_min, int force_zero_flux_max, int weno_scheme,
                               int weno_version, double sensor_threshold, int weno_size, const double *cv_coeff,
                               const double *cp_coeff, int nsetcv, const double *trange, int indx_cp_l,
                               int indx_cp_r, const int *ep_ord_change, double tol_iter_nr, double rho0,
                               double u0, int flux_splitting, int eul_imin, int eul_imax, double *fl,
                               void *stream) {
  if (weno_size != 2 * weno_scheme) return STREAMS_E_OPTION;
  return euler_common(1, n_s, nv, nv_aux, nx, ny, nz, ng, istart_face, iend_face, eul_imin, eul_imax, lmax_base,
                      nkeep, rgas, w_aux, coeff_deriv1, dcsidx, fhat, force_zero_flux_min, force_zero_flux_max,
                      weno_scheme, weno_version, sensor_threshold, cv_coeff, cp_coeff, nsetcv, trange,
                      indx_cp_l, indx_cp_r, ep_ord_change, tol_iter_nr, rho0, u0, flux_splitting, fl, stream);
}

extern "C" int streams_euler_y(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng, int eul_jmin,
                               int eul_jmax, int lmax_base, int nkeep, const double *rgas, const double *w_aux,
                               double *fl, const double *coeff_deriv1, const double *detady, double *fhat,
                               int force_zero_flux_min, int force_zero_flux_max, int weno_scheme,
                               int weno_version, double sensor_threshold, int weno_size, const double *cv_coeff,
                               const double *cp_coeff, int nsetcv, const double *trange, int indx_cp_l,
                               int indx_cp_r, const int *ep_ord_change, double tol_iter_nr, double rho0,
                               double u0, int flux_splitting, void *stream) {
  if (weno_size != 2 * weno_scheme) return STREAMS_E_OPTION;
  return euler_common(2, n_s, nv, nv_aux, nx, ny, nz, ng, eul_jmin - 1, eul_jmax, eul_jmin, eul_jmax, lmax_base,
                      nkeep, rgas, w_aux, coeff_deriv1, detady, fhat, force_zero_flux_min, force_zero_flux_max,
                      weno_scheme, weno_version, sensor_threshold, cv_coeff, cp_coeff, nsetcv, trange,
                      indx_cp_l, indx_cp_r, ep_ord_change, tol_iter_nr, rho0, u0, flux_splitting, fl, stream);
}

extern "C" int streams_euler_z(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng, int eul_kmin,
                               int eul_kmax, int lmax_base, int nkeep, const double *rgas, const double *w_aux,
                               double *fl, const double *coeff_deriv1, const double *dzitdz, double *fhat,
                               int force_zero_flux_min, int force_zero_flux_max, int weno_scheme,
                               int weno_version, double sensor_threshold, int weno_size, const double *cv_coeff,
                               const double *cp_coeff, int nsetcv, const double *trange, int indx_cp_l,
                               int indx_cp_r, const int *ep_ord_change, double tol_iter_nr, double rho0,
                               double u0, int flux_splitting, void *stream) {
  if (weno_size != 2 * weno_scheme) return STREAMS_E_OPTION;
  return euler_common(3, n_s, nv, nv_aux, nx, ny, nz, ng, eul_kmin - 1, eul_kmax, eul_kmin, eul_kmax, lmax_base,
                      nkeep, rgas, w_aux, coeff_deriv1, dzitdz, fhat, force_zero_flux_min, force_zero_flux_max,
                      weno_scheme, weno_version, sensor_threshold, cv_coeff, cp_coeff, nsetcv, trange,
                      indx_cp_l, indx_cp_r, ep_ord_change, tol_iter_nr, rho0, u0, flux_splitting, fl, stream);
}
