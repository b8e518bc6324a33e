// euler.hip — hybrid KEEP-central / WENO convective flux along x, y and z.
//
// Structure (per direction):
//   1. shock pre-pass (only when the sensor threshold is below 1, the Ducros sensor's
//      upper bound), in two kernels: one thread per face evaluates the sensor window and
//      appends the shocked faces to a worklist (one atomic per workgroup); then every
//      lane of a persistent grid takes listed faces and parks their characteristic (Roe)
//      or component-wise (Rusanov) WENO flux in fhat.  Shocked faces are a tenth of the
//      faces in a typical shear layer but touch a third of the 64-face waves, so the
//      list keeps the WENO lanes full instead of idling 60 % of each active wave;
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
// Face t of the pre-pass enumeration -> (face f, transverse t1, t2): x faces fastest
// (contiguous along x); y and z with i fastest, then the face, then the other index.
template <int DIR>
__device__ inline void prepass_face(long t, int fs, int nf, int t1n, int &f, int &t1, int &t2) {
  if (DIR == 1) {
    f = fs + (int)(t % nf);
    long r = t / nf;
    t1 = 1 + (int)(r % t1n);
    t2 = 1 + (int)(r / t1n);
  } else {
    t1 = 1 + (int)(t % t1n);
    long r = t / t1n;
    f = fs + (int)(r % nf);
    t2 = 1 + (int)(r / nf);
  }
}

__global__ void zero_count_kernel(unsigned *c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0u;
}

// Sensor window of every face.  Shocked faces go to the worklist: with split set, faces
// reconstructed at the full order WS_FAST fill it from the front (count[0]) and the
// order-reduced ones from the back (count[1]); otherwise all from the front.  Order
// within the list is arbitrary, which changes nothing: each face is computed on its own.
constexpr int FLAG_BS = 256;
constexpr int WS_FAST = 3;   // WENO5, specialised at compile time
template <int NS, int DIR>
__global__ void __launch_bounds__(FLAG_BS) weno_flag_kernel(Box b, EulerArgs A, int t1n, long nfaces, bool split,
                                                            unsigned *__restrict__ list, unsigned *count) {
  __shared__ unsigned wave_n[2][FLAG_BS / 64];
  __shared__ unsigned block_base[2];
  const long t = (long)blockIdx.x * FLAG_BS + threadIdx.x;
  bool s = false, front = true;
  if (t < nfaces) {
    int f, t1, t2;
    prepass_face<DIR>(t, A.fs, A.fe - A.fs + 1, t1n, f, t1, t2);
    Line L{A.w_aux, 0, 0, b.pstride()};
    line_base<DIR>(b, t1, t2, L.base, L.st);
    s = shocked<NS>(L, f, A.weno_scheme, A.sensor_threshold);
    if (s && split) front = epo_at<DIR>(b, A.epo, t1, t2, f) == 0;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long below_mask = (1ull << lane) - 1ull;
  const unsigned long long mf = __ballot(s && front), mb = __ballot(s && !front);
  if (lane == 0) {
    wave_n[0][w] = (unsigned)__popcll(mf);
    wave_n[1][w] = (unsigned)__popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    unsigned tot = 0;
    for (int q = 0; q < FLAG_BS / 64; ++q) tot += wave_n[threadIdx.x][q];
    block_base[threadIdx.x] = tot ? atomicAdd(count + threadIdx.x, tot) : 0u;
  }
  __syncthreads();
  if (s) {
    const int k = front ? 0 : 1;
    unsigned off = block_base[k] + (unsigned)__popcll((front ? mf : mb) & below_mask);
    for (int q = 0; q < w; ++q) off += wave_n[k][q];
    if (front) list[off] = (unsigned)t;
    else list[nfaces - 1 - (long)off] = (unsigned)t;
  }
}

// WENO flux of the listed faces: entries q < count[k] of the front list (k = 0) or of the
// back list (k = 1, stored from list[nfaces - 1] downwards).  The grid is a multiple of the
// XCD count; the workgroups of one XCD share one contiguous eighth of the list, so faces
// that share stencil cells (neighbours in the list) are read through the same L2.
// WS > 0 fixes weno_scheme and the reconstruction order at compile time.
constexpr int LIST_BS = 128;
constexpr unsigned LIST_MAX_BLOCKS = 2048;
template <int NS, int DIR, bool RUS, int WS>
__global__ void __launch_bounds__(LIST_BS) weno_list_kernel(Box b, EulerArgs A, int t1n, long nfaces, int k,
                                                            const unsigned *__restrict__ list, const unsigned *count) {
  constexpr int NV = NS + 4;
  const unsigned n = count[k];
  const unsigned per_xcd = gridDim.x / NXCD;
  const unsigned x = blockIdx.x % NXCD, r = blockIdx.x / NXCD;
  const unsigned seg = (n + NXCD - 1) / NXCD;
  const unsigned lo = x * seg;
  const unsigned hi = lo + seg < n ? lo + seg : n;
  const int nf = A.fe - A.fs + 1;
  for (unsigned q = lo + r * LIST_BS + threadIdx.x; q < hi; q += per_xcd * LIST_BS) {
    const long t = (long)(k == 0 ? list[q] : list[nfaces - 1 - (long)q]);
    int f, t1, t2;
    prepass_face<DIR>(t, A.fs, nf, t1n, f, t1, t2);
    Line L{A.w_aux, 0, 0, b.pstride()};
    line_base<DIR>(b, t1, t2, L.base, L.st);
    double fh[NV];
    if constexpr (WS > 0) {
      weno_flux<NS, DIR, RUS, WS, WS>(L, f, A, WS, fh);
    } else {
      int ord = A.weno_scheme + epo_at<DIR>(b, A.epo, t1, t2, f);
      ord = ord < 1 ? 1 : ord;
      weno_flux<NS, DIR, RUS>(L, f, A, ord, fh);
    }
#pragma unroll
    for (int m = 0; m < NV; ++m) A.fhat[L.base + (long)f * L.st + b.pstride() * m] = fh[m];
  }
}

template <int NS, int DIR, bool RUS>
inline void list_launch(const Box &b, const EulerArgs &A, int t1n, long nfaces, bool split, const unsigned *list,
                        const unsigned *count, hipStream_t st) {
  unsigned nb = blocks_for(nfaces, LIST_BS);
  nb = xcd_grid(nb < LIST_MAX_BLOCKS ? nb : LIST_MAX_BLOCKS);
  if (split) {
    weno_list_kernel<NS, DIR, RUS, WS_FAST><<<nb, LIST_BS, 0, st>>>(b, A, t1n, nfaces, 0, list, count);
    weno_list_kernel<NS, DIR, RUS, 0><<<NXCD * 8, LIST_BS, 0, st>>>(b, A, t1n, nfaces, 1, list, count);
  } else {
    weno_list_kernel<NS, DIR, RUS, 0><<<nb, LIST_BS, 0, st>>>(b, A, t1n, nfaces, 0, list, count);
  }
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
inline int prepass_launch(const Box &b, const EulerArgs &A, int t1n, long nfaces, hipStream_t st) {
  if (nfaces > (long)0xFFFFFF00u) return STREAMS_E_DIMS;
  // stream-ordered worklist: calls on different streams each get their own
  unsigned *buf = nullptr;
  if (hipMallocAsync((void **)&buf, ((size_t)nfaces + 64) * sizeof(unsigned), st) != hipSuccess)
    return STREAMS_E_NOMEM;
  unsigned *count = buf, *list = buf + 64;
  const bool split = A.weno_scheme == WS_FAST;
  zero_count_kernel<<<1, 64, 0, st>>>(count);
  weno_flag_kernel<NS, DIR><<<blocks_for(nfaces, FLAG_BS), FLAG_BS, 0, st>>>(b, A, t1n, nfaces, split, list, count);
  if (A.flux_splitting == 1)
    list_launch<NS, DIR, true>(b, A, t1n, nfaces, split, list, count, st);
  else
    list_launch<NS, DIR, false>(b, A, t1n, nfaces, split, list, count, st);
  const int rc = hip_status();
  if (hipFreeAsync(buf, st) != hipSuccess && rc == 0) return hip_status();
  return rc;
}

int weno_prepass_launch(int dir, int ns, const Box &b, const EulerArgs &A, hipStream_t st) {
  const int nf = A.fe - A.fs + 1;
  const int t1n = dir == 1 ? b.ny : b.nx;
  const int t2n = dir == 3 ? b.ny : b.nz;
  const long nfaces = (long)nf * t1n * t2n;
  if (nfaces <= 0) return 0;
  int rc = 0;
  if (dir == 1) { STREAMS_DISPATCH_NS(ns, (rc = prepass_launch<NS, 1>(b, A, t1n, nfaces, st))); }
  else if (dir == 2) { STREAMS_DISPATCH_NS(ns, (rc = prepass_launch<NS, 2>(b, A, t1n, nfaces, st))); }
  else { STREAMS_DISPATCH_NS(ns, (rc = prepass_launch<NS, 3>(b, A, t1n, nfaces, st))); }
  return rc;
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
    int rc = prepass_launch<NS, DIR>(b, A, t1n, nfaces, st);
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
