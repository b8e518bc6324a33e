// dt_reduce.h — the stable-time-step reduction of compute_dt_cuf (kernels_gpu.F90:6075-6152):
// per-block maxima of the nine inverse time steps, then one block that reduces the block
// results.  Shared by streams_compute_dt (scratch from the stream-ordered allocator) and its
// allocation-free variant (scratch in a fixed device array).
#pragma once
#include "common.h"

namespace streams {

// partial results hold at most this many blocks of nine maxima; the grid-stride loop covers
// any box with that many blocks
constexpr unsigned kDtMaxBlocks = 2048;

inline unsigned dt_blocks(long n) {
  long b = (n + 255) / 256;
  if (b > (long)kDtMaxBlocks) b = kDtMaxBlocks;
  return (unsigned)(b < 1 ? 1 : b);
}

// nine maxima per block, then a second pass over the block results
template <int NS>
__global__ void __launch_bounds__(256) compute_dt_kernel(Box b, const double *__restrict__ w_aux,
                                                         const double *dcsidx, const double *detady,
                                                         const double *dzitdz, const double *dcsidxs,
                                                         const double *detadys, const double *dzitdzs,
                                                         Thermo th, const int *__restrict__ mask,
                                                         double *__restrict__ partial) {
  using X = Idx<NS>;
  double mx[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const long n = b.fstride(), ps = b.pstride();
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i = 1 + (int)(t % b.nx);
    long r = t / b.nx;
    int j = 1 + (int)(r % b.ny);
    int k = 1 + (int)(r / b.ny);
    long pt = b.p(i, j, k);
    if (mask != nullptr && mask[pt] != 0) continue;
    double rho = w_aux[pt + ps * (X::J_R - 1)], ri = 1.0 / rho;
    double uu = w_aux[pt + ps * (X::J_U - 1)], vv = w_aux[pt + ps * (X::J_V - 1)];
    double ww = w_aux[pt + ps * (X::J_W - 1)], tt = w_aux[pt + ps * (X::J_T - 1)];
    double mu = w_aux[pt + ps * (X::J_MU - 1)], c = w_aux[pt + ps * (X::J_C - 1)];
    double kc = w_aux[pt + ps * (X::J_K_COND - 1)];
    double yy[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) yy[s] = w_aux[pt + ps * s];
    double cploc = mixture_cp<NS>(tt, th, yy);
    double nu = ri * mu, kr = kc * ri / cploc;
    double v[9] = {(fabs(uu) + c) * dcsidx[i - 1], (fabs(vv) + c) * detady[j - 1], (fabs(ww) + c) * dzitdz[k - 1],
                   nu * dcsidxs[i - 1], nu * detadys[j - 1], nu * dzitdzs[k - 1],
                   kr * dcsidxs[i - 1], kr * detadys[j - 1], kr * dzitdzs[k - 1]};
#pragma unroll
    for (int q = 0; q < 9; ++q) mx[q] = fmax(mx[q], v[q]);
  }
  __shared__ double red[9][256];
#pragma unroll
  for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = mx[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = fmax(red[q][threadIdx.x], red[q][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 9; ++q) partial[(long)blockIdx.x * 9 + q] = red[q][0];
}

static __global__ void __launch_bounds__(256) reduce9_kernel(int nblk, const double *__restrict__ partial,
                                                      double *__restrict__ out) {
  __shared__ double red[9][256];
  double mx[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < nblk; t += blockDim.x)
#pragma unroll
    for (int q = 0; q < 9; ++q) mx[q] = fmax(mx[q], partial[(long)t * 9 + q]);
#pragma unroll
  for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = mx[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = fmax(red[q][threadIdx.x], red[q][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 9; ++q) out[q] = red[q][0];
}

// Launches the block maxima and the final reduction of compute_dt on `st` into `partial`
// (at least 9 * dt_blocks(n) + 9 doubles of device memory; every block writes its own nine,
// so nothing needs clearing), copies the nine results to host `dtmax9` and waits for `st`.
inline int compute_dt_into(int n_s, const Box &b, const double *dcsidx, const double *detady,
                           const double *dzitdz, const double *dcsidxs, const double *detadys,
                           const double *dzitdzs, const double *w_aux, double *dtmax9, const Thermo &th,
                           const int *fluid_mask, double *partial, hipStream_t st) {
  const unsigned nblk = dt_blocks(b.fstride());
  STREAMS_DISPATCH_NS(n_s, (compute_dt_kernel<NS><<<nblk, 256, 0, st>>>(b, w_aux, dcsidx, detady, dzitdz, dcsidxs,
                                                                          detadys, dzitdzs, th, fluid_mask,
                                                                          partial)));
  reduce9_kernel<<<1, 256, 0, st>>>((int)nblk, partial, partial + 9 * (size_t)nblk);
  int rc = hip_status();
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(dtmax9, partial + 9 * (size_t)nblk, 9 * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return (int)e;
  e = hipStreamSynchronize(st);
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace streams
