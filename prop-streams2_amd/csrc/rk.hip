// rk.hip — Wray low-storage RK3 bookkeeping and the stable-time-step reduction.
//
// The bookkeeping kernels are pure streams over the interior; they are written as
// grid-stride loops over (cell, variable) with i fastest so that each wave moves 512
// contiguous bytes per array.
#include "common.h"

namespace streams {

__global__ void __launch_bounds__(256) init_flux_kernel(long n, double *__restrict__ fl, double *__restrict__ fln,
                                                        double rhodt) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    double f = fl[t];
    fln[t] = -rhodt * f;
    fl[t] = 0.0;
  }
}

__global__ void __launch_bounds__(256) update_flux_kernel(long n, const double *__restrict__ fl,
                                                          double *__restrict__ fln, double gamdt) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    fln[t] = fln[t] - gamdt * fl[t];
}

__global__ void __launch_bounds__(256) update_field_kernel(Box b, int nv, double *__restrict__ w,
                                                           const double *__restrict__ fln,
                                                           const int *__restrict__ mask) {
  const long n = b.fstride();
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n * nv; t += (long)gridDim.x * blockDim.x) {
    long c = t % n;
    int m = (int)(t / n);
    int i = 1 + (int)(c % b.nx);
    long r = c / b.nx;
    int j = 1 + (int)(r % b.ny);
    int k = 1 + (int)(r / b.ny);
    long pt = b.p(i, j, k);
    if (mask == nullptr || mask[pt] == 0) w[pt + b.pstride() * m] = w[pt + b.pstride() * m] + fln[t];
  }
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

__global__ void __launch_bounds__(256) reduce9_kernel(int nblk, const double *__restrict__ partial,
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

static unsigned stream_blocks(long n) {
  long b = (n + 255) / 256;
  return (unsigned)(b < 65536 ? (b < 1 ? 1 : b) : 65536);
}

}  // namespace streams

using namespace streams;

extern "C" int streams_init_flux(int nx, int ny, int nz, int nv, double *fl, double *fln, double rhodt,
                                 void *stream) {
  long n = (long)nx * ny * nz * nv;
  if (n <= 0) return 0;
  init_flux_kernel<<<stream_blocks(n), 256, 0, as_stream(stream)>>>(n, fl, fln, rhodt);
  return hip_status();
}

extern "C" int streams_update_flux(int nx, int ny, int nz, int nv, const double *fl, double *fln,
                                   double gamdt, void *stream) {
  long n = (long)nx * ny * nz * nv;
  if (n <= 0) return 0;
  update_flux_kernel<<<stream_blocks(n), 256, 0, as_stream(stream)>>>(n, fl, fln, gamdt);
  return hip_status();
}

extern "C" int streams_update_field(int nx, int ny, int nz, int ng, int nv, double *w, const double *fln,
                                    const int *fluid_mask, void *stream) {
  Box b{nx, ny, nz, ng};
  long n = b.fstride() * nv;
  if (n <= 0) return 0;
  update_field_kernel<<<stream_blocks(n), 256, 0, as_stream(stream)>>>(b, nv, w, fln, fluid_mask);
  return hip_status();
}

extern "C" int streams_compute_dt(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                                  const double *dcsidx, const double *detady, const double *dzitdz,
                                  const double *dcsidxs, const double *detadys, const double *dzitdzs,
                                  const double *w, const double *w_aux, double *dtmax9, int indx_cp_l,
                                  int indx_cp_r, const double *cp_coeff, int nsetcv, const double *trange,
                                  const int *fluid_mask, void *stream) {
  (void)w; (void)nv;
  if (nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  Thermo th{nullptr, cp_coeff, trange, nsetcv, indx_cp_l, indx_cp_r};
  long n = b.fstride();
  unsigned nblk = (unsigned)((n + 255) / 256);
  if (nblk > 2048) nblk = 2048;
  if (nblk < 1) nblk = 1;
  double *partial = nullptr;
  hipStream_t st = as_stream(stream);
  if (hipMallocAsync((void **)&partial, sizeof(double) * (9 * (size_t)nblk + 9), st) != hipSuccess)
    return STREAMS_E_NOMEM;
  int rc = 0;
  auto body = [&]() -> int {
    STREAMS_DISPATCH_NS(n_s, (compute_dt_kernel<NS><<<nblk, 256, 0, st>>>(b, w_aux, dcsidx, detady, dzitdz,
                                                                            dcsidxs, detadys, dzitdzs, th,
                                                                            fluid_mask, partial)));
    return 0;
  };
  rc = body();
  if (rc == 0) {
    reduce9_kernel<<<1, 256, 0, st>>>((int)nblk, partial, partial + 9 * (size_t)nblk);
    rc = hip_status();
  }
  if (rc == 0 && hipMemcpyAsync(dtmax9, partial + 9 * (size_t)nblk, 9 * sizeof(double), hipMemcpyDeviceToHost,
                                st) != hipSuccess)
    rc = hip_status();
  hipError_t e1 = hipFreeAsync(partial, st);
  hipError_t e2 = hipStreamSynchronize(st);
  if (rc) return rc;
  if (e1 != hipSuccess) return (int)e1;
  return e2 == hipSuccess ? 0 : (int)e2;
}
