// monitor.hip — the run loop's field monitors: the momentum residual, the min/max of
// density, temperature and pressure, and the NaN search that reports where a run blew up
// (compute_residual_cuf, compute_rho_t_p_minmax_cuf, find_boom_cuf; kernels_gpu.F90:5991-6012,
// 6038-6073, 10334-10358).
//
// Each is a two-pass reduction over the interior: a grid-stride pass leaves one partial
// result per block, one block combines them in a fixed order, so a result does not depend
// on the launch (the reference's CUF reductions combine in an unspecified order).
#include "common.h"

namespace streams {

constexpr unsigned kMonBlocks = 1024;

inline unsigned mon_blocks(long n) {
  long b = (n + 255) / 256;
  if (b > (long)kMonBlocks) b = kMonBlocks;
  return (unsigned)(b < 1 ? 1 : b);
}

// Interior point t (i fastest) of the box: its padded index.
__device__ inline long interior_point(const Box &b, long t) {
  const int i = 1 + (int)(t % b.nx);
  const long r = t / b.nx;
  const int j = 1 + (int)(r % b.ny);
  const int k = 1 + (int)(r / b.ny);
  return b.p(i, j, k);
}

// sum over fluid points of (fln(:,:,:,I_U) / dt)^2
__global__ void __launch_bounds__(256) residual_kernel(Box b, const double *__restrict__ flu, double dt,
                                                       const int *__restrict__ mask, double *__restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const long n = b.fstride();
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    if (mask != nullptr && mask[interior_point(b, t)] != 0) continue;
    const double r = flu[t] / dt;
    acc = acc + r * r;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) sum_kernel(int nblk, const double *__restrict__ partial,
                                                  double *__restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int t = threadIdx.x; t < nblk; t += blockDim.x) acc = acc + partial[t];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// (min rho, max rho, min T, max T, min p, max p) over fluid points; the reference starts
// the minima at huge() and the maxima at -100
__global__ void __launch_bounds__(256) minmax_kernel(Box b, const double *__restrict__ rho_,
                                                     const double *__restrict__ t_, const double *__restrict__ p_,
                                                     const int *__restrict__ mask, double *__restrict__ partial) {
  __shared__ double red[6][256];
  double v[6] = {1.7976931348623157e308, -100.0, 1.7976931348623157e308, -100.0, 1.7976931348623157e308, -100.0};
  const long n = b.fstride();
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long pt = interior_point(b, t);
    if (mask != nullptr && mask[pt] != 0) continue;
    const double a[3] = {rho_[pt], t_[pt], p_[pt]};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      v[2 * q] = fmin(v[2 * q], a[q]);
      v[2 * q + 1] = fmax(v[2 * q + 1], a[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        red[2 * q][threadIdx.x] = fmin(red[2 * q][threadIdx.x], red[2 * q][threadIdx.x + s]);
        red[2 * q + 1][threadIdx.x] = fmax(red[2 * q + 1][threadIdx.x], red[2 * q + 1][threadIdx.x + s]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 6; ++q) partial[6 * (long)blockIdx.x + q] = red[q][0];
}

__global__ void __launch_bounds__(256) minmax_combine_kernel(int nblk, const double *__restrict__ partial,
                                                             double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double v[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) v[q] = partial[q];
  for (int t = 1; t < nblk; ++t)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      v[2 * q] = fmin(v[2 * q], partial[6 * (long)t + 2 * q]);
      v[2 * q + 1] = fmax(v[2 * q + 1], partial[6 * (long)t + 2 * q + 1]);
    }
#pragma unroll
  for (int q = 0; q < 6; ++q) out[q] = v[q];
}

// NaN cells of (fln(:,:,:,2) / dt)^2: their count and the smallest interior index among them
__global__ void __launch_bounds__(256) boom_kernel(Box b, const double *__restrict__ fl2, double dt,
                                                   const int *__restrict__ mask, unsigned long long *__restrict__ out) {
  const long n = b.fstride();
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    if (mask != nullptr && mask[interior_point(b, t)] != 0) continue;
    const double r = fl2[t] / dt;
    const double res = r * r;
    if (res != res) {
      atomicAdd(&out[0], 1ull);
      atomicMin(&out[1], (unsigned long long)t);
    }
  }
}

__global__ void boom_init_kernel(unsigned long long *out) {
  out[0] = 0ull;
  out[1] = ~0ull;
}

template <class F>
int with_scratch(size_t bytes, hipStream_t st, F &&body) {
  void *p = nullptr;
  if (hipMallocAsync(&p, bytes, st) != hipSuccess) return STREAMS_E_NOMEM;
  int rc = body(p);
  hipError_t e = hipFreeAsync(p, st);
  if (rc) return rc;
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace streams

using namespace streams;

extern "C" int streams_compute_residual(int nx, int ny, int nz, int ng, int nv, const double *fln, double dt,
                                        double *residual_rhou, const int *fluid_mask, void *stream) {
  if (nx < 1 || ny < 1 || nz < 1 || ng < 0 || nv < 1) return STREAMS_E_DIMS;
  // I_U = N_S + 1 = nv - 3 (1-based): the x momentum
  if (nv < 4) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  const unsigned nblk = mon_blocks(b.fstride());
  hipStream_t st = as_stream(stream);
  return with_scratch(sizeof(double) * (nblk + 1), st, [&](void *p) -> int {
    double *partial = static_cast<double *>(p);
    residual_kernel<<<nblk, 256, 0, st>>>(b, fln + b.fstride() * (long)(nv - 4), dt, fluid_mask, partial);
    sum_kernel<<<1, 256, 0, st>>>((int)nblk, partial, partial + nblk);
    int rc = hip_status();
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(residual_rhou, partial + nblk, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? 0 : (int)e;
  });
}

extern "C" int streams_compute_rho_t_p_minmax(int n_s, int nx, int ny, int nz, int ng, int nv_aux,
                                              const double *w_aux, double *rhomin, double *rhomax, double *tmin,
                                              double *tmax, double *pmin, double *pmax, const int *fluid_mask,
                                              void *stream) {
  if (nx < 1 || ny < 1 || nz < 1 || ng < 0 || n_s < 1) return STREAMS_E_DIMS;
  if (nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  const long ps = b.pstride();
  const unsigned nblk = mon_blocks(b.fstride());
  hipStream_t st = as_stream(stream);
  // J_R = N_S + 11, J_T = N_S + 5, J_P = N_S + 6 (1-based)
  const double *rho_ = w_aux + ps * (long)(n_s + 10);
  const double *t_ = w_aux + ps * (long)(n_s + 4);
  const double *p_ = w_aux + ps * (long)(n_s + 5);
  return with_scratch(sizeof(double) * (6 * (size_t)nblk + 6), st, [&](void *p) -> int {
    double *partial = static_cast<double *>(p);
    minmax_kernel<<<nblk, 256, 0, st>>>(b, rho_, t_, p_, fluid_mask, partial);
    minmax_combine_kernel<<<1, 64, 0, st>>>((int)nblk, partial, partial + 6 * (size_t)nblk);
    int rc = hip_status();
    if (rc) return rc;
    double out[6];
    hipError_t e = hipMemcpyAsync(out, partial + 6 * (size_t)nblk, sizeof(out), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return (int)e;
    *rhomin = out[0];
    *rhomax = out[1];
    *tmin = out[2];
    *tmax = out[3];
    *pmin = out[4];
    *pmax = out[5];
    return 0;
  });
}

extern "C" int streams_find_boom(int nx, int ny, int nz, int ng, int nv, const double *fln, double dt,
                                 const int *fluid_mask, int ncoords1, int ncoords3, int *nboom, int *first_ijk,
                                 void *stream) {
  if (nx < 1 || ny < 1 || nz < 1 || ng < 0 || nv < 2) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  hipStream_t st = as_stream(stream);
  return with_scratch(2 * sizeof(unsigned long long), st, [&](void *p) -> int {
    unsigned long long *out = static_cast<unsigned long long *>(p);
    boom_init_kernel<<<1, 1, 0, st>>>(out);
    // variable 2 of fln, as the reference's find_boom_cuf reads it
    boom_kernel<<<mon_blocks(b.fstride()), 256, 0, st>>>(b, fln + b.fstride(), dt, fluid_mask, out);
    int rc = hip_status();
    if (rc) return rc;
    unsigned long long h[2];
    hipError_t e = hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return (int)e;
    *nboom = (int)(h[0] > 0x7fffffffull ? 0x7fffffffull : h[0]);
    if (h[0] > 0) {
      const long t = (long)h[1];
      first_ijk[0] = 1 + (int)(t % nx) + ncoords1 * nx;
      first_ijk[1] = 1 + (int)((t / nx) % ny);
      first_ijk[2] = 1 + (int)(t / ((long)nx * ny)) + ncoords3 * nz;
    } else {
      first_ijk[0] = first_ijk[1] = first_ijk[2] = 0;
    }
    return 0;
  });
}
