// rk.hip — Wray low-storage RK3 bookkeeping (plain and simpler splitting) and the
// stable-time-step reduction.
//
// The bookkeeping kernels are pure streams over the interior; they are written as
// grid-stride loops over (cell, variable) with i fastest so that each wave moves 512
// contiguous bytes per array.
#include "common.h"
#include "dt_reduce.h"

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

// simpler splitting (operator_splitting = 1): the transport forcing fl_sav is taken out of
// the flux that the RK stages accumulate
__global__ void __launch_bounds__(256) init_flux_simpler_kernel(long n, double *__restrict__ fl,
                                                                double *__restrict__ fln,
                                                                const double *__restrict__ fl_sav, double rhodt) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    double f = fl[t] - fl_sav[t];
    fln[t] = -rhodt * f;
    fl[t] = 0.0;
  }
}

__global__ void __launch_bounds__(256) update_simpler_flux_kernel(long n, const double *__restrict__ fl,
                                                                  double *__restrict__ fln,
                                                                  const double *__restrict__ fl_sav, double gamdt) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    fln[t] = fln[t] - gamdt * (fl[t] - fl_sav[t]);
}

__global__ void __launch_bounds__(256) copy_kernel(long n, const double *__restrict__ src, double *__restrict__ dst) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    dst[t] = src[t];
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

extern "C" int streams_init_flux_simpler(int nx, int ny, int nz, int nv, double *fl, double *fln,
                                         const double *fl_sav, double rhodt, void *stream) {
  long n = (long)nx * ny * nz * nv;
  if (n <= 0) return 0;
  init_flux_simpler_kernel<<<stream_blocks(n), 256, 0, as_stream(stream)>>>(n, fl, fln, fl_sav, rhodt);
  return hip_status();
}

extern "C" int streams_sav_flx(int nx, int ny, int nz, int nv, const double *fl, double *fl_sav, void *stream) {
  long n = (long)nx * ny * nz * nv;
  if (n <= 0) return 0;
  copy_kernel<<<stream_blocks(n), 256, 0, as_stream(stream)>>>(n, fl, fl_sav);
  return hip_status();
}

extern "C" int streams_update_simpler_flux(int nx, int ny, int nz, int nv, const double *fl, double *fln,
                                           const double *fl_sav, double gamdt, void *stream) {
  long n = (long)nx * ny * nz * nv;
  if (n <= 0) return 0;
  update_simpler_flux_kernel<<<stream_blocks(n), 256, 0, as_stream(stream)>>>(n, fl, fln, fl_sav, gamdt);
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
  double *partial = nullptr;
  hipStream_t st = as_stream(stream);
  if (hipMallocAsync((void **)&partial, sizeof(double) * (9 * (size_t)dt_blocks(b.fstride()) + 9), st) != hipSuccess)
    return STREAMS_E_NOMEM;
  int rc = compute_dt_into(n_s, b, dcsidx, detady, dzitdz, dcsidxs, detadys, dzitdzs, w_aux, dtmax9, th, fluid_mask,
                           partial, st);
  hipError_t e = hipFreeAsync(partial, st);
  if (rc) return rc;
  return e == hipSuccess ? 0 : (int)e;
}
