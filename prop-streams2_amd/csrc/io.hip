// io.hip — host <-> device copies of the conserved field with the change of layout that the
// reference does on the host (copy_cpu_gpu / copy_gpu_cpu, base_gpu.F90:1266-1325): the
// host field is w(nv, 1-ng:nx+ng, 1-ng:ny+ng, 1-ng:nz+ng) (variable fastest, field.F90:289),
// the device field w_gpu(1-ng:.., .., .., nv) (point fastest).  Here the host array moves
// as one contiguous copy and the transposition runs on the device: a workgroup takes 64
// consecutive points, reads their nv-long records (one contiguous stretch of 64*nv doubles)
// into LDS and writes each variable's 64 values as one coalesced segment (and the reverse).
// Used for restarts (rst1_*.bin holds the host layout of the interior) and for moving a
// whole state between host and device.
#include "common.h"

namespace streams {

constexpr int TP = 64;          // points per workgroup
constexpr int TP_MAXV = 32;     // largest nv staged (N_S + 4 with N_S <= 28)

// AoS (nv, npts) -> SoA (npts, nv) with point stride ps_out and variable stride
__global__ void __launch_bounds__(256) aos_to_soa_kernel(long npts, int nv, const double *__restrict__ aos,
                                                         double *__restrict__ soa, long vstride) {
  __shared__ double t[TP * TP_MAXV];
  const long p0 = (long)blockIdx.x * TP;
  const int np = (int)(npts - p0 < TP ? npts - p0 : TP);
  for (int q = threadIdx.x; q < np * nv; q += blockDim.x) t[q] = aos[p0 * nv + q];
  __syncthreads();
  for (int q = threadIdx.x; q < TP * nv; q += blockDim.x) {
    const int m = q / TP, p = q % TP;
    if (p < np) soa[p0 + p + vstride * m] = t[p * nv + m];
  }
}

__global__ void __launch_bounds__(256) soa_to_aos_kernel(long npts, int nv, const double *__restrict__ soa,
                                                         double *__restrict__ aos, long vstride) {
  __shared__ double t[TP * TP_MAXV];
  const long p0 = (long)blockIdx.x * TP;
  const int np = (int)(npts - p0 < TP ? npts - p0 : TP);
  for (int q = threadIdx.x; q < TP * nv; q += blockDim.x) {
    const int m = q / TP, p = q % TP;
    if (p < np) t[p * nv + m] = soa[p0 + p + vstride * m];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < np * nv; q += blockDim.x) aos[p0 * nv + q] = t[q];
}

// interior of a padded SoA field <-> a dense SoA block (nx, ny, nz, nv)
__global__ void __launch_bounds__(256) interior_copy_kernel(Box b, int nv, const double *__restrict__ src,
                                                            double *__restrict__ dst, int to_dense) {
  const long n = b.fstride();
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n * nv; t += (long)gridDim.x * blockDim.x) {
    const long c = t % n;
    const int m = (int)(t / n);
    const int i = 1 + (int)(c % b.nx);
    const long r = c / b.nx;
    const int j = 1 + (int)(r % b.ny);
    const int k = 1 + (int)(r / b.ny);
    const long pt = b.p(i, j, k) + b.pstride() * m;
    if (to_dense) dst[t] = src[pt];
    else dst[pt] = src[t];
  }
}

inline unsigned grid_for(long n) {
  long b = (n + 255) / 256;
  return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace streams

using namespace streams;

extern "C" int streams_copy_cpu_gpu(int nx, int ny, int nz, int ng, int nv, const double *w_host, double *w_gpu,
                                    void *stream) {
  if (nx < 1 || ny < 1 || nz < 1 || ng < 0 || nv < 1 || nv > TP_MAXV) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  const long npts = b.pstride();
  hipStream_t st = as_stream(stream);
  double *stage = nullptr;
  if (hipMallocAsync((void **)&stage, sizeof(double) * npts * nv, st) != hipSuccess) return STREAMS_E_NOMEM;
  hipError_t e = hipMemcpyAsync(stage, w_host, sizeof(double) * npts * nv, hipMemcpyHostToDevice, st);
  int rc = e == hipSuccess ? 0 : (int)e;
  if (rc == 0) {
    aos_to_soa_kernel<<<(unsigned)((npts + TP - 1) / TP), 256, 0, st>>>(npts, nv, stage, w_gpu, npts);
    rc = hip_status();
  }
  e = hipFreeAsync(stage, st);
  if (rc) return rc;
  if (e != hipSuccess) return (int)e;
  e = hipStreamSynchronize(st);
  return e == hipSuccess ? 0 : (int)e;
}

extern "C" int streams_copy_gpu_cpu(int nx, int ny, int nz, int ng, int nv, const double *w_gpu, double *w_host,
                                    void *stream) {
  if (nx < 1 || ny < 1 || nz < 1 || ng < 0 || nv < 1 || nv > TP_MAXV) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  const long npts = b.pstride();
  hipStream_t st = as_stream(stream);
  double *stage = nullptr;
  if (hipMallocAsync((void **)&stage, sizeof(double) * npts * nv, st) != hipSuccess) return STREAMS_E_NOMEM;
  soa_to_aos_kernel<<<(unsigned)((npts + TP - 1) / TP), 256, 0, st>>>(npts, nv, w_gpu, stage, npts);
  int rc = hip_status();
  if (rc == 0) {
    hipError_t e = hipMemcpyAsync(w_host, stage, sizeof(double) * npts * nv, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) rc = (int)e;
  }
  hipError_t e = hipFreeAsync(stage, st);
  hipError_t s = hipStreamSynchronize(st);
  if (rc) return rc;
  if (e != hipSuccess) return (int)e;
  return s == hipSuccess ? 0 : (int)s;
}

extern "C" int streams_field_interior(int nx, int ny, int nz, int ng, int nv, double *w_gpu, double *dense,
                                      int to_dense, void *stream) {
  if (nx < 1 || ny < 1 || nz < 1 || ng < 0 || nv < 1) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  interior_copy_kernel<<<grid_for(b.fstride() * nv), 256, 0, as_stream(stream)>>>(
      b, nv, to_dense ? w_gpu : dense, to_dense ? dense : w_gpu, to_dense);
  return hip_status();
}
