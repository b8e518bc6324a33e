// noalloc.hip — compute_dt without any device allocation, for callers that cannot allocate
// (code under stream capture, serving loops, harnesses that forbid hipMalloc).
//
// The partial maxima live in a fixed device array of this library, `scratch`: 9 doubles for
// each of at most kDtMaxBlocks blocks plus the 9 results (147,528 bytes per GPU, held until
// the process ends).  The host never uses the array's name as an address: on the first call
// on a device, with that device current, hipGetSymbolAddress gives that device's copy of
// the array, and that pointer is kept per device and passed to the kernels.  Every call on
// one GPU shares its array, so calls that overlap on one GPU (other streams, other host
// threads) race; each call rewrites every partial result it reads, so nothing needs
// clearing between calls.
#include <mutex>

#include "common.h"
#include "dt_reduce.h"

namespace streams {

constexpr size_t kDtScratchDoubles = 9 * (size_t)kDtMaxBlocks + 9;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kDtScratchDoubles];

static std::mutex g_scratch_mutex;
static double *g_scratch_ptr[kMaxDevices] = {};

// This device's address of `scratch`, looked up once per device (the current one).
static int scratch_address(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return STREAMS_E_DIMS;
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace streams

using namespace streams;

extern "C" int streams_compute_dt_noalloc(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                                          const double *dcsidx, const double *detady, const double *dzitdz,
                                          const double *dcsidxs, const double *detadys, const double *dzitdzs,
                                          const double *w, const double *w_aux, double *dtmax9, int indx_cp_l,
                                          int indx_cp_r, const double *cp_coeff, int nsetcv, const double *trange,
                                          const int *fluid_mask, void *stream) {
  (void)w; (void)nv;
  if (nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  if (nx < 0 || ny < 0 || nz < 0 || ng < 0) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  // the grid is capped at kDtMaxBlocks blocks (any box is covered by the grid-stride loop),
  // so the partial results always fit; a launch that needed more would be refused here
  if (9 * (size_t)dt_blocks(b.fstride()) + 9 > kDtScratchDoubles) return STREAMS_E_DIMS;
  double *partial = nullptr;
  int rc = scratch_address(&partial);
  if (rc) return rc;
  Thermo th{nullptr, cp_coeff, trange, nsetcv, indx_cp_l, indx_cp_r};
  return compute_dt_into(n_s, b, dcsidx, detady, dzitdz, dcsidxs, detadys, dzitdzs, w_aux, dtmax9, th, fluid_mask,
                         partial, as_stream(stream));
}
