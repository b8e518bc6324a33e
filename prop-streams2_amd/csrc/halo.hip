// halo.hip — ghost-cell fills: single-rank loopback exchange and physical boundaries.
//
// The reference exchanges ng-deep face slabs through pack -> MPI (or a device copy when
// the neighbour is the rank itself) -> unpack (base_gpu.F90:1132-1264).  On one GPU the
// periodic exchange is a direct copy from the opposite interior slab into the ghost slab,
// one kernel per axis with the two sides and all variables in one grid.  When an axis is
// shorter than ng the source slab overlaps the ghost slab of the same axis; that case is
// staged through a device buffer so that every source is read before any ghost is written.

#include "common.h"

namespace streams {

// axis 0: (i in 1..ng, j in 1..ny, k in 1..nz); axis 1: (i 1..nx, j 1..ng, k 1..nz);
// axis 2: (i 1..nx, j 1..ny, k 1..ng).  The flat index covers side, variable, slab.
struct Slab {
  int ax, nv;
  int e0, e1, e2;  // extents of the slab in i, j, k
  __device__ inline void decode(long t, int &side, int &m, int &a, int &bq, int &c) const {
    a = 1 + (int)(t % e0); t /= e0;
    bq = 1 + (int)(t % e1); t /= e1;
    c = 1 + (int)(t % e2); t /= e2;
    m = (int)(t % nv); t /= nv;
    side = (int)t;
  }
};

__device__ inline void slab_src_dst(const Box &b, int ax, int side, int a, int bq, int c, long &src,
                                    long &dst) {
  // side 0: left ghost <- right interior slab; side 1: right ghost <- left interior slab
  int i = a, j = bq, k = c;
  if (ax == 0) {
    if (side == 0) { src = b.p(b.nx - b.ng + i, j, k); dst = b.p(i - b.ng, j, k); }
    else { src = b.p(i, j, k); dst = b.p(b.nx + i, j, k); }
  } else if (ax == 1) {
    if (side == 0) { src = b.p(i, b.ny - b.ng + j, k); dst = b.p(i, j - b.ng, k); }
    else { src = b.p(i, j, k); dst = b.p(i, b.ny + j, k); }
  } else {
    if (side == 0) { src = b.p(i, j, b.nz - b.ng + k); dst = b.p(i, j, k - b.ng); }
    else { src = b.p(i, j, k); dst = b.p(i, j, b.nz + k); }
  }
}

__global__ void __launch_bounds__(256) swap_direct_kernel(Box b, Slab s, long n, double *__restrict__ w) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int side, m, a, bq, c;
  s.decode(t, side, m, a, bq, c);
  long src, dst;
  slab_src_dst(b, s.ax, side, a, bq, c, src, dst);
  const long ps = b.pstride() * m;
  w[dst + ps] = w[src + ps];
}

__global__ void __launch_bounds__(256) swap_pack_kernel(Box b, Slab s, long n, const double *__restrict__ w,
                                                        double *__restrict__ buf) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int side, m, a, bq, c;
  s.decode(t, side, m, a, bq, c);
  long src, dst;
  slab_src_dst(b, s.ax, side, a, bq, c, src, dst);
  buf[t] = w[src + b.pstride() * m];
}

__global__ void __launch_bounds__(256) swap_unpack_kernel(Box b, Slab s, long n, double *__restrict__ w,
                                                          const double *__restrict__ buf) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int side, m, a, bq, c;
  s.decode(t, side, m, a, bq, c);
  long src, dst;
  slab_src_dst(b, s.ax, side, a, bq, c, src, dst);
  w[dst + b.pstride() * m] = buf[t];
}

// bcextr_var_cuf / bcextr_cuf: zero-order extrapolation of face ilat (1..6)
__global__ void __launch_bounds__(256) extr_kernel(Box b, int ilat, int nv, long n, double *__restrict__ w) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int l, a, c, m;
  // flat index: l fastest, then the two face coordinates, then the variable
  l = 1 + (int)(t % b.ng); long r = t / b.ng;
  int e0 = (ilat <= 2) ? b.ny : b.nx;
  int e1 = (ilat <= 4) ? b.nz : b.ny;
  a = 1 + (int)(r % e0); r /= e0;
  c = 1 + (int)(r % e1); r /= e1;
  m = (int)r;
  long src, dst;
  switch (ilat) {
    case 1: src = b.p(1, a, c); dst = b.p(1 - l, a, c); break;
    case 2: src = b.p(b.nx, a, c); dst = b.p(b.nx + l, a, c); break;
    case 3: src = b.p(a, 1, c); dst = b.p(a, 1 - l, c); break;
    case 4: src = b.p(a, b.ny, c); dst = b.p(a, b.ny + l, c); break;
    case 5: src = b.p(a, c, 1); dst = b.p(a, c, 1 - l); break;
    default: src = b.p(a, c, b.nz); dst = b.p(a, c, b.nz + l); break;
  }
  const long ps = b.pstride() * m;
  w[dst + ps] = w[src + ps];
}

// bcwall_ad_cuf: mirror with the three momenta reversed
__global__ void __launch_bounds__(256) wall_ad_kernel(Box b, int ilat, int ns, long n, double *__restrict__ w) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int nv = ns + 4;
  int l = 1 + (int)(t % b.ng); long r = t / b.ng;
  int e0 = (ilat <= 2) ? b.ny : b.nx;
  int e1 = (ilat <= 4) ? b.nz : b.ny;
  int a = 1 + (int)(r % e0); r /= e0;
  int c = 1 + (int)(r % e1); r /= e1;
  int m = (int)r;  // 0-based variable
  long src, dst;
  switch (ilat) {
    case 1: src = b.p(l, a, c); dst = b.p(1 - l, a, c); break;
    case 2: src = b.p(b.nx + 1 - l, a, c); dst = b.p(b.nx + l, a, c); break;
    case 3: src = b.p(a, l, c); dst = b.p(a, 1 - l, c); break;
    case 4: src = b.p(a, b.ny + 1 - l, c); dst = b.p(a, b.ny + l, c); break;
    case 5: src = b.p(a, c, l); dst = b.p(a, c, 1 - l); break;
    default: src = b.p(a, c, b.nz + 1 - l); dst = b.p(a, c, b.nz + l); break;
  }
  const long ps = b.pstride() * m;
  double v = w[src + ps];
  w[dst + ps] = (m >= ns && m < ns + 3) ? -v : v;
  (void)nv;
}

int extr_faces(const Box &b, int ilat, int nv, double *w, hipStream_t st) {
  int e0 = (ilat <= 2) ? b.ny : b.nx;
  int e1 = (ilat <= 4) ? b.nz : b.ny;
  long n = (long)b.ng * e0 * e1 * nv;
  if (n <= 0) return 0;
  extr_kernel<<<blocks_for(n, 256), 256, 0, st>>>(b, ilat, nv, n, w);
  return hip_status();
}

int swap_axis(const Box &b, int ax, int nv, double *w, hipStream_t st) {
  int n_ax = ax == 0 ? b.nx : ax == 1 ? b.ny : b.nz;
  Slab s{ax, nv, ax == 0 ? b.ng : b.nx, ax == 1 ? b.ng : b.ny, ax == 2 ? b.ng : b.nz};
  long n = 2L * s.e0 * s.e1 * s.e2 * nv;
  if (n <= 0) return 0;
  if (n_ax >= b.ng) {
    swap_direct_kernel<<<blocks_for(n, 256), 256, 0, st>>>(b, s, n, w);
  } else {
    // stream-ordered staging: calls on different streams each get their own buffer
    double *buf = nullptr;
    if (hipMallocAsync((void **)&buf, (size_t)n * sizeof(double), st) != hipSuccess) return STREAMS_E_NOMEM;
    swap_pack_kernel<<<blocks_for(n, 256), 256, 0, st>>>(b, s, n, w, buf);
    swap_unpack_kernel<<<blocks_for(n, 256), 256, 0, st>>>(b, s, n, w, buf);
    const int rc = hip_status();
    const hipError_t e = hipFreeAsync(buf, st);
    if (rc) return rc;
    return e == hipSuccess ? 0 : (int)e;
  }
  return hip_status();
}

}  // namespace streams

using namespace streams;

extern "C" int streams_bcswap(int nx, int ny, int nz, int ng, int nv, double *w, const int *periodic,
                              void *stream) {
  if (nx < 1 || ny < 1 || nz < 1 || ng < 1 || nv < 1 || !periodic) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  hipStream_t st = as_stream(stream);
  // The reference packs all six slabs before unpacking any.  Sources of one axis never lie
  // in the ghost slabs of another axis (their tangential indices are interior), so the
  // axes are independent; only the short-axis self-overlap needs staging.
  for (int ax = 0; ax < 3; ++ax) {
    if (!periodic[ax]) continue;
    int rc = swap_axis(b, ax, nv, w, st);
    if (rc) return rc;
  }
  return 0;
}

extern "C" int streams_bcextr_var(int nx, int ny, int nz, int ng, double *w_var, void *stream) {
  Box b{nx, ny, nz, ng};
  for (int ilat = 1; ilat <= 6; ++ilat) {
    int rc = extr_faces(b, ilat, 1, w_var, as_stream(stream));
    if (rc) return rc;
  }
  return 0;
}

extern "C" int streams_bcextr(int ilat, int nx, int ny, int nz, int ng, int nv, double *w, void *stream) {
  if (ilat < 1 || ilat > 6) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  return extr_faces(b, ilat, nv, w, as_stream(stream));
}

extern "C" int streams_bcwall_ad(int n_s, int ilat, int nx, int ny, int nz, int ng, double *w, void *stream) {
  if (ilat < 1 || ilat > 6 || n_s < 1) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  int e0 = (ilat <= 2) ? ny : nx;
  int e1 = (ilat <= 4) ? nz : ny;
  long n = (long)ng * e0 * e1 * (n_s + 4);
  if (n <= 0) return 0;
  wall_ad_kernel<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(b, ilat, n_s, n, w);
  return hip_status();
}

extern "C" int streams_species_supported(int *list, int cap) {
  static const int compiled[] = {
#define STREAMS_LIST_ENTRY(n) n,
      STREAMS_NS_LIST(STREAMS_LIST_ENTRY)
#undef STREAMS_LIST_ENTRY
  };
  int n = (int)(sizeof(compiled) / sizeof(compiled[0]));
  for (int q = 0; q < n && q < cap; ++q) list[q] = compiled[q];
  return n;
}
