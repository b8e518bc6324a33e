// common.h — shared device-side definitions for libstreams_hip (gfx950).
//
// Layouts follow the reference's Fortran arrays (column-major, first index fastest):
//   padded field  a(1-ng:nx+ng, 1-ng:ny+ng, 1-ng:nz+ng, 1:nvar)
//   interior field f(1:nx, 1:ny, 1:nz, 1:nv)
//   cp/cv tables   c(indx_cp_l:indx_cp_r+2, N_S, nsetcv), trange(N_S, nsetcv+1)
// N_S is a template parameter; the C-ABI dispatches on the runtime value.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/streams_hip.h"

namespace streams {

// --------------------------------------------------------------------- geometry
struct Box {
  int nx, ny, nz, ng;
  __host__ __device__ inline long px() const { return nx + 2 * ng; }
  __host__ __device__ inline long py() const { return ny + 2 * ng; }
  __host__ __device__ inline long pz() const { return nz + 2 * ng; }
  // padded point (i,j,k) with Fortran indices
  __host__ __device__ inline long p(int i, int j, int k) const {
    return (long)(i + ng - 1) + px() * ((long)(j + ng - 1) + py() * (long)(k + ng - 1));
  }
  __host__ __device__ inline long pstride() const { return px() * py() * pz(); }
  // interior point
  __host__ __device__ inline long f(int i, int j, int k) const {
    return (long)(i - 1) + (long)nx * ((long)(j - 1) + (long)ny * (long)(k - 1));
  }
  __host__ __device__ inline long fstride() const { return (long)nx * ny * nz; }
};

// Hides the value of x from the optimiser, so that values derived from it after this point
// are recomputed rather than kept live (in VGPRs) across the code in between.
template <class T> __device__ inline void opaque(T &x) { asm volatile("" : "+v"(x)); }

// 1/d from the hardware reciprocal estimate refined by two Newton steps (about one ulp,
// not correctly rounded): a fraction of the instructions of an IEEE division, for the
// transport sums that divide by table values dozens of times per point.
__device__ inline double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// index of variables (index_define.h), 1-based as in the reference
template <int NS> struct Idx {
  static constexpr int I_U = NS + 1, I_V = NS + 2, I_W = NS + 3, I_E = NS + 4;
  static constexpr int J_U = NS + 1, J_V = NS + 2, J_W = NS + 3, J_H = NS + 4, J_T = NS + 5;
  static constexpr int J_P = NS + 6, J_C = NS + 7, J_MU = NS + 8, J_DUC = NS + 9, J_DIV = NS + 10;
  static constexpr int J_R = NS + 11, J_Z = NS + 12, J_K_COND = NS + 13, J_D_START = NS + 13;
  // LES eddy viscosity, production rates and heat release (index_define.h:22-26; the PaSR
  // factor shares the heat-release slot)
  static constexpr int J_LES1 = 2 * NS + 14, J_WDOT_START = 2 * NS + 14, J_HRR = 3 * NS + 15;
  static constexpr int NV = NS + 4;
};

// ------------------------------------------------------------------- thermo tables
struct Thermo {
  const double *cv_coeff, *cp_coeff, *trange;
  int nsetcv, lo, hi;  // indx_cp_l, indx_cp_r
  __device__ inline int ncol() const { return hi + 2 - lo + 1; }
};

template <int NS>
__device__ inline double cpc(const double *tab, const Thermo &th, int l, int s, int m) {
  // tab(l, s, m) with l in [lo, hi+2], s,m 1-based
  return tab[(l - th.lo) + th.ncol() * ((s - 1) + NS * (m - 1))];
}

// Numerical-Recipes locate on trange(s,:), clamped to [1, nsetcv]
// (kernels_gpu.F90:6400-6418)
template <int NS>
__device__ inline int locate_range(double tt, const Thermo &th, int s) {
  if (th.nsetcv <= 1) return 1;
  int jl = 0, ju = th.nsetcv + 2;
  while (ju - jl > 1) {
    int jm = (ju + jl) >> 1;
    if (tt >= th.trange[(s - 1) + NS * (jm - 1)]) jl = jm; else ju = jm;
  }
  jl = jl < 1 ? 1 : jl;
  return jl > th.nsetcv ? th.nsetcv : jl;
}

__device__ inline double ipow(double x, int n) {
  // Fortran x**n with an integer exponent
  if (n == 0) return 1.0;
  bool neg = n < 0;
  if (neg) n = -n;
  double r = 1.0;
  for (int q = 0; q < n; ++q) r *= x;
  return neg ? 1.0 / r : r;
}

// Newton inversion e -> T (get_mixture_temperature_from_e_dev, kernels_gpu.F90:6532-6622).
// The mixture sums sum_s cv(l,s,range_s)*Y_s are rebuilt only when a species changes
// temperature range between iterations; with unchanged ranges they are the same numbers
// the reference recomputes every iteration.
constexpr int NCOLMAX = 9;   // NASA9: indx_cp_l = -2 .. indx_cp_r + 2 = 6

template <int NS>
__device__ double temperature_from_e(double ee, double t_start, const Thermo &th,
                                     double tol_iter_nr, const double *yy) {
  int nrange[NS];
#pragma unroll
  for (int s = 1; s <= NS; ++s) nrange[s - 1] = locate_range<NS>(t_start, th, s);
  double cvl[NCOLMAX];
  double sumb = 0.0;
  bool dirty = true;
  const int nl = th.hi - th.lo + 1;
  double t_old = t_start, tt = t_start;
  for (int iter = 1; iter <= 10; ++iter) {
    if (th.nsetcv > 1) {
#pragma unroll
      for (int s = 1; s <= NS; ++s) {
        int nr = nrange[s - 1];
        const int nr0 = nr;
        if (t_old < th.trange[(s - 1) + NS * (nr - 1)]) nr = nr - 1;
        else if (t_old > th.trange[(s - 1) + NS * nr]) nr = nr + 1;
        nr = nr < 1 ? 1 : nr;
        nr = nr > th.nsetcv ? th.nsetcv : nr;
        dirty = dirty || (nr != nr0);
        nrange[s - 1] = nr;
      }
    }
    if (dirty) {
      sumb = 0.0;
#pragma unroll
      for (int s = 1; s <= NS; ++s) sumb = sumb + cpc<NS>(th.cv_coeff, th, th.hi + 1, s, nrange[s - 1]) * yy[s - 1];
      for (int q = 0; q < nl; ++q) {
        double c = 0.0;
#pragma unroll
        for (int s = 1; s <= NS; ++s) c = c + cpc<NS>(th.cv_coeff, th, th.lo + q, s, nrange[s - 1]) * yy[s - 1];
        cvl[q] = c;
      }
      dirty = false;
    }
    const double ebar = ee - sumb;
    double den = 0.0, num = 0.0;
    for (int q = 0; q < nl; ++q) {
      const int l = th.lo + q;
      const double t_pow = ipow(t_old, l);
      const double tnum = (l == -1) ? log(t_old) : (t_old * t_pow) / (l + 1.0);
      num = num + cvl[q] * tnum;
      den = den + cvl[q] * t_pow;
    }
    tt = t_old + (ebar - num) / den;
    if (fabs(tt - t_old) < tol_iter_nr) break;
    t_old = tt;
  }
  return tt;
}

// Copy the thermo tables into a block's LDS buffer (all threads call, then the caller
// synchronises).  Returns a Thermo whose pointers refer to LDS when the tables fit in
// `cap` doubles, the unchanged global view otherwise.
template <int NS>
__device__ Thermo thermo_to_lds(const Thermo &g, double *buf, int cap) {
  const int ncol = g.hi + 2 - g.lo + 1;
  const int ntab = ncol * NS * g.nsetcv;
  const int ntr = NS * (g.nsetcv + 1);
  const bool has_cv = g.cv_coeff != nullptr, has_cp = g.cp_coeff != nullptr;
  const int need = (has_cv ? ntab : 0) + (has_cp ? ntab : 0) + ntr;
  if (need > cap) return g;
  Thermo t = g;
  double *p = buf;
  for (int q = threadIdx.x; q < ntr; q += blockDim.x) p[q] = g.trange[q];
  t.trange = p;
  p += ntr;
  if (has_cv) {
    for (int q = threadIdx.x; q < ntab; q += blockDim.x) p[q] = g.cv_coeff[q];
    t.cv_coeff = p;
    p += ntab;
  }
  if (has_cp) {
    for (int q = threadIdx.x; q < ntab; q += blockDim.x) p[q] = g.cp_coeff[q];
    t.cp_coeff = p;
  }
  return t;
}

template <int NS>
__host__ __device__ constexpr int thermo_lds_doubles() {
  return 2 * NCOLMAX * NS * 3 + NS * 4;   // cp and cv with up to 3 ranges, trange
}

// mixture c_p (get_cp_dev, kernels_gpu.F90:6384-6432)
template <int NS>
__device__ double mixture_cp(double tt, const Thermo &th, const double *yy) {
  int nrange[NS];
#pragma unroll
  for (int s = 1; s <= NS; ++s) nrange[s - 1] = locate_range<NS>(tt, th, s);
  double cploc = 0.0;
  for (int l = th.lo; l <= th.hi; ++l) {
    double tpow = ipow(tt, l);
    double cp_l = 0.0;
#pragma unroll
    for (int s = 1; s <= NS; ++s) cp_l = cp_l + cpc<NS>(th.cp_coeff, th, l, s, nrange[s - 1]) * yy[s - 1];
    cploc = cploc + cp_l * tpow;
  }
  return cploc;
}

// species enthalpy from cp_coeff (get_species_h_from_temperature_dev, :7144-7182); with
// the cv table the same expression is the species energy (:7104-7142)
template <int NS>
__device__ double species_hcoef(double tt, const double *tab, const Thermo &th, int s) {
  int nr = locate_range<NS>(tt, th, s);
  double hh = cpc<NS>(tab, th, th.hi + 1, s, nr);
  for (int l = th.lo; l <= th.hi; ++l) {
    if (l == -1) hh = hh + cpc<NS>(tab, th, l, s, nr) * log(tt);
    else hh = hh + cpc<NS>(tab, th, l, s, nr) / (l + 1.0) * ipow(tt, l + 1);
  }
  return hh;
}

// ------------------------------------------------------------ XCD-aware block order
// The MI355X dispatches consecutive workgroups of a launch round-robin to its 8 XCDs, each
// with its own L2.  Tiles that share cache lines (halo rows, the misaligned ends of a
// padded row) reuse them only when they run on the same XCD at about the same time, so a
// launch over nblk tiles uses a grid of xcd_grid(nblk) workgroups and maps its workgroup
// to the tile xcd_tile(): XCD x takes the x-th contiguous range of tiles, in order.
constexpr unsigned NXCD = 8;

__host__ __device__ inline unsigned xcd_grid(unsigned nblk) { return (nblk + NXCD - 1) / NXCD * NXCD; }

__device__ inline unsigned xcd_tile(unsigned b, unsigned grid) {
  return (b % NXCD) * (grid / NXCD) + b / NXCD;   // a permutation of 0..grid-1
}

// Row order for pointwise kernels with a +-l stencil in y and z: one workgroup per
// 256-point chunk of an i-row, the rows visited in bands of SB j-rows with k fastest within
// a band, and the bands split into contiguous ranges per XCD.  The rows an XCD has in
// flight then form a compact (j, k) block whose y and z neighbours sit in that XCD's L2.
struct StencilRows {
  static constexpr int SB = 8;                       // j-rows per band
  int nib, nbands;                                   // i-chunks per row, bands
  __host__ __device__ inline unsigned tiles(int nz) const { return (unsigned)nib * SB * nz * nbands; }
};
inline StencilRows stencil_rows(const Box &b) {
  StencilRows r;
  r.nib = (b.nx + 255) / 256;
  r.nbands = (b.ny + StencilRows::SB - 1) / StencilRows::SB;
  return r;
}
// Interior point (i, j, k) of this thread; false for padding tiles and points past nx / ny.
__device__ inline bool stencil_point(const Box &b, const StencilRows &r, int &i, int &j, int &k) {
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  if (tile >= r.tiles(b.nz)) return false;
  const unsigned jj = tile % StencilRows::SB;
  const unsigned q = tile / StencilRows::SB;
  const unsigned ib = q % r.nib;
  const unsigned q2 = q / r.nib;
  k = 1 + (int)(q2 % b.nz);
  j = 1 + (int)((q2 / b.nz) * StencilRows::SB + jj);
  i = 1 + (int)(ib * 256 + threadIdx.x);
  return i <= b.nx && j <= b.ny;
}

// ------------------------------------------------------------------ launch helpers
inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline int hip_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace streams

// N_S dispatch: the compiled species counts.  Each kernel family instantiates its
// templates for every entry of this list.
#define STREAMS_NS_LIST(X) X(1) X(2) X(3) X(4) X(5) X(9) X(11) X(13) X(14)

#define STREAMS_DISPATCH_NS(ns, CALL)                         \
  switch (ns) {                                               \
    case 1: { constexpr int NS = 1; CALL; } break;            \
    case 2: { constexpr int NS = 2; CALL; } break;            \
    case 3: { constexpr int NS = 3; CALL; } break;            \
    case 4: { constexpr int NS = 4; CALL; } break;            \
    case 5: { constexpr int NS = 5; CALL; } break;            \
    case 9: { constexpr int NS = 9; CALL; } break;            \
    case 11: { constexpr int NS = 11; CALL; } break;          \
    case 13: { constexpr int NS = 13; CALL; } break;          \
    case 14: { constexpr int NS = 14; CALL; } break;          \
    default: return STREAMS_E_NSPECIES;                       \
  }
