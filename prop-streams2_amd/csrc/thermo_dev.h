// thermo_dev.h — NASA-polynomial thermodynamics read from a per-workgroup LDS copy of the
// tables, with the powers of T formed once per temperature.
//
// The reference evaluates, per species s with its temperature range m
// (get_cp_dev, get_mixture_temperature_from_e_dev, get_species_h_from_temperature_dev;
// kernels_gpu.F90:6384-6432, :6532-6622, :7144-7182):
//   c_p = sum_{l=lo..hi} a_l T^l
//   h   = a_{hi+1} + sum_{l=lo..hi} a_l T^(l+1)/(l+1)   (a_{-1} ln T for l = -1)
// with integer powers by repeated multiplication inside every term.  Here the tables are
// copied into LDS once per workgroup together with a_l/(l+1), the powers T^l are formed
// once per temperature by the same repeated multiplication, and each species term is one
// multiply-add.  The polynomial form is a template parameter: LO = 0 for NASA7 (l = 0..4),
// LO = -2 for NASA9 (l = -2..4); hi = 4 in both (multideal.F90:3017-3026).
#pragma once
#include "common.h"

namespace streams {

constexpr int TSTRIDE = 8;   // doubles per (range, species) row: NL terms + the constant
constexpr int TMAXSET = 4;   // largest number of temperature ranges staged in LDS

template <int LO>
struct Nasa {
  static constexpr int HI = 4;
  static constexpr int NL = HI - LO + 1;          // polynomial terms l = LO..HI
  static constexpr int NP = HI + 2 - LO;          // powers T^l, l = LO..HI+1
};

// Which tables a kernel stages; each is [range][species][TSTRIDE].
enum : int { TH_CP = 1, TH_CPH = 2, TH_CV = 4, TH_CVH = 8 };

template <int NS>
__host__ __device__ constexpr int thermo_lds_doubles(int which) {
  return ((which & TH_CP) ? 1 : 0) * TMAXSET * NS * TSTRIDE + ((which & TH_CPH) ? 1 : 0) * TMAXSET * NS * TSTRIDE +
         ((which & TH_CV) ? 1 : 0) * TMAXSET * NS * TSTRIDE + ((which & TH_CVH) ? 1 : 0) * TMAXSET * NS * TSTRIDE +
         NS * (TMAXSET + 1);
}

struct ThermoLds {
  const double *cp, *cph, *cv, *cvh;  // LDS rows, nullptr when not staged
  const double *tr;                   // trange, [s * (TMAXSET + 1) + (m - 1)]
  int nsetcv;
};

// Copy the requested tables of g (global, Fortran layouts) into buf; all threads of the
// block call this, the caller synchronises before use.
template <int NS, int LO>
__device__ ThermoLds thermo_stage(const Thermo &g, double *buf, int which) {
  using N = Nasa<LO>;
  ThermoLds t{nullptr, nullptr, nullptr, nullptr, nullptr, g.nsetcv};
  const int ncol = g.hi + 2 - g.lo + 1;          // rows of the Fortran table: lo..hi+2
  const int rows = g.nsetcv * NS;
  double *p = buf;
  auto stage = [&](const double *src, bool divide) {
    for (int q = threadIdx.x; q < rows * TSTRIDE; q += blockDim.x) {
      const int row = q / TSTRIDE, col = q % TSTRIDE;
      const int m = row / NS, s = row % NS;
      double v = 0.0;
      if (col < N::NL) {
        const int l = LO + col;
        v = src[col + ncol * (s + NS * m)];
        if (divide && l != -1) v = v / (l + 1.0);
      } else if (col == N::NL) {
        v = src[N::NL + ncol * (s + NS * m)];      // a_{hi+1}
      }
      p[q] = v;
    }
    const double *out = p;
    p += TMAXSET * NS * TSTRIDE;
    return out;
  };
  if (which & TH_CP) t.cp = stage(g.cp_coeff, false);
  if (which & TH_CPH) t.cph = stage(g.cp_coeff, true);
  if (which & TH_CV) t.cv = stage(g.cv_coeff, false);
  if (which & TH_CVH) t.cvh = stage(g.cv_coeff, true);
  for (int q = threadIdx.x; q < NS * (TMAXSET + 1); q += blockDim.x) {
    const int s = q / (TMAXSET + 1), m = q % (TMAXSET + 1);
    p[q] = m <= g.nsetcv ? g.trange[s + NS * m] : 0.0;
  }
  t.tr = p;
  return t;
}

// Range of species s (0-based) at temperature tt: the Numerical-Recipes locate of
// kernels_gpu.F90:6400-6418 on an ascending trange, clamped to [1, nsetcv], as a scan.
template <int NS>
__device__ inline int range_of(const ThermoLds &t, int s, double tt) {
  int nr = 1;
  const double *tr = t.tr + s * (TMAXSET + 1);
  for (int m = 2; m <= t.nsetcv; ++m) nr = tt >= tr[m - 1] ? m : nr;
  return nr;
}

// T^l for l = LO..5 by repeated multiplication (negative powers as 1/T^|l|) and ln T.
template <int LO>
struct TPow {
  double p[Nasa<LO>::NP];
  double lnT;
  __device__ inline explicit TPow(double tt) {
    constexpr int z = -LO;
    p[z] = 1.0;
#pragma unroll
    for (int l = 1; l <= 5; ++l) p[z + l] = p[z + l - 1] * tt;
#pragma unroll
    for (int l = -1; l >= LO; --l) p[z + l] = 1.0 / p[z - l];
    lnT = LO <= -1 ? log(tt) : 0.0;
  }
  __device__ inline double operator[](int l) const { return p[l - LO]; }
};

// species enthalpy (cph rows) or energy (cvh rows): a_{hi+1} + sum_l a_l/(l+1) T^(l+1)
template <int NS, int LO>
__device__ inline double species_h(const double *rows, int s, int nr, const TPow<LO> &P) {
  using N = Nasa<LO>;
  const double *c = rows + ((nr - 1) * NS + s) * TSTRIDE;
  double hh = c[N::NL];
#pragma unroll
  for (int q = 0; q < N::NL; ++q) {
    const int l = LO + q;
    hh = hh + c[q] * (l == -1 ? P.lnT : P[l + 1]);
  }
  return hh;
}

// mixture c_p = sum_l (sum_s a_l(s) Y_s) T^l
template <int NS, int LO>
__device__ inline double mixture_cp_lds(const ThermoLds &t, const int *nrange, const double *yy, const TPow<LO> &P) {
  using N = Nasa<LO>;
  double cp_l[N::NL];
#pragma unroll
  for (int q = 0; q < N::NL; ++q) cp_l[q] = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {          // species by species, fenced as in the Newton rebuild
    const int row = ((nrange[s] - 1) * NS + s) * TSTRIDE;
#pragma unroll
    for (int q = 0; q < N::NL; ++q) cp_l[q] = cp_l[q] + t.cp[row + q] * yy[s];
    __builtin_amdgcn_sched_barrier(0);
  }
  double cploc = 0.0;
#pragma unroll
  for (int q = 0; q < N::NL; ++q) cploc = cploc + cp_l[q] * P[LO + q];
  return cploc;
}

// Newton inversion e -> T (get_mixture_temperature_from_e_dev, kernels_gpu.F90:6532-6622):
// per-species range tracking by one step per iteration from the locate at t_start, at
// most 10 iterations, stop at |dT| < tol.  The mixture sums of the coefficients are rebuilt
// only when a range changes.  nrange returns the ranges of the last iteration.
// The same Newton iteration on any pair of staged tables: (cv, cv/(l+1)) inverts the
// internal energy, (cp, cp/(l+1)) the enthalpy.
template <int NS, int LO>
__device__ inline double temperature_newton_lds(double ee, double t_start, const ThermoLds &t, const double *c_rows,
                                                const double *ch_rows, double tol_iter_nr, const double *yy,
                                                int *nrange);

template <int NS, int LO>
__device__ inline double temperature_from_e_lds(double ee, double t_start, const ThermoLds &t, double tol_iter_nr,
                                                const double *yy, int *nrange) {
  return temperature_newton_lds<NS, LO>(ee, t_start, t, t.cv, t.cvh, tol_iter_nr, yy, nrange);
}

template <int NS, int LO>
__device__ inline double temperature_newton_lds(double ee, double t_start, const ThermoLds &t, const double *c_rows,
                                                const double *ch_rows, double tol_iter_nr, const double *yy,
                                                int *nrange) {
  using N = Nasa<LO>;
#pragma unroll
  for (int s = 0; s < NS; ++s) nrange[s] = range_of<NS>(t, s, t_start);
  double cvl[N::NL], cvh[N::NL];
  double sumb = 0.0;
  bool dirty = true;
  double t_old = t_start, tt = t_start;
  for (int iter = 1; iter <= 10; ++iter) {
    if (t.nsetcv > 1) {
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const double *tr = t.tr + s * (TMAXSET + 1);
        int nr = nrange[s];
        const int nr0 = nr;
        if (t_old < tr[nr - 1]) nr = nr - 1;
        else if (t_old > tr[nr]) nr = nr + 1;
        nr = nr < 1 ? 1 : nr;
        nr = nr > t.nsetcv ? t.nsetcv : nr;
        dirty = dirty || (nr != nr0);
        nrange[s] = nr;
      }
    }
    if (dirty) {
      // species by species (the same sums in the same order as per coefficient), with a
      // scheduling fence after each: left free, the scheduler issues all 11*NS LDS reads
      // at once and the kernel's register count doubles
      sumb = 0.0;
#pragma unroll
      for (int q = 0; q < N::NL; ++q) cvl[q] = cvh[q] = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int row = ((nrange[s] - 1) * NS + s) * TSTRIDE;
        sumb = sumb + c_rows[row + N::NL] * yy[s];
#pragma unroll
        for (int q = 0; q < N::NL; ++q) {
          cvl[q] = cvl[q] + c_rows[row + q] * yy[s];
          cvh[q] = cvh[q] + ch_rows[row + q] * yy[s];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      dirty = false;
    }
    const TPow<LO> P(t_old);
    const double ebar = ee - sumb;
    double den = 0.0, num = 0.0;
#pragma unroll
    for (int q = 0; q < N::NL; ++q) {
      const int l = LO + q;
      num = num + cvh[q] * (l == -1 ? P.lnT : P[l + 1]);
      den = den + cvl[q] * P[l];
    }
    tt = t_old + (ebar - num) / den;
    if (fabs(tt - t_old) < tol_iter_nr) break;
    t_old = tt;
  }
  return tt;
}

}  // namespace streams

// Dispatch on the polynomial form (indx_cp_l, indx_cp_r): NASA7 (0, 4) or NASA9 (-2, 4).
#define STREAMS_DISPATCH_NASA(lo, hi, CALL)                                   \
  if ((hi) != 4) return STREAMS_E_OPTION;                                     \
  switch (lo) {                                                               \
    case 0: { constexpr int LO = 0; CALL; } break;                            \
    case -2: { constexpr int LO = -2; CALL; } break;                          \
    default: return STREAMS_E_OPTION;                                         \
  }
