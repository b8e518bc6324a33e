// kinpack.h — compact per-reaction records of a mechanism's kinetics.
//
// The reference keeps the kinetics as dense tables indexed (reaction, species)
// (arr_a/b/ea(nr,2), falloff_coeffs(nr,5), tb_eff/r_coeffs/p_coeffs(nr,N_S),
// reac_ty/isRev(nr); multideal.F90:2763-2926).  Every cell walks the same reactions in the
// same order, so a reaction's data is wave-uniform: the fused chemistry reads one record
// per reaction through the constant address space, which the compiler turns into scalar
// (SMEM) loads into SGPRs, and touches only the species that take part in the reaction.
//
// One record is KREC doubles (640 bytes):
//   the first 16 doubles hold 32 int32:
//     [0] reac_ty   [1] isRev   [2] n_reac   [3] n_prod
//     [5] 1 when b or Ea of the (low-pressure) rate is nonzero
//     [6] 1 when the Troe T2 term is present (falloff(4) != -3.14)
//     [8..11]  reactant species (0-based), [12..15] product species
//   doubles:
//     [16..18] A, b, Ea of the rate (the low-pressure limit of a fall-off reaction)
//     [19..21] A, b, Ea of the high-pressure limit
//     [22..26] falloff_coeffs(1..5)
//     [28..31] reactant exponents, [32..35] product exponents
//     [36..51] (p_s - r_s) * W_s of species 1..N_S (zero for species the reaction leaves
//              unchanged: subtracting q*0 leaves the accumulator as it is)
//     [52..67] third-body efficiencies of species 1..N_S
#pragma once
#include "common.h"

namespace streams {

constexpr int KREC = 80, KSIDE = 4, KTB = 16;
constexpr int KI_TY = 0, KI_REV = 1, KI_NRE = 2, KI_NPR = 3, KI_ARR = 5, KI_T2 = 6;
constexpr int KI_RE = 8, KI_PR = 12;
constexpr int KD_A = 16, KD_AINF = 19, KD_FC = 22, KD_REX = 28, KD_PEX = 32, KD_NET = 36, KD_TB = 52;

typedef __attribute__((address_space(4))) const double kdouble;
typedef __attribute__((address_space(4))) const int kint;

// The tables the fused chemistry needs beside the packed records.
struct KinPacked {
  const double *pack;     // nr records of KREC doubles
  const double *kc_tab;   // kc_tab(num_t_tab+1, nr)
  int nr, num_t_tab;
  double t_min_tab, dt_tab, r_univ;
};

// c**ex for a stoichiometric exponent.  Exponents other than 1, 2 and 3 go through
// exp(ex ln c) rather than pow: the same value to a few ulp for c > 0 and the same special
// values for c = 0 (0 or inf) and c < 0 (NaN for a fractional exponent), with a fraction of
// pow's register footprint, which the whole chemistry loop would otherwise carry.
__device__ inline double conc_power(double c, double ex) {
  if (ex == 1.0) return c;
  if (ex == 2.0) return c * c;
  if (ex == 3.0) return c * c * c;
  return exp(ex * log(c));
}

// The record's (reactant, product) exponents of species s (0 when s does not take part):
// the values of r_coeffs(lr,s) and p_coeffs(lr,s).
template <int S>
__device__ inline void record_coeffs(kdouble *R, kint *I, int nre, int npr, double &rr, double &pp) {
  rr = 0.0;
  pp = 0.0;
#pragma unroll
  for (int e = 0; e < KSIDE; ++e) {
    if (e < nre && I[KI_RE + e] == S) rr = R[KD_REX + e];
    if (e < npr && I[KI_PR + e] == S) pp = R[KD_PEX + e];
  }
}

template <int S, int NS>
struct RefSpecies {
  // q1 = prod c^r, q2 = prod c^p in ascending species order with the reference's
  // exponent cases (1: multiply, other non-zero: pow)
  __device__ static inline void products(kdouble *R, kint *I, int nre, int npr, const double *conc, double &q1,
                                         double &q2) {
    double rr, pp;
    record_coeffs<S>(R, I, nre, npr, rr, pp);
    if (rr == 1.0) q1 = q1 * conc[S];
    else if (rr != 0.0) q1 = q1 * pow(conc[S], rr);
    if (pp == 1.0) q2 = q2 * conc[S];
    else if (pp != 0.0) q2 = q2 * pow(conc[S], pp);
    RefSpecies<S + 1, NS>::products(R, I, nre, npr, conc, q1, q2);
  }
  // acc_s -= ((p - r) q) W_s for the species of the reaction
  __device__ static inline void accumulate(kdouble *R, kint *I, int nre, int npr, double q, const double *mw,
                                           double *acc) {
    double rr, pp;
    record_coeffs<S>(R, I, nre, npr, rr, pp);
    if (rr != 0.0 || pp != 0.0) acc[S] = acc[S] - ((pp - rr) * q) * mw[S];
    RefSpecies<S + 1, NS>::accumulate(R, I, nre, npr, q, mw, acc);
  }
};
template <int NS>
struct RefSpecies<NS, NS> {
  __device__ static inline void products(kdouble *, kint *, int, int, const double *, double &, double &) {}
  __device__ static inline void accumulate(kdouble *, kint *, int, int, double, const double *, double *) {}
};

// Species source -omega_s W_s accumulated into acc[0..NS) for one cell at temperature tt
// with molar concentrations conc (compute_chemistry_cuf's reaction loop,
// kernels_gpu.F90:9551-9649).
//  * REF = false (the fused RHS and the fast reactor paths): rate constants through
//    A exp(b ln T - Ea/RT), log10(x) = ln(x)/ln(10), 10^x = exp(x ln 10), c^2 = c*c, and the
//    net coefficient (p - r) W_s taken from the record; these identities change results in
//    the last digits only.
//  * REF = true (the plain FP64 reactor integrator, the yardstick of the others): the
//    reference's own forms and operation order (rosenbrock_rhs, :9249-9339): A T^b exp(-Ea/RT)
//    with the b = 0 and Ea = 0 cases skipped, log10, pow(10, .), pow for exponents other
//    than 1, products over the species in ascending order, acc -= ((p - r) q) W_s with
//    mw = W_s.
template <int NS, bool REF = false>
__device__ inline void chemistry_source_packed(const KinPacked &K, double tt, const double *conc, double *acc,
                                               const double *mw = nullptr) {
  static_assert(NS <= KTB, "third-body efficiencies of a record hold at most KTB species");
  constexpr double LN10 = 2.302585092994045684;
  constexpr double ILN10 = 0.434294481903251827651;
  const double Rtti = 1.0 / (K.r_univ * tt);
  const double lnT = log(tt);
  int itt = (int)((tt - K.t_min_tab) / K.dt_tab) + 1;
  itt = itt < 1 ? 1 : itt;
  itt = itt > K.num_t_tab ? K.num_t_tab : itt;
  const double ttleft = (itt - 1) * K.dt_tab + K.t_min_tab;
  const double dtt = (tt - ttleft) / K.dt_tab;
  const long ntab = (long)K.num_t_tab + 1;
  for (int lr = 0; lr < K.nr; ++lr) {
    kdouble *R = (kdouble *)K.pack + KREC * lr;
    kint *I = (kint *)R;
    const int ty = I[KI_TY];
    double kf, tbv = 1.0;
    if (REF) {
      if (ty <= 1) {
        kf = R[KD_A];
        if (R[KD_A + 1] != 0.0) kf = kf * pow(tt, R[KD_A + 1]);
        if (R[KD_A + 2] != 0.0) kf = kf * exp(-R[KD_A + 2] * Rtti);
        if (ty == 1) {
          tbv = 0.0;
#pragma unroll
          for (int s = 0; s < NS; ++s) tbv = tbv + R[KD_TB + s] * conc[s];
        }
      } else {
        const double k0 = R[KD_A] * pow(tt, R[KD_A + 1]) * exp(-R[KD_A + 2] * Rtti);
        const double kinf = R[KD_AINF] * pow(tt, R[KD_AINF + 1]) * exp(-R[KD_AINF + 2] * Rtti);
        double tb = 0.0;
#pragma unroll
        for (int s = 0; s < NS; ++s) tb = tb + R[KD_TB + s] * conc[s];
        const double rpres = k0 * tb / kinf;
        if (ty == 2) {
          kf = kinf * (rpres / (1.0 + rpres));
        } else if (ty == 3) {
          if (rpres < 1e-30) {
            kf = k0 * tb;
          } else {
            double fcent = (1.0 - R[KD_FC]) * exp(-tt / R[KD_FC + 1]) + R[KD_FC] * exp(-tt / R[KD_FC + 2]);
            if (I[KI_T2]) fcent = fcent + exp(-R[KD_FC + 3] / tt);
            fcent = log10(fcent);
            const double coff = -0.4 - 0.67 * fcent;
            const double noff = 0.75 - 1.27 * fcent;
            const double doff = 0.14;
            const double lr10 = log10(rpres);
            const double fo = (lr10 + coff) / (noff - doff * (lr10 + coff));
            const double foff = pow(10.0, pow(1.0 + pow(fo, 2.0), -1.0) * fcent);
            kf = kinf * (rpres / (1.0 + rpres)) * foff;
          }
        } else {
          const double xoff = 1.0 / (1.0 + pow(log10(rpres), 2.0));
          const double foff = R[KD_FC + 3] * pow(R[KD_FC] * exp(-R[KD_FC + 1] / tt) + exp(-tt / R[KD_FC + 2]), xoff) *
                              pow(tt, R[KD_FC + 4]);
          kf = kinf * (rpres / (1.0 + rpres)) * foff;
        }
      }
    } else if (ty <= 1) {
      kf = R[KD_A];
      if (I[KI_ARR]) kf = kf * exp(R[KD_A + 1] * lnT - R[KD_A + 2] * Rtti);
      if (ty == 1) {
        tbv = 0.0;
#pragma unroll
        for (int s = 0; s < NS; ++s) tbv = tbv + R[KD_TB + s] * conc[s];
      }
    } else {
      const double k0 = R[KD_A] * exp(R[KD_A + 1] * lnT - R[KD_A + 2] * Rtti);
      const double kinf = R[KD_AINF] * exp(R[KD_AINF + 1] * lnT - R[KD_AINF + 2] * Rtti);
      double tb = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) tb = tb + R[KD_TB + s] * conc[s];
      const double rpres = k0 * tb / kinf;
      if (ty == 2) {
        kf = kinf * (rpres / (1.0 + rpres));
      } else if (ty == 3) {
        if (rpres < 1e-30) {
          kf = k0 * tb;
        } else {
          double fcent = (1.0 - R[KD_FC]) * exp(-tt / R[KD_FC + 1]) + R[KD_FC] * exp(-tt / R[KD_FC + 2]);
          if (I[KI_T2]) fcent = fcent + exp(-R[KD_FC + 3] / tt);
          fcent = log(fcent) * ILN10;
          const double coff = -0.4 - 0.67 * fcent;
          const double noff = 0.75 - 1.27 * fcent;
          const double lr10 = log(rpres) * ILN10;
          const double fo = (lr10 + coff) / (noff - 0.14 * (lr10 + coff));
          kf = kinf * (rpres / (1.0 + rpres)) * exp(LN10 * (fcent / (1.0 + fo * fo)));
        }
      } else {
        const double l10 = log(rpres) * ILN10;
        const double xoff = 1.0 / (1.0 + l10 * l10);
        const double sri = R[KD_FC] * exp(-R[KD_FC + 1] / tt) + exp(-tt / R[KD_FC + 2]);
        const double foff = R[KD_FC + 3] * exp(xoff * log(sri)) * exp(R[KD_FC + 4] * lnT);
        kf = kinf * (rpres / (1.0 + rpres)) * foff;
      }
    }
    double kb = 0.0;
    if (I[KI_REV] == 1) {
      const double *kp = K.kc_tab + (itt - 1) + ntab * lr;
      kb = kf / (kp[0] + (kp[1] - kp[0]) * dtt);
    }
    double q1 = 1.0, q2 = 1.0;
    const int nre = I[KI_NRE], npr = I[KI_NPR];
    if (REF) {
      RefSpecies<0, NS>::products(R, I, nre, npr, conc, q1, q2);
      const double q = tbv * (kf * q1 - kb * q2);
      RefSpecies<0, NS>::accumulate(R, I, nre, npr, q, mw, acc);
    } else {
#pragma unroll
      for (int e = 0; e < KSIDE; ++e)
        if (e < nre) q1 = q1 * conc_power(conc[I[KI_RE + e]], R[KD_REX + e]);
#pragma unroll
      for (int e = 0; e < KSIDE; ++e)
        if (e < npr) q2 = q2 * conc_power(conc[I[KI_PR + e]], R[KD_PEX + e]);
      const double q = tbv * (kf * q1 - kb * q2);
#pragma unroll
      for (int s = 0; s < NS; ++s) acc[s] = acc[s] - q * R[KD_NET + s];
    }
  }
}

}  // namespace streams
