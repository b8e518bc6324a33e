// chem_common.h — finite-rate kinetics and the FP64 Rosenbrock step, device side.
//
// Rate expressions restate compute_chemistry_cuf / rosenbrock_rhs
// (kernels_gpu.F90:9551-9649, :9249-9339): Arrhenius, three-body, Lindemann, Troe (F_cent
// with the T2 sentinel -3.14) and SRI forward rates, backward rates through the tabulated
// K_c(T) (linear interpolation), mass-action products with integer or real exponents.
// The integrator restates rosenbrock_kernel/_step/_jacobian (:8928-9193) with the
// Numerical-Recipes ROS4 constants and ludcmp/lubksb (:9395-9501).
#pragma once
#include <cfloat>

#include "common.h"

namespace streams {

struct Kin {
  const double *mw, *arr_a, *arr_b, *arr_ea, *tb_eff, *falloff, *r_co, *p_co, *kc_tab;
  const int *reac_ty, *isrev;
  int nr, num_t_tab;
  double t_min_tab, dt_tab, r_univ;
};

template <int NS>
__device__ inline double forward_rate(const Kin &K, int lr, double tt, double Rtti, const double *conc) {
  const int nr = K.nr;
  const int ty = K.reac_ty[lr];
  double kf = 0.0;
  if (ty <= 1) {
    kf = K.arr_a[lr];
    if (K.arr_b[lr] != 0.0) kf = kf * pow(tt, K.arr_b[lr]);
    if (K.arr_ea[lr] != 0.0) kf = kf * exp(-K.arr_ea[lr] * Rtti);
    return kf;
  }
  const double k0 = K.arr_a[lr] * pow(tt, K.arr_b[lr]) * exp(-K.arr_ea[lr] * Rtti);
  const double kinf = K.arr_a[lr + nr] * pow(tt, K.arr_b[lr + nr]) * exp(-K.arr_ea[lr + nr] * Rtti);
  double tb = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) tb = tb + K.tb_eff[lr + nr * s] * conc[s];
  const double rpres = k0 * tb / kinf;
  if (ty == 2) {
    kf = kinf * (rpres / (1.0 + rpres));
  } else if (ty == 3) {
    if (rpres < 1e-30) {
      kf = k0 * tb;
    } else {
      const double *fc = K.falloff;
      double fcent = (1.0 - fc[lr]) * exp(-tt / fc[lr + nr]) + fc[lr] * exp(-tt / fc[lr + 2 * nr]);
      if (fc[lr + 3 * nr] != -3.14) fcent = fcent + exp(-fc[lr + 3 * nr] / tt);
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
    const double *fc = K.falloff;
    const double xoff = 1.0 / (1.0 + pow(log10(rpres), 2.0));
    const double foff = fc[lr + 3 * nr] * pow(fc[lr] * exp(-fc[lr + nr] / tt) + exp(-tt / fc[lr + 2 * nr]), xoff) *
                        pow(tt, fc[lr + 4 * nr]);
    kf = kinf * (rpres / (1.0 + rpres)) * foff;
  }
  return kf;
}

template <int NS>
__device__ inline double progress_rate(const Kin &K, int lr, double tt, double Rtti, int itt, double dtt,
                                       const double *conc) {
  const int nr = K.nr;
  const double kf = forward_rate<NS>(K, lr, tt, Rtti, conc);
  double kb = 0.0;
  if (K.isrev[lr] == 1) {
    const double *kp = K.kc_tab + (itt - 1) + (long)(K.num_t_tab + 1) * lr;
    const double dkc = kp[1] - kp[0];
    const double kc = kp[0] + dkc * dtt;
    kb = kf / kc;
  }
  double q1 = 1.0, q2 = 1.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double rr = K.r_co[lr + nr * s], pp = K.p_co[lr + nr * s];
    if (rr == 1.0) q1 = q1 * conc[s];
    else if (rr != 0.0) q1 = q1 * pow(conc[s], rr);
    if (pp == 1.0) q2 = q2 * conc[s];
    else if (pp != 0.0) q2 = q2 * pow(conc[s], pp);
  }
  double tb = 1.0;
  if (K.reac_ty[lr] == 1) {
    tb = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) tb = tb + K.tb_eff[lr + nr * s] * conc[s];
  }
  return tb * (kf * q1 - kb * q2);
}

// rosenbrock_rhs on a cell state wv[NV] (species clipped in place, T guess updated)
// The PaSR closure of the LES path (rosenbrock_rhs, kernels_gpu.F90:9359-9379, and
// rosenbrock_krylov_rhs, :10954-10974): tau_c is the geometric mean of the slowest and the
// fastest species time Y_s / |wdot_s / rho| over the species with |wdot_s / rho| > 1e-8,
// and the species sources are scaled by tau_c / (tau_c + tau_k).  With tau_k not above the
// smallest normal double the factor is 1 and the call is skipped.
template <int NS>
__device__ inline void pasr_scale(double tau_k, double ri, const double *yy, double *rhs) {
  double tau_cf = 0.0, tau_cs = DBL_MAX;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double a = fabs(rhs[s] * ri);
    if (a > 1.0e-8) {
      const double tau = yy[s] / a;
      tau_cf = tau > tau_cf ? tau : tau_cf;
      if (tau > 0.0 && tau < tau_cs) tau_cs = tau;
    }
  }
  const double tau_c = sqrt(tau_cs * tau_cf);
  const double gam = tau_c / (tau_c + tau_k);
#pragma unroll
  for (int s = 0; s < NS; ++s) rhs[s] = gam * rhs[s];
}

// The cell's mixing time of the PaSR closure (kernels_gpu.F90:8970-8980):
// Delta = (dxi/dx deta/dy dzeta/dz)^(-1/3), eps_sgs = c_eps c_yoshi^(3/2) nu_sgs^3 / Delta^4,
// tau_k = c_mix sqrt(nu_eff / eps_sgs) with nu_sgs = mu_sgs / rho and nu_eff = mu / rho.
__device__ inline double pasr_tau_k(double rho, double mu_sgs, double mu, double dx, double dy, double dz,
                                    double c_yoshi, double c_mix, double c_eps) {
  const double ri = 1.0 / rho;
  const double delta = pow(dx * dy * dz, -1.0 / 3.0);
  const double nu_sgs = mu_sgs * ri;
  const double d2 = delta * delta;
  const double eps_sgs = c_eps * pow(c_yoshi, 1.5) * (nu_sgs * nu_sgs * nu_sgs) / (d2 * d2);
  const double r = (mu * ri) / (eps_sgs + DBL_MIN);
  return c_mix * sqrt(r > 0.0 ? r : 0.0);
}

// The closure's constants and metrics as the integrator kernels take them (enable_pasr,
// les_c_yoshi, les_c_mix, les_c_eps, dcsidx, detady, dzitdz of rosenbrock_kernel).
struct Pasr {
  bool on;
  double c_yoshi, c_mix, c_eps;
  const double *dcsidx, *detady, *dzitdz;
  // tau_k of interior cell (i, j, k) from rho, mu_sgs and mu in w_aux (point pt, stride ps)
  template <int NS>
  __device__ inline double tau_k(const double *w_aux, long pt, long ps, int i, int j, int k) const {
    using X = Idx<NS>;
    return pasr_tau_k(w_aux[pt + ps * (X::J_R - 1)], w_aux[pt + ps * (X::J_LES1 - 1)], w_aux[pt + ps * (X::J_MU - 1)],
                      dcsidx[i - 1], detady[j - 1], dzitdz[k - 1], c_yoshi, c_mix, c_eps);
  }
};

template <int NS>
__device__ void ros_rhs(const Kin &K, const Thermo &th, double tol_iter_nr, double *wv, double &tg, double *rhs,
                        double tau_k = 0.0) {
  constexpr int NV = NS + 4;
  using X = Idx<NS>;
  double yy[NS], conc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) wv[s] = wv[s] > 0.0 ? wv[s] : 0.0;
  double rho = wv[0];
#pragma unroll
  for (int s = 1; s < NS; ++s) rho = rho + wv[s];
  const double ri = 1.0 / rho;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    yy[s] = wv[s] * ri;
    conc[s] = rho * yy[s] / K.mw[s];
  }
  const double uu = wv[X::I_U - 1] * ri, vv = wv[X::I_V - 1] * ri, ww = wv[X::I_W - 1] * ri;
  const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
  const double ee = wv[X::I_E - 1] * ri - qq;
  const double tt = temperature_from_e<NS>(ee, tg, th, tol_iter_nr, yy);
  tg = tt;
  const double Rtti = 1.0 / (K.r_univ * tt);
  int itt = (int)((tt - K.t_min_tab) / K.dt_tab) + 1;
  itt = itt < 1 ? 1 : itt;
  itt = itt > K.num_t_tab ? K.num_t_tab : itt;
  const double ttleft = (itt - 1) * K.dt_tab + K.t_min_tab;
  const double dtt = (tt - ttleft) / K.dt_tab;
#pragma unroll
  for (int m = 0; m < NV; ++m) rhs[m] = 0.0;
  for (int lr = 0; lr < K.nr; ++lr) {
    const double q = progress_rate<NS>(K, lr, tt, Rtti, itt, dtt, conc);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double rr = K.r_co[lr + K.nr * s], pp = K.p_co[lr + K.nr * s];
      if (rr != 0.0 || pp != 0.0) rhs[s] = rhs[s] + ((pp - rr) * q) * K.mw[s];
    }
  }
  if (tau_k > DBL_MIN) pasr_scale<NS>(tau_k, ri, yy, rhs);
}

// ludcmp: Crout LU with implicit partial pivoting, a is n x n column-major.  A row whose
// largest magnitude is below tiny (or NaN) ends the factorisation early, as in the
// reference, which then leaves indx unset; here indx starts as the identity permutation so
// that lubksb never indexes outside its vector (the result is then not a solve, as there).
template <int N, class T = double>
__device__ void ludcmp(T *a, int *indx) {
  const T tiny = (T)1.5e-16;
  T vv[N];
#define A_(r, c) a[(r) + N * (c)]
  for (int i = 0; i < N; ++i) indx[i] = i;
  for (int i = 0; i < N; ++i) {
    T amax = 0;
    for (int j = 0; j < N; ++j) amax = fabs(A_(i, j)) > amax ? fabs(A_(i, j)) : amax;
    if (amax < tiny) return;
    vv[i] = (T)1 / amax;
  }
  int imax = 0;
  for (int j = 0; j < N; ++j) {
    for (int i = 0; i < j; ++i) {
      T summ = A_(i, j);
      for (int k = 0; k < i; ++k) summ = summ - A_(i, k) * A_(k, j);
      A_(i, j) = summ;
    }
    T amax = 0;
    for (int i = j; i < N; ++i) {
      T summ = A_(i, j);
      for (int k = 0; k < j; ++k) summ = summ - A_(i, k) * A_(k, j);
      A_(i, j) = summ;
      T dum = vv[i] * fabs(summ);
      if (dum >= amax) { imax = i; amax = dum; }
    }
    if (j != imax) {
      for (int k = 0; k < N; ++k) { T d = A_(imax, k); A_(imax, k) = A_(j, k); A_(j, k) = d; }
      vv[imax] = vv[j];
    }
    indx[j] = imax;
    if (fabs(A_(j, j)) < tiny) A_(j, j) = tiny;
    if (j != N - 1) {
      T dum = (T)1 / A_(j, j);
      for (int i = j + 1; i < N; ++i) A_(i, j) = A_(i, j) * dum;
    }
  }
#undef A_
}

template <int N, class T = double>
__device__ void lubksb(const T *a, const int *indx, T *bv) {
#define A_(r, c) a[(r) + N * (c)]
  int ii = -1;
  for (int i = 0; i < N; ++i) {
    int ll = indx[i];
    T summ = bv[ll];
    bv[ll] = bv[i];
    if (ii >= 0) {
      for (int j = ii; j < i; ++j) summ = summ - A_(i, j) * bv[j];
    } else if (summ != (T)0) {
      ii = i;
    }
    bv[i] = summ;
  }
  for (int i = N - 1; i >= 0; --i) {
    T summ = bv[i];
    for (int j = i + 1; j < N; ++j) summ = summ - A_(i, j) * bv[j];
    bv[i] = summ / A_(i, i);
  }
#undef A_
}

// Step counts out; for a resumable integration (resume = true, the batched reactors with a
// step budget) the first step is clipped to the interval like every later one, and the
// time reached, the next step size and whether time_end was reached come back.
struct RosStats {
  int nacc = 0, nrej = 0;
  bool resume = false;
  int done = 0;
  double t_end = 0.0, dt_next = 0.0;
};

// One cell from time_start to time_end (rosenbrock_kernel body); rhs(wv, tg, out) is the
// chemistry right-hand side of rosenbrock_rhs (it may clip wv's species and updates tg).
template <int NS, class Rhs>
__device__ void rosenbrock_cell_f(const Rhs &rhs_fn, double *wv, double &tg, double dttry, double time_start,
                                  double time_end, int maxsteps, int maxtry, double tol, RosStats &st) {
  constexpr int NV = NS + 4;
  const double safety = 0.9, grow = 1.5, pgrow = -0.25, shrnk = 0.5, pshrnk = -1.0 / 3.0, errcon = 0.1296;
  const double gam = 1.0 / 2.0, a21 = 2.0, a31 = 48.0 / 25.0, a32 = 6.0 / 25.0;
  const double c21 = -8.0, c31 = 372.0 / 25.0, c32 = 12.0 / 5.0;
  const double c41 = -112.0 / 125.0, c42 = -54.0 / 125.0, c43 = -2.0 / 5.0;
  const double b1 = 19.0 / 9.0, b2 = 1.0 / 2.0, b3 = 25.0 / 108.0, b4 = 125.0 / 108.0;
  const double e1 = 17.0 / 54.0, e2 = 7.0 / 36.0, e3 = 0.0, e4 = 125.0 / 108.0;
  double jac[NV * NV], amat[NV * NV];
  double w_sav[NV], rhs_sav[NV], yscal[NV], rhs[NV], rp[NV], rm[NV];
  double g1[NV], g2[NV], g3[NV], g4[NV];
  int indx[NV];
  int finish = 0;
  double time = time_start, dt = dttry, itol = 1.0 / tol, dtnext = dt;
  if (st.resume && dt * 1.5 > time_end - time) { dt = time_end - time; finish = 1; }
  for (int iter = 1; iter <= maxsteps; ++iter) {
    const double t0 = time;
    for (int n = 0; n < NV; ++n) {
      const double wtmp = wv[n];
      const double h = wtmp * 1.0e-06 + 1.0e-14;
      wv[n] = wtmp + h;
      rhs_fn(wv, tg, rp);
      wv[n] = wtmp - h;
      rhs_fn(wv, tg, rm);
      wv[n] = wtmp;
      for (int m = 0; m < NV; ++m) jac[m + NV * n] = (rp[m] - rm[m]) / (2.0 * h);
    }
    rhs_fn(wv, tg, rhs_sav);
    for (int m = 0; m < NV; ++m) {
      w_sav[m] = wv[m];
      yscal[m] = fmax(1.0, fabs(w_sav[m]));
    }
    for (int jtry = 1; jtry <= maxtry; ++jtry) {
      for (int n = 0; n < NV; ++n) {
        for (int m = 0; m < NV; ++m) amat[m + NV * n] = -jac[m + NV * n];
        amat[n + NV * n] = amat[n + NV * n] + 1.0 / (gam * dt);
      }
      ludcmp<NV>(amat, indx);
      for (int m = 0; m < NV; ++m) g1[m] = rhs_sav[m];
      lubksb<NV>(amat, indx, g1);
      for (int m = 0; m < NV; ++m) wv[m] = w_sav[m] + a21 * g1[m];
      time = t0 + dt;
      rhs_fn(wv, tg, rhs);
      for (int m = 0; m < NV; ++m) g2[m] = rhs[m] + c21 * g1[m] / dt;
      lubksb<NV>(amat, indx, g2);
      for (int m = 0; m < NV; ++m) wv[m] = w_sav[m] + a31 * g1[m] + a32 * g2[m];
      time = t0 + 3.0 / 5.0 * dt;
      rhs_fn(wv, tg, rhs);
      for (int m = 0; m < NV; ++m) g3[m] = rhs[m] + (c31 * g1[m] + c32 * g2[m]) / dt;
      lubksb<NV>(amat, indx, g3);
      for (int m = 0; m < NV; ++m) g4[m] = rhs[m] + (c41 * g1[m] + c42 * g2[m] + c43 * g3[m]) / dt;
      lubksb<NV>(amat, indx, g4);
      double errmax = 0.0;
      for (int m = 0; m < NV; ++m) {
        wv[m] = w_sav[m] + b1 * g1[m] + b2 * g2[m] + b3 * g3[m] + b4 * g4[m];
        const double err = e1 * g1[m] + e2 * g2[m] + e3 * g3[m] + e4 * g4[m];
        errmax = fmax(errmax, fabs(err / yscal[m]));
      }
      time = t0 + dt;
      errmax = errmax * itol;
      if (errmax < 1.0 || finish == 1) {
        if (errmax > errcon) dtnext = safety * dt * pow(errmax, pgrow);
        else dtnext = grow * dt;
        break;
      } else {
        st.nrej += 1;
        dtnext = safety * dt * pow(errmax, pshrnk);
        dt = copysign(fmax(fabs(dtnext), shrnk * fabs(dt)), dt);
      }
    }
    st.nacc += 1;
    const double remaining = time_end - time;
    if (finish == 1 || remaining < 1.0e-12 * time_end) { st.done = 1; break; }
    if (dtnext * 1.5 > remaining) { dt = remaining; finish = 1; }
    else { dt = dtnext; finish = 0; }
  }
  st.t_end = time;
  st.dt_next = st.done ? dtnext : dt;
}

// The reference-form right-hand side (rosenbrock_rhs) bound to its tables.
template <int NS>
struct RosRhsRef {
  const Kin &K;
  const Thermo &th;
  double tol_iter_nr;
  double tau_k;   // PaSR mixing time; 0 without the closure
  __device__ inline void operator()(double *wv, double &tg, double *out) const {
    ros_rhs<NS>(K, th, tol_iter_nr, wv, tg, out, tau_k);
  }
};

template <int NS>
__device__ void rosenbrock_cell(const Kin &K, const Thermo &th, double tol_iter_nr, double *wv, double &tg,
                                double dttry, double time_start, double time_end, int maxsteps, int maxtry,
                                double tol, RosStats &st, double tau_k = 0.0) {
  const RosRhsRef<NS> f{K, th, tol_iter_nr, tau_k};
  rosenbrock_cell_f<NS>(f, wv, tg, dttry, time_start, time_end, maxsteps, maxtry, tol, st);
}

}  // namespace streams

namespace streams {

// Species source term -omega_s*W_s accumulated into acc[] (compute_chemistry_cuf's reaction
// loop, kernels_gpu.F90:9551-9649) with the rate constants evaluated through exp/log
// identities: A T^b exp(-Ea/RT) = A exp(b ln T - Ea/RT), log10(x) = ln(x)/ln(10),
// 10^x = exp(x ln 10), c^2 = c*c.  Each identity changes a result by a few units in the
// last place only.
template <int NS>
__device__ inline void chemistry_source(const Kin &K, double tt, const double *conc, double *acc) {
  const int nr = K.nr;
  const double Rtti = 1.0 / (K.r_univ * tt);
  const double lnT = log(tt);
  constexpr double LN10 = 2.302585092994045684;
  constexpr double ILN10 = 0.434294481903251827651;
  int itt = (int)((tt - K.t_min_tab) / K.dt_tab) + 1;
  itt = itt < 1 ? 1 : itt;
  itt = itt > K.num_t_tab ? K.num_t_tab : itt;
  const double ttleft = (itt - 1) * K.dt_tab + K.t_min_tab;
  const double dtt = (tt - ttleft) / K.dt_tab;
  for (int lr = 0; lr < nr; ++lr) {
    const int ty = K.reac_ty[lr];
    double kf;
    double tb = 0.0;
    if (ty <= 1) {
      const double b0 = K.arr_b[lr], e0 = K.arr_ea[lr];
      kf = K.arr_a[lr];
      if (b0 != 0.0 || e0 != 0.0) kf = kf * exp(b0 * lnT - e0 * Rtti);
    } else {
      const double k0 = K.arr_a[lr] * exp(K.arr_b[lr] * lnT - K.arr_ea[lr] * Rtti);
      const double kinf = K.arr_a[lr + nr] * exp(K.arr_b[lr + nr] * lnT - K.arr_ea[lr + nr] * Rtti);
#pragma unroll
      for (int s = 0; s < NS; ++s) tb = tb + K.tb_eff[lr + nr * s] * conc[s];
      const double rpres = k0 * tb / kinf;
      if (ty == 2) {
        kf = kinf * (rpres / (1.0 + rpres));
      } else if (ty == 3) {
        if (rpres < 1e-30) {
          kf = k0 * tb;
        } else {
          const double *fc = K.falloff;
          double fcent = (1.0 - fc[lr]) * exp(-tt / fc[lr + nr]) + fc[lr] * exp(-tt / fc[lr + 2 * nr]);
          if (fc[lr + 3 * nr] != -3.14) fcent = fcent + exp(-fc[lr + 3 * nr] / tt);
          fcent = log(fcent) * ILN10;
          const double coff = -0.4 - 0.67 * fcent;
          const double noff = 0.75 - 1.27 * fcent;
          const double lr10 = log(rpres) * ILN10;
          const double fo = (lr10 + coff) / (noff - 0.14 * (lr10 + coff));
          const double foff = exp(LN10 * (fcent / (1.0 + fo * fo)));
          kf = kinf * (rpres / (1.0 + rpres)) * foff;
        }
      } else {
        const double *fc = K.falloff;
        const double l10 = log(rpres) * ILN10;
        const double xoff = 1.0 / (1.0 + l10 * l10);
        const double foff = fc[lr + 3 * nr] * pow(fc[lr] * exp(-fc[lr + nr] / tt) + exp(-tt / fc[lr + 2 * nr]), xoff) *
                            exp(fc[lr + 4 * nr] * lnT);
        kf = kinf * (rpres / (1.0 + rpres)) * foff;
      }
    }
    double kb = 0.0;
    if (K.isrev[lr] == 1) {
      const double *kp = K.kc_tab + (itt - 1) + (long)(K.num_t_tab + 1) * lr;
      kb = kf / (kp[0] + (kp[1] - kp[0]) * dtt);
    }
    double q1 = 1.0, q2 = 1.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double rr = K.r_co[lr + nr * s], pp = K.p_co[lr + nr * s];
      if (rr == 1.0) q1 = q1 * conc[s];
      else if (rr == 2.0) q1 = q1 * (conc[s] * conc[s]);
      else if (rr != 0.0) q1 = q1 * pow(conc[s], rr);
      if (pp == 1.0) q2 = q2 * conc[s];
      else if (pp == 2.0) q2 = q2 * (conc[s] * conc[s]);
      else if (pp != 0.0) q2 = q2 * pow(conc[s], pp);
    }
    if (ty == 1) {
      tb = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) tb = tb + K.tb_eff[lr + nr * s] * conc[s];
    } else {
      tb = 1.0;
    }
    const double q = tb * (kf * q1 - kb * q2);
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = acc[s] - (K.p_co[lr + nr * s] - K.r_co[lr + nr * s]) * q * K.mw[s];
  }
}

}  // namespace streams
