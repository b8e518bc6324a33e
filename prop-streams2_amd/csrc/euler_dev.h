// euler_dev.h — device functions of the hybrid KEEP / WENO convective flux, shared by
// the per-direction convective kernels and the fused convective+viscous sweeps.
// Arithmetic follows euler_x_fluxes_hybrid_kernel (kernels_gpu.F90:88-425),
// euler_y_hybrid_kernel (:1427-1759), euler_z_hybrid_kernel (:795-1123), the Rusanov
// variants (:427-730, :1125-1425, :1761-2063), compute_roe_average (:2986-3295),
// eigenvectors_x/y/z (:2703-2984), wenorec_1d[_rusanov] (:2065-2637).
#pragma once
#include "common.h"

namespace streams {


struct EulerArgs {
  const double *w_aux, *cd1, *metric, *rgas;
  double *fhat, *fl;
  const int *epo;
  Thermo th;
  int lmax_base, nkeep, zmin, zmax, weno_scheme, weno_version, flux_splitting;
  double sensor_threshold, tol_iter_nr, rho0, u0;
  int fs, fe;      // face range
  int n;           // cells along the sweep
  int check_shock_flag = 0;  // sensor threshold below 1 (the fused sweep consults fhat)
};

// A line of the padded w_aux along direction DIR: cell p at base + p*st.
struct Line {
  const double *a;
  long base, st, ps;
  __device__ inline double operator()(int p, int m) const { return a[base + (long)p * st + ps * (m - 1)]; }
};

template <int DIR>
__device__ inline void line_base(const Box &b, int t1, int t2, long &base, long &st) {
  if (DIR == 1) { base = b.p(0, t1, t2); st = 1; }
  else if (DIR == 2) { base = b.p(t1, 0, t2); st = b.px(); }
  else { base = b.p(t1, t2, 0); st = b.px() * b.py(); }
}

template <int DIR>
__device__ inline int epo_at(const Box &b, const int *epo, int t1, int t2, int f) {
  int i, j, k;
  if (DIR == 1) { i = f; j = t1; k = t2; }
  else if (DIR == 2) { i = t1; j = f; k = t2; }
  else { i = t1; j = t2; k = f; }
  return epo[(long)i + (long)(b.nx + 1) * ((long)j + (long)(b.ny + 1) * ((long)k + (long)(b.nz + 1) * (DIR - 1)))];
}

template <int NS>
__device__ inline bool shocked(const Line &L, int f, int ws, double thr) {
  using X = Idx<NS>;
  bool s = false;
  for (int ii = f - ws + 1; ii <= f + ws; ++ii) s = s || (L(ii, X::J_DUC) > thr);
  return s;
}

// KEEP split-form central flux at face f (kernels_gpu.F90:150-337); LineT maps (cell
// position along the sweep, reference variable index) to a value (global or LDS)
template <int NS, int DIR, class LineT>
__device__ void keep_flux(const LineT &L, int f, int lmax, int nkeep, const double *cd1, bool zero,
                          double *fh) {
  using X = Idx<NS>;
  constexpr int JN = DIR == 1 ? X::J_U : DIR == 2 ? X::J_V : X::J_W;
  constexpr int IN = DIR == 1 ? X::I_U : DIR == 2 ? X::I_V : X::I_W;
  // pair factors (rho_a+rho_b)(un_a+un_b)*drhof, shared by the species fluxes
  double pf[10];
  int q = 0;
  for (int l = 1; l <= lmax; ++l) {
    for (int m = 0; m <= l - 1; ++m, ++q) {
      double rhoi = L(f - m, X::J_R), rhoip = L(f - m + l, X::J_R);
      double uni = L(f - m, JN), unip = L(f - m + l, JN);
      double rhom = rhoi + rhoip;
      double drhof = 1.0;
      if (nkeep > 0) {
        double drho = 2.0 * (rhoip - rhoi) / rhom;
        double ts = (0.5 * drho) * (0.5 * drho);
        double t2 = ts;
        double sden = 1.0 + ts / 3.0;
        for (int n = 2; n <= nkeep; ++n) { ts = t2 * ts; sden = sden + ts / (1.0 + 2 * n); }
        drhof = 1.0 / sden;
      }
      pf[q] = (uni + unip) * rhom * drhof;
    }
  }
#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    double ft1 = 0.0;
    int qq = 0;
    for (int l = 1; l <= lmax; ++l) {
      double uvs1 = 0.0;
      for (int m = 0; m <= l - 1; ++m, ++qq) uvs1 = uvs1 + pf[qq] * (L(f - m, s) + L(f - m + l, s));
      ft1 = ft1 + cd1[(l - 1) + 4 * (lmax - 1)] * uvs1;
    }
    fh[s - 1] = zero ? 0.0 : 0.25 * ft1;
  }
  double ft2 = 0, ft3 = 0, ft4 = 0, ft5 = 0, ft6 = 0;
  q = 0;
  for (int l = 1; l <= lmax; ++l) {
    double uvs2 = 0, uvs3 = 0, uvs4 = 0, uvs5 = 0, uvs5_i = 0, uvs5_k = 0, uvs5_p = 0, uvs6 = 0;
    for (int m = 0; m <= l - 1; ++m, ++q) {
      const int a = f - m, c = f - m + l;
      double rhoi = L(a, X::J_R), uui = L(a, X::J_U), vvi = L(a, X::J_V), wwi = L(a, X::J_W);
      double enti = L(a, X::J_H), ppi = L(a, X::J_P);
      double rhoip = L(c, X::J_R), uuip = L(c, X::J_U), vvip = L(c, X::J_V), wwip = L(c, X::J_W);
      double entip = L(c, X::J_H), ppip = L(c, X::J_P);
      double uni = DIR == 1 ? uui : DIR == 2 ? vvi : wwi;
      double unip = DIR == 1 ? uuip : DIR == 2 ? vvip : wwip;
      if (nkeep >= 0) {
        double rhom = rhoi + rhoip;
        double eei = enti - ppi / rhoi - 0.5 * (uui * uui + vvi * vvi + wwi * wwi);
        double eeip = entip - ppip / rhoip - 0.5 * (uuip * uuip + vvip * vvip + wwip * wwip);
        double eem = eei + eeip;
        double deef = 1.0;
        if (nkeep > 0) {
          double dee = 2.0 * (eeip - eei) / eem;
          double te = (0.5 * dee) * (0.5 * dee);
          double te2 = te;
          double sde = 1.0 + te, sne = 1.0 + te / 3.0;
          for (int n = 2; n <= nkeep; ++n) { te = te2 * te; sde = sde + te; sne = sne + te / (1.0 + 2 * n); }
          deef = sne / sde;
        }
        double uv_part = pf[q];
        uvs2 = uvs2 + uv_part * (uui + uuip);
        uvs3 = uvs3 + uv_part * (vvi + vvip);
        uvs4 = uvs4 + uv_part * (wwi + wwip);
        uvs5_i = uvs5_i + uv_part * eem * deef;
        uvs5_k = uvs5_k + uv_part * (uui * uuip + vvi * vvip + wwi * wwip);
        uvs5_p = uvs5_p + 4.0 * (uni * ppip + unip * ppi);
        uvs6 = uvs6 + 2.0 * (ppi + ppip);
      } else {
        double uv_part = (uni + unip) * (rhoi + rhoip);
        uvs2 = uvs2 + uv_part * (uui + uuip);
        uvs3 = uvs3 + uv_part * (vvi + vvip);
        uvs4 = uvs4 + uv_part * (wwi + wwip);
        uvs5 = uvs5 + uv_part * (enti + entip);
        uvs6 = uvs6 + 2.0 * (ppi + ppip);
      }
    }
    double cl = cd1[(l - 1) + 4 * (lmax - 1)];
    ft2 = ft2 + cl * uvs2;
    ft3 = ft3 + cl * uvs3;
    ft4 = ft4 + cl * uvs4;
    ft5 = ft5 + cl * (nkeep >= 0 ? (uvs5_i + uvs5_k + uvs5_p) : uvs5);
    ft6 = ft6 + cl * uvs6;
  }
  double fh2 = 0.25 * ft2, fh3 = 0.25 * ft3, fh4 = 0.25 * ft4, fh5 = 0.25 * ft5;
  if (zero) { fh2 = 0.0; fh3 = 0.0; fh4 = 0.0; fh5 = 0.0; }
  fh[X::I_U - 1] = fh2;
  fh[X::I_V - 1] = fh3;
  fh[X::I_W - 1] = fh4;
  fh[X::I_E - 1] = fh5;
  fh[IN - 1] = fh[IN - 1] + 0.5 * ft6;
}

// ---------------------------------------------------------------------------- WENO
template <int NS>
__device__ void wenorec(const double *vp, const double *vm, double *vhat, int iweno, int ord, int version,
                        double rho0, double u0, bool rusanov) {
  constexpr int NV = NS + 4;
  double bsc[NV];
  const double u0_2 = u0 * u0, r2u2 = rho0 * rho0 * u0_2, r2u4 = r2u2 * u0_2;
  if (!rusanov) {
#pragma unroll
    for (int s = 0; s < NS + 1; ++s) bsc[s] = 1.0 / r2u2;
    bsc[NS + 1] = 1.0 / r2u4;
    bsc[NS + 2] = bsc[NS + 1];
    bsc[NS + 3] = 1.0 / r2u2;
  } else {
    const double r2u4i = 1.0 / r2u4;
#pragma unroll
    for (int s = 0; s < NS; ++s) bsc[s] = 1.0 / r2u2;
    bsc[NS] = r2u4i; bsc[NS + 1] = r2u4i; bsc[NS + 2] = r2u4i;
    bsc[NS + 3] = r2u4i / u0_2;
  }
#define VP(m, l) vp[(m) + NV * ((l) - 1)]
#define VM(m, l) vm[(m) + NV * ((l) - 1)]
  const int i = iweno;
  for (int m = 0; m < NV; ++m) {
    if (ord == 1) {
      vhat[m] = VP(m, i) + VM(m, i + 1);
    } else if (ord == 2) {
      double bp0 = (VP(m, i) - VP(m, i - 1)) * (VP(m, i) - VP(m, i - 1));
      double bp1 = (VP(m, i + 1) - VP(m, i)) * (VP(m, i + 1) - VP(m, i));
      bp1 = bsc[m] * bp1; bp0 = bsc[m] * bp0;
      double bm0 = (VM(m, i + 2) - VM(m, i + 1)) * (VM(m, i + 2) - VM(m, i + 1));
      double bm1 = (VM(m, i + 1) - VM(m, i)) * (VM(m, i + 1) - VM(m, i));
      bm1 = bsc[m] * bm1; bm0 = bsc[m] * bm0;
      const double d0 = 1.0 / 3.0, d1 = 2.0 / 3.0;
      double t;
      t = 0.000001 + bp0; bp0 = d0 / (t * t);
      t = 0.000001 + bm0; bm0 = d0 / (t * t);
      t = 0.000001 + bp1; bp1 = d1 / (t * t);
      t = 0.000001 + bm1; bm1 = d1 / (t * t);
      double sump = bp0 + bp1, summ = bm0 + bm1;
      bp0 = bp0 / sump; bp1 = bp1 / sump; bm0 = bm0 / summ; bm1 = bm1 / summ;
      double vminus = bp0 * (-VP(m, i - 1) + 3 * VP(m, i)) + bp1 * (VP(m, i) + VP(m, i + 1));
      double vplus = bm0 * (-VM(m, i + 2) + 3 * VM(m, i + 1)) + bm1 * (VM(m, i) + VM(m, i + 1));
      vhat[m] = 0.5 * (vminus + vplus);
    } else if (ord == 3) {
      const double d0 = 13.0 / 12.0, d1 = 1.0 / 4.0;
      const double c0 = 1.0 / 3.0, c1 = 5.0 / 6.0, c2 = -1.0 / 6.0, c3 = -7.0 / 6.0, c4 = 11.0 / 6.0;
      const double dw0 = 1.0 / 10.0, dw1 = 6.0 / 10.0, dw2 = 3.0 / 10.0;
      double a, bq;
      a = VP(m, i) - 2.0 * VP(m, i + 1) + VP(m, i + 2); bq = 3.0 * VP(m, i) - 4.0 * VP(m, i + 1) + VP(m, i + 2);
      double bp2 = d0 * a * a + d1 * bq * bq;
      a = VP(m, i - 1) - 2.0 * VP(m, i) + VP(m, i + 1); bq = VP(m, i - 1) - VP(m, i + 1);
      double bp1 = d0 * a * a + d1 * bq * bq;
      a = VP(m, i) - 2.0 * VP(m, i - 1) + VP(m, i - 2); bq = 3.0 * VP(m, i) - 4.0 * VP(m, i - 1) + VP(m, i - 2);
      double bp0 = d0 * a * a + d1 * bq * bq;
      bp2 = bsc[m] * bp2; bp1 = bsc[m] * bp1; bp0 = bsc[m] * bp0;
      a = VM(m, i + 1) - 2.0 * VM(m, i) + VM(m, i - 1); bq = 3.0 * VM(m, i + 1) - 4.0 * VM(m, i) + VM(m, i - 1);
      double bm2 = d0 * a * a + d1 * bq * bq;
      a = VM(m, i + 2) - 2.0 * VM(m, i + 1) + VM(m, i); bq = VM(m, i + 2) - VM(m, i);
      double bm1 = d0 * a * a + d1 * bq * bq;
      a = VM(m, i + 1) - 2.0 * VM(m, i + 2) + VM(m, i + 3); bq = 3.0 * VM(m, i + 1) - 4.0 * VM(m, i + 2) + VM(m, i + 3);
      double bm0 = d0 * a * a + d1 * bq * bq;
      bm2 = bsc[m] * bm2; bm1 = bsc[m] * bm1; bm0 = bsc[m] * bm0;
      if (version == 0) {
        double t;
        t = 0.000001 + bp0; bp0 = dw0 / (t * t);
        t = 0.000001 + bm0; bm0 = dw0 / (t * t);
        t = 0.000001 + bp1; bp1 = dw1 / (t * t);
        t = 0.000001 + bm1; bm1 = dw1 / (t * t);
        t = 0.000001 + bp2; bp2 = dw2 / (t * t);
        t = 0.000001 + bm2; bm2 = dw2 / (t * t);
      } else {
        const double eps40 = 1.0e-40;
        double tau5p = fabs(bp0 - bp2) + eps40, tau5m = fabs(bm0 - bm2) + eps40;
        bp0 = (bp0 + eps40) / (bp0 + tau5p); bm0 = (bm0 + eps40) / (bm0 + tau5m);
        bp1 = (bp1 + eps40) / (bp1 + tau5p); bm1 = (bm1 + eps40) / (bm1 + tau5m);
        bp2 = (bp2 + eps40) / (bp2 + tau5p); bm2 = (bm2 + eps40) / (bm2 + tau5m);
        bp0 = dw0 / bp0; bm0 = dw0 / bm0;
        bp1 = dw1 / bp1; bm1 = dw1 / bm1;
        bp2 = dw2 / bp2; bm2 = dw2 / bm2;
      }
      double sump = bp0 + bp1 + bp2, summ = bm0 + bm1 + bm2;
      bp0 = bp0 / sump; bp1 = bp1 / sump; bp2 = bp2 / sump;
      bm0 = bm0 / summ; bm1 = bm1 / summ; bm2 = bm2 / summ;
      double vminus = bp2 * (c0 * VP(m, i) + c1 * VP(m, i + 1) + c2 * VP(m, i + 2)) +
                      bp1 * (c2 * VP(m, i - 1) + c1 * VP(m, i) + c0 * VP(m, i + 1)) +
                      bp0 * (c0 * VP(m, i - 2) + c3 * VP(m, i - 1) + c4 * VP(m, i));
      double vplus = bm2 * (c0 * VM(m, i + 1) + c1 * VM(m, i) + c2 * VM(m, i - 1)) +
                     bm1 * (c2 * VM(m, i + 2) + c1 * VM(m, i + 1) + c0 * VM(m, i)) +
                     bm0 * (c0 * VM(m, i + 3) + c3 * VM(m, i + 2) + c4 * VM(m, i + 1));
      vhat[m] = vminus + vplus;
    } else {  // ord == 4, WENO-7
      const double d1 = 1.0 / 36.0, d2 = 13.0 / 12.0, d3 = 781.0 / 720.0;
      const double dw[4] = {1.0 / 35.0, 12.0 / 35.0, 18.0 / 35.0, 4.0 / 35.0};
      double bp[4], bm[4], a, bq, c;
      a = -11 * VP(m, i) + 18 * VP(m, i + 1) - 9 * VP(m, i + 2) + 2 * VP(m, i + 3);
      bq = 2 * VP(m, i) - 5 * VP(m, i + 1) + 4 * VP(m, i + 2) - VP(m, i + 3);
      c = -VP(m, i) + 3 * VP(m, i + 1) - 3 * VP(m, i + 2) + VP(m, i + 3);
      bp[3] = d1 * a * a + d2 * bq * bq + d3 * c * c;
      a = -2 * VP(m, i - 1) - 3 * VP(m, i) + 6 * VP(m, i + 1) - VP(m, i + 2);
      bq = VP(m, i - 1) - 2 * VP(m, i) + VP(m, i + 1);
      c = -VP(m, i - 1) + 3 * VP(m, i) - 3 * VP(m, i + 1) + VP(m, i + 2);
      bp[2] = d1 * a * a + d2 * bq * bq + d3 * c * c;
      a = VP(m, i - 2) - 6 * VP(m, i - 1) + 3 * VP(m, i) + 2 * VP(m, i + 1);
      bq = VP(m, i - 1) - 2 * VP(m, i) + VP(m, i + 1);
      c = -VP(m, i - 2) + 3 * VP(m, i - 1) - 3 * VP(m, i) + VP(m, i + 1);
      bp[1] = d1 * a * a + d2 * bq * bq + d3 * c * c;
      a = -2 * VP(m, i - 3) + 9 * VP(m, i - 2) - 18 * VP(m, i - 1) + 11 * VP(m, i);
      bq = -VP(m, i - 3) + 4 * VP(m, i - 2) - 5 * VP(m, i - 1) + 2 * VP(m, i);
      c = -VP(m, i - 3) + 3 * VP(m, i - 2) - 3 * VP(m, i - 1) + VP(m, i);
      bp[0] = d1 * a * a + d2 * bq * bq + d3 * c * c;
      a = -11 * VM(m, i + 1) + 18 * VM(m, i) - 9 * VM(m, i - 1) + 2 * VM(m, i - 2);
      bq = 2 * VM(m, i + 1) - 5 * VM(m, i) + 4 * VM(m, i - 1) - VM(m, i - 2);
      c = -VM(m, i + 1) + 3 * VM(m, i) - 3 * VM(m, i - 1) + VM(m, i - 2);
      bm[3] = d1 * a * a + d2 * bq * bq + d3 * c * c;
      a = -2 * VM(m, i + 2) - 3 * VM(m, i + 1) + 6 * VM(m, i) - VM(m, i - 1);
      bq = VM(m, i + 2) - 2 * VM(m, i + 1) + VM(m, i);
      c = -VM(m, i + 2) + 3 * VM(m, i + 1) - 3 * VM(m, i) + VM(m, i - 1);
      bm[2] = d1 * a * a + d2 * bq * bq + d3 * c * c;
      a = VM(m, i + 3) - 6 * VM(m, i + 2) + 3 * VM(m, i + 1) + 2 * VM(m, i);
      bq = VM(m, i + 2) - 2 * VM(m, i + 1) + VM(m, i);
      c = -VM(m, i + 3) + 3 * VM(m, i + 2) - 3 * VM(m, i + 1) + VM(m, i);
      bm[1] = d1 * a * a + d2 * bq * bq + d3 * c * c;
      a = -2 * VM(m, i + 4) + 9 * VM(m, i + 3) - 18 * VM(m, i + 2) + 11 * VM(m, i + 1);
      bq = -VM(m, i + 4) + 4 * VM(m, i + 3) - 5 * VM(m, i + 2) + 2 * VM(m, i + 1);
      c = -VM(m, i + 4) + 3 * VM(m, i + 3) - 3 * VM(m, i + 2) + VM(m, i + 1);
      bm[0] = d1 * a * a + d2 * bq * bq + d3 * c * c;
      double sump = 0, summ = 0;
      for (int l = 0; l < 4; ++l) {
        bp[l] = bsc[m] * bp[l];
        bm[l] = bsc[m] * bm[l];
      }
      for (int l = 0; l < 4; ++l) {
        double t = 0.000001 + bp[l];
        bp[l] = dw[l] / (t * t);
        t = 0.000001 + bm[l];
        bm[l] = dw[l] / (t * t);
        sump = sump + bp[l];
        summ = summ + bm[l];
      }
      for (int l = 0; l < 4; ++l) { bp[l] = bp[l] / sump; bm[l] = bm[l] / summ; }
      double vminus = bp[3] * (6 * VP(m, i) + 26 * VP(m, i + 1) - 10 * VP(m, i + 2) + 2 * VP(m, i + 3)) +
                      bp[2] * (-2 * VP(m, i - 1) + 14 * VP(m, i) + 14 * VP(m, i + 1) - 2 * VP(m, i + 2)) +
                      bp[1] * (2 * VP(m, i - 2) - 10 * VP(m, i - 1) + 26 * VP(m, i) + 6 * VP(m, i + 1)) +
                      bp[0] * (-6 * VP(m, i - 3) + 26 * VP(m, i - 2) - 46 * VP(m, i - 1) + 50 * VP(m, i));
      double vplus = bm[3] * (6 * VM(m, i + 1) + 26 * VM(m, i) - 10 * VM(m, i - 1) + 2 * VM(m, i - 2)) +
                     bm[2] * (-2 * VM(m, i + 2) + 14 * VM(m, i + 1) + 14 * VM(m, i) - 2 * VM(m, i - 1)) +
                     bm[1] * (2 * VM(m, i + 3) - 10 * VM(m, i + 2) + 26 * VM(m, i + 1) + 6 * VM(m, i)) +
                     bm[0] * (-6 * VM(m, i + 4) + 26 * VM(m, i + 3) - 46 * VM(m, i + 2) + 50 * VM(m, i + 1));
      vhat[m] = (vminus + vplus) / 24.0;
    }
  }
#undef VP
#undef VM
}

// WENO face flux (characteristic Roe/LLF or component-wise Rusanov)
template <int NS, int DIR>
__device__ void weno_flux(const Line &L, int f, const EulerArgs &A, int ord, double *fh) {
  using X = Idx<NS>;
  constexpr int NV = NS + 4;
  constexpr int JN = DIR == 1 ? X::J_U : DIR == 2 ? X::J_V : X::J_W;
  constexpr int IN = DIR == 1 ? X::I_U : DIR == 2 ? X::I_V : X::I_W;
  const int ws = A.weno_scheme, wsize = 2 * A.weno_scheme;
  double gp[NV * 8], gm[NV * 8];
  if (A.flux_splitting == 1) {
    double evmax = -1.0;
    for (int l = 1; l <= wsize; ++l) {
      int ll = f + l - ws;
      double un = L(ll, JN), c = L(ll, X::J_C);
      evmax = fmax(fmax(fabs(un - c), fabs(un + c)), evmax);
    }
    for (int l = 1; l <= wsize; ++l) {
      int ll = f + l - ws;
      double rho = L(ll, X::J_R), uu = L(ll, X::J_U), vv = L(ll, X::J_V), ww = L(ll, X::J_W);
      double h = L(ll, X::J_H), pp = L(ll, X::J_P), un = L(ll, JN);
      double rhou = rho * un, rhoevm = rho * evmax;
#pragma unroll
      for (int s = 1; s <= NS; ++s) {
        double yy = L(ll, s), evm = rhou * yy, c = 0.5 * (evm + rhoevm * yy);
        gp[(s - 1) + NV * (l - 1)] = c;
        gm[(s - 1) + NV * (l - 1)] = evm - c;
      }
      double vel[3] = {uu, vv, ww};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double evm = vel[d] * rhou + ((X::I_U + d) == IN ? pp : 0.0);
        double c = 0.5 * (evm + rhoevm * vel[d]);
        gp[(X::I_U + d - 1) + NV * (l - 1)] = c;
        gm[(X::I_U + d - 1) + NV * (l - 1)] = evm - c;
      }
      double evm = h * rhou, c = 0.5 * (evm + evmax * (rho * h - pp));
      gp[(X::I_E - 1) + NV * (l - 1)] = c;
      gm[(X::I_E - 1) + NV * (l - 1)] = evm - c;
    }
    wenorec<NS>(gp, gm, fh, ws, ord, A.weno_version, A.rho0, A.u0, true);
    return;
  }
  // ---- Roe average (compute_roe_average)
  const Thermo &th = A.th;
  double yyroe[NS], prhoi[NS], drho[NS];
  double rho = L(f, X::J_R), uu = L(f, X::J_U), vv = L(f, X::J_V), ww = L(f, X::J_W);
  double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
  double h = L(f, X::J_H), pp = L(f, X::J_P);
  double ee = h - pp / rho - qq;
  const double rho_l = rho, ee_l = ee;
  double rhop = L(f + 1, X::J_R), up = L(f + 1, X::J_U), vp = L(f + 1, X::J_V), wp = L(f + 1, X::J_W);
  double qqp = 0.5 * (up * up + vp * vp + wp * wp);
  double hp = L(f + 1, X::J_H), ppp = L(f + 1, X::J_P);
  double eep = hp - ppp / rhop - qqp;
  double r = sqrt(rhop / rho);
  rho = r * rho_l;
  double rp1 = 1.0 / (r + 1.0);
  uu = (r * up + uu) * rp1;
  vv = (r * vp + vv) * rp1;
  ww = (r * wp + ww) * rp1;
  h = (r * hp + h) * rp1;
  qq = 0.5 * (uu * uu + vv * vv + ww * ww);
  ee = (r * eep + ee) * rp1;
  const double p_over_rho = h - ee - qq;
  double rmix = 0.0;
#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    double yy = L(f, s), yyp = L(f + 1, s);
    drho[s - 1] = rhop * yyp - rho_l * yy;
    yy = (r * yyp + yy) * rp1;
    rmix = rmix + A.rgas[s - 1] * yy;
    yyroe[s - 1] = yy;
  }
  double tt = temperature_from_e<NS>(ee, p_over_rho / rmix, th, A.tol_iter_nr, yyroe);
  double cploc = mixture_cp<NS>(tt, th, yyroe);
  double gamloc = cploc / (cploc - rmix);
  double gm1 = gamloc - 1.0;
  double p_e = rho * gm1;
#pragma unroll
  for (int s = 1; s <= NS; ++s) prhoi[s - 1] = species_hcoef<NS>(tt, th.cv_coeff, th, s);
#pragma unroll
  for (int s = 1; s <= NS; ++s) prhoi[s - 1] = A.rgas[s - 1] * tt + gm1 * (ee - prhoi[s - 1]);
  const double etot = ee + qq;
  double pede = p_e * (eep - ee_l);
  double dp_err = (ppp - pp) - pede;
  double den = pede * pede;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double pr = prhoi[s] * drho[s];
    dp_err = dp_err - pr;
    den = den + pr * pr;
  }
  double dpfact = dp_err / fmax(den, 0.000001);
  p_e = p_e * (1.0 + pede * dpfact);
#pragma unroll
  for (int s = 0; s < NS; ++s) prhoi[s] = prhoi[s] * (1.0 + prhoi[s] * drho[s] * dpfact);
  double cc = p_over_rho / rho * p_e;
#pragma unroll
  for (int s = 0; s < NS; ++s) cc = cc + yyroe[s] * prhoi[s];
  const double c = sqrt(cc), ci = 1.0 / c, cci = ci * ci;
  const double b3 = etot, b2 = p_e / (rho * cc), b1 = -b2 * (etot - 2.0 * qq);
  const double un = DIR == 1 ? uu : DIR == 2 ? vv : ww;
  const double ta = DIR == 1 ? vv : uu, tb = DIR == 3 ? vv : ww;
  constexpr int IA = DIR == 1 ? X::I_V : X::I_U;
  constexpr int IB = DIR == 3 ? X::I_V : X::I_W;
  // ---- left eigenvectors el(row, field) (eigenvectors_x/y/z)
  double el[NV * NV];
#pragma unroll
  for (int q = 0; q < NV * NV; ++q) el[q] = 0.0;
#define EL(rr, cq) el[((rr) - 1) + NV * ((cq) - 1)]
#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    double b1i = b1 + prhoi[s - 1] * cci;
    EL(s, 1) = 0.5 * (b1i + un * ci);
    EL(s, NS + 2) = -ta;
    EL(s, NS + 3) = -tb;
    EL(s, NS + 4) = 0.5 * (b1i - un * ci);
    EL(s, s + 1) = 1.0;
    EL(X::I_U, 1 + s) = b2 * uu * yyroe[s - 1];
    EL(X::I_V, 1 + s) = b2 * vv * yyroe[s - 1];
    EL(X::I_W, 1 + s) = b2 * ww * yyroe[s - 1];
    EL(X::I_E, 1 + s) = -b2 * yyroe[s - 1];
#pragma unroll
    for (int m = 1; m <= NS; ++m) {
      double b1m = b1 + prhoi[m - 1] * cci;
      EL(m, 1 + s) = EL(m, 1 + s) - b1m * yyroe[s - 1];
    }
  }
  EL(X::I_U, 1) = -0.5 * (b2 * uu + (DIR == 1 ? ci : 0.0));
  EL(X::I_V, 1) = -0.5 * (b2 * vv + (DIR == 2 ? ci : 0.0));
  EL(X::I_W, 1) = -0.5 * (b2 * ww + (DIR == 3 ? ci : 0.0));
  EL(X::I_E, 1) = 0.5 * b2;
  EL(IA, NS + 2) = 1.0;
  EL(IB, NS + 3) = 1.0;
  EL(X::I_U, NS + 4) = -0.5 * (b2 * uu - (DIR == 1 ? ci : 0.0));
  EL(X::I_V, NS + 4) = -0.5 * (b2 * vv - (DIR == 2 ? ci : 0.0));
  EL(X::I_W, NS + 4) = -0.5 * (b2 * ww - (DIR == 3 ? ci : 0.0));
  EL(X::I_E, NS + 4) = 0.5 * b2;
  // ---- LLF eigenvalue bounds over the stencil
  double ev1 = -1.0, ev2 = -1.0, ev4 = -1.0;
  for (int l = 1; l <= wsize; ++l) {
    int ll = f + l - ws;
    double unl = L(ll, JN), cl = L(ll, X::J_C);
    ev1 = fmax(fabs(unl - cl), ev1);
    ev2 = fmax(fabs(unl), ev2);
    ev4 = fmax(fabs(unl + cl), ev4);
  }
  // ---- characteristic projection of the split fluxes
  for (int l = 1; l <= wsize; ++l) {
    int ll = f + l - ws;
    double rl = L(ll, X::J_R), ul = L(ll, X::J_U), vl = L(ll, X::J_V), wl = L(ll, X::J_W);
    double hl = L(ll, X::J_H), pl = L(ll, X::J_P), unl = L(ll, JN);
    double rhou = rl * unl;
    double fi[NV], yl[NS];
#pragma unroll
    for (int s = 1; s <= NS; ++s) { yl[s - 1] = L(ll, s); fi[s - 1] = rhou * yl[s - 1]; }
    fi[X::I_U - 1] = ul * rhou; fi[X::I_V - 1] = vl * rhou; fi[X::I_W - 1] = wl * rhou; fi[X::I_E - 1] = hl * rhou;
    fi[IN - 1] = fi[IN - 1] + pl;
#pragma unroll
    for (int m = 1; m <= NV; ++m) {
      double wc = 0.0, gc = 0.0;
#pragma unroll
      for (int s = 1; s <= NS; ++s) {
        wc = wc + EL(s, m) * rl * yl[s - 1];
        gc = gc + EL(s, m) * fi[s - 1];
      }
      wc = wc + EL(X::I_U, m) * rl * ul; gc = gc + EL(X::I_U, m) * fi[X::I_U - 1];
      wc = wc + EL(X::I_V, m) * rl * vl; gc = gc + EL(X::I_V, m) * fi[X::I_V - 1];
      wc = wc + EL(X::I_W, m) * rl * wl; gc = gc + EL(X::I_W, m) * fi[X::I_W - 1];
      wc = wc + EL(X::I_E, m) * (rl * hl - pl); gc = gc + EL(X::I_E, m) * fi[X::I_E - 1];
      double evm = (m == 1) ? ev1 : (m == NV ? ev4 : ev2);
      double cq = 0.5 * (gc + evm * wc);
      gp[(m - 1) + NV * (l - 1)] = cq;
      gm[(m - 1) + NV * (l - 1)] = gc - cq;
    }
  }
#undef EL
  double fc[NV];
  wenorec<NS>(gp, gm, fc, ws, ord, A.weno_version, A.rho0, A.u0, false);
  // ---- back to conserved fluxes with the right eigenvectors er(field, row)
  // er(1,.) acoustic-, er(1+s,.) species, er(NS+2/NS+3,.) shear, er(NS+4,.) acoustic+
  double fsum_species = 0.0;  // unused accumulator kept explicit below
  (void)fsum_species;
#pragma unroll
  for (int m = 1; m <= NV; ++m) {
    double s = 0.0;
    // er(mm, m) for mm = 1..NV, in the reference's order
    // mm = 1
    double e1;
    if (m <= NS) e1 = yyroe[m - 1];
    else if (m == X::I_U) e1 = uu - (DIR == 1 ? c : 0.0);
    else if (m == X::I_V) e1 = vv - (DIR == 2 ? c : 0.0);
    else if (m == X::I_W) e1 = ww - (DIR == 3 ? c : 0.0);
    else e1 = h - un * c;
    s = s + e1 * fc[0];
#pragma unroll
    for (int sp = 1; sp <= NS; ++sp) {
      double e;
      if (m <= NS) e = (m == sp) ? 1.0 : 0.0;
      else if (m == X::I_U) e = uu;
      else if (m == X::I_V) e = vv;
      else if (m == X::I_W) e = ww;
      else e = b3 - prhoi[sp - 1] / b2 * cci;
      s = s + e * fc[sp];
    }
    double e2, e3, e4;
    if (m <= NS) { e2 = 0.0; e3 = 0.0; e4 = yyroe[m - 1]; }
    else if (m == X::I_E) { e2 = ta; e3 = tb; e4 = h + un * c; }
    else {
      e2 = (m == IA) ? 1.0 : 0.0;
      e3 = (m == IB) ? 1.0 : 0.0;
      e4 = (m == X::I_U ? uu : m == X::I_V ? vv : ww) + ((m == IN) ? c : 0.0);
    }
    s = s + e2 * fc[NS + 1];
    s = s + e3 * fc[NS + 2];
    s = s + e4 * fc[NS + 3];
    fh[m - 1] = s;
  }
}

template <int NS, int DIR>
__device__ inline void face_flux(const Box &b, const EulerArgs &A, const Line &L, int t1, int t2, int f,
                                 bool check_shock, double *fh) {
  constexpr int NV = NS + 4;
  if (check_shock && shocked<NS>(L, f, A.weno_scheme, A.sensor_threshold)) {
#pragma unroll
    for (int m = 0; m < NV; ++m) fh[m] = A.fhat[L.base + (long)f * L.st + b.pstride() * m];
    return;
  }
  int lmax = A.lmax_base + epo_at<DIR>(b, A.epo, t1, t2, f);
  lmax = lmax < 1 ? 1 : lmax;
  bool zero = (f == 0 && A.zmin == 1) || (f == A.n && A.zmax == 1);
  keep_flux<NS, DIR>(L, f, lmax, A.nkeep, A.cd1, zero, fh);
}


// WENO pre-pass over all faces of one direction (defined in euler.hip): shocked faces get
// their WENO flux parked in fhat.
int weno_prepass_launch(int dir, int ns, const Box &b, const EulerArgs &A, hipStream_t st);

}  // namespace streams
