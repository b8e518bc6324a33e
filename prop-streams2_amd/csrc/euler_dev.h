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
// One field of wenorec_1d (kernels_gpu.F90:2353-2637): vp[k + 3] and vm[k + 3] hold the
// plus / minus split flux at stencil offset k (the cell f + k of face f, k = -3..4), bsc
// the field's smoothness-indicator scale.  With offsets fixed at compile time the stencil
// lives in registers.
__device__ inline double wenorec1(const double *vp, const double *vm, double bsc, int ord, int version) {
#define VP(k) vp[(k) + 3]
#define VM(k) vm[(k) + 3]
  if (ord == 1) {
    return VP(0) + VM(+1);
  } else if (ord == 2) {
    double bp0 = (VP(0) - VP(-1)) * (VP(0) - VP(-1));
    double bp1 = (VP(+1) - VP(0)) * (VP(+1) - VP(0));
    bp1 = bsc * bp1; bp0 = bsc * bp0;
    double bm0 = (VM(+2) - VM(+1)) * (VM(+2) - VM(+1));
    double bm1 = (VM(+1) - VM(0)) * (VM(+1) - VM(0));
    bm1 = bsc * bm1; bm0 = bsc * bm0;
    const double d0 = 1.0 / 3.0, d1 = 2.0 / 3.0;
    double t;
    t = 0.000001 + bp0; bp0 = d0 / (t * t);
    t = 0.000001 + bm0; bm0 = d0 / (t * t);
    t = 0.000001 + bp1; bp1 = d1 / (t * t);
    t = 0.000001 + bm1; bm1 = d1 / (t * t);
    double sump = bp0 + bp1, summ = bm0 + bm1;
    bp0 = bp0 / sump; bp1 = bp1 / sump; bm0 = bm0 / summ; bm1 = bm1 / summ;
    double vminus = bp0 * (-VP(-1) + 3 * VP(0)) + bp1 * (VP(0) + VP(+1));
    double vplus = bm0 * (-VM(+2) + 3 * VM(+1)) + bm1 * (VM(0) + VM(+1));
    return 0.5 * (vminus + vplus);
  } else if (ord == 3) {
    const double d0 = 13.0 / 12.0, d1 = 1.0 / 4.0;
    const double c0 = 1.0 / 3.0, c1 = 5.0 / 6.0, c2 = -1.0 / 6.0, c3 = -7.0 / 6.0, c4 = 11.0 / 6.0;
    const double dw0 = 1.0 / 10.0, dw1 = 6.0 / 10.0, dw2 = 3.0 / 10.0;
    double a, bq;
    a = VP(0) - 2.0 * VP(+1) + VP(+2); bq = 3.0 * VP(0) - 4.0 * VP(+1) + VP(+2);
    double bp2 = d0 * a * a + d1 * bq * bq;
    a = VP(-1) - 2.0 * VP(0) + VP(+1); bq = VP(-1) - VP(+1);
    double bp1 = d0 * a * a + d1 * bq * bq;
    a = VP(0) - 2.0 * VP(-1) + VP(-2); bq = 3.0 * VP(0) - 4.0 * VP(-1) + VP(-2);
    double bp0 = d0 * a * a + d1 * bq * bq;
    bp2 = bsc * bp2; bp1 = bsc * bp1; bp0 = bsc * bp0;
    a = VM(+1) - 2.0 * VM(0) + VM(-1); bq = 3.0 * VM(+1) - 4.0 * VM(0) + VM(-1);
    double bm2 = d0 * a * a + d1 * bq * bq;
    a = VM(+2) - 2.0 * VM(+1) + VM(0); bq = VM(+2) - VM(0);
    double bm1 = d0 * a * a + d1 * bq * bq;
    a = VM(+1) - 2.0 * VM(+2) + VM(+3); bq = 3.0 * VM(+1) - 4.0 * VM(+2) + VM(+3);
    double bm0 = d0 * a * a + d1 * bq * bq;
    bm2 = bsc * bm2; bm1 = bsc * bm1; bm0 = bsc * bm0;
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
    double vminus = bp2 * (c0 * VP(0) + c1 * VP(+1) + c2 * VP(+2)) +
                    bp1 * (c2 * VP(-1) + c1 * VP(0) + c0 * VP(+1)) +
                    bp0 * (c0 * VP(-2) + c3 * VP(-1) + c4 * VP(0));
    double vplus = bm2 * (c0 * VM(+1) + c1 * VM(0) + c2 * VM(-1)) +
                   bm1 * (c2 * VM(+2) + c1 * VM(+1) + c0 * VM(0)) +
                   bm0 * (c0 * VM(+3) + c3 * VM(+2) + c4 * VM(+1));
    return vminus + vplus;
  } else {  // ord == 4, WENO-7
    const double d1 = 1.0 / 36.0, d2 = 13.0 / 12.0, d3 = 781.0 / 720.0;
    const double dw[4] = {1.0 / 35.0, 12.0 / 35.0, 18.0 / 35.0, 4.0 / 35.0};
    double bp[4], bm[4], a, bq, c;
    a = -11 * VP(0) + 18 * VP(+1) - 9 * VP(+2) + 2 * VP(+3);
    bq = 2 * VP(0) - 5 * VP(+1) + 4 * VP(+2) - VP(+3);
    c = -VP(0) + 3 * VP(+1) - 3 * VP(+2) + VP(+3);
    bp[3] = d1 * a * a + d2 * bq * bq + d3 * c * c;
    a = -2 * VP(-1) - 3 * VP(0) + 6 * VP(+1) - VP(+2);
    bq = VP(-1) - 2 * VP(0) + VP(+1);
    c = -VP(-1) + 3 * VP(0) - 3 * VP(+1) + VP(+2);
    bp[2] = d1 * a * a + d2 * bq * bq + d3 * c * c;
    a = VP(-2) - 6 * VP(-1) + 3 * VP(0) + 2 * VP(+1);
    bq = VP(-1) - 2 * VP(0) + VP(+1);
    c = -VP(-2) + 3 * VP(-1) - 3 * VP(0) + VP(+1);
    bp[1] = d1 * a * a + d2 * bq * bq + d3 * c * c;
    a = -2 * VP(-3) + 9 * VP(-2) - 18 * VP(-1) + 11 * VP(0);
    bq = -VP(-3) + 4 * VP(-2) - 5 * VP(-1) + 2 * VP(0);
    c = -VP(-3) + 3 * VP(-2) - 3 * VP(-1) + VP(0);
    bp[0] = d1 * a * a + d2 * bq * bq + d3 * c * c;
    a = -11 * VM(+1) + 18 * VM(0) - 9 * VM(-1) + 2 * VM(-2);
    bq = 2 * VM(+1) - 5 * VM(0) + 4 * VM(-1) - VM(-2);
    c = -VM(+1) + 3 * VM(0) - 3 * VM(-1) + VM(-2);
    bm[3] = d1 * a * a + d2 * bq * bq + d3 * c * c;
    a = -2 * VM(+2) - 3 * VM(+1) + 6 * VM(0) - VM(-1);
    bq = VM(+2) - 2 * VM(+1) + VM(0);
    c = -VM(+2) + 3 * VM(+1) - 3 * VM(0) + VM(-1);
    bm[2] = d1 * a * a + d2 * bq * bq + d3 * c * c;
    a = VM(+3) - 6 * VM(+2) + 3 * VM(+1) + 2 * VM(0);
    bq = VM(+2) - 2 * VM(+1) + VM(0);
    c = -VM(+3) + 3 * VM(+2) - 3 * VM(+1) + VM(0);
    bm[1] = d1 * a * a + d2 * bq * bq + d3 * c * c;
    a = -2 * VM(+4) + 9 * VM(+3) - 18 * VM(+2) + 11 * VM(+1);
    bq = -VM(+4) + 4 * VM(+3) - 5 * VM(+2) + 2 * VM(+1);
    c = -VM(+4) + 3 * VM(+3) - 3 * VM(+2) + VM(+1);
    bm[0] = d1 * a * a + d2 * bq * bq + d3 * c * c;
    double sump = 0, summ = 0;
    for (int l = 0; l < 4; ++l) {
      bp[l] = bsc * bp[l];
      bm[l] = bsc * bm[l];
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
    double vminus = bp[3] * (6 * VP(0) + 26 * VP(+1) - 10 * VP(+2) + 2 * VP(+3)) +
                    bp[2] * (-2 * VP(-1) + 14 * VP(0) + 14 * VP(+1) - 2 * VP(+2)) +
                    bp[1] * (2 * VP(-2) - 10 * VP(-1) + 26 * VP(0) + 6 * VP(+1)) +
                    bp[0] * (-6 * VP(-3) + 26 * VP(-2) - 46 * VP(-1) + 50 * VP(0));
    double vplus = bm[3] * (6 * VM(+1) + 26 * VM(0) - 10 * VM(-1) + 2 * VM(-2)) +
                   bm[2] * (-2 * VM(+2) + 14 * VM(+1) + 14 * VM(0) - 2 * VM(-1)) +
                   bm[1] * (2 * VM(+3) - 10 * VM(+2) + 26 * VM(+1) + 6 * VM(0)) +
                   bm[0] * (-6 * VM(+4) + 26 * VM(+3) - 46 * VM(+2) + 50 * VM(+1));
    return (vminus + vplus) / 24.0;
  }
#undef VP
#undef VM
}

// wenorec over all NV fields of vp / vm (NV x 8, column-major by stencil index l = 1..8,
// cell f + l - iweno)
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
  for (int m = 0; m < NV; ++m) {
    double p8[8], m8[8];
#pragma unroll
    for (int k = -3; k <= 4; ++k) {
      const int l = iweno + k;
      p8[k + 3] = (l >= 1 && l <= 8) ? vp[m + NV * (l - 1)] : 0.0;
      m8[k + 3] = (l >= 1 && l <= 8) ? vm[m + NV * (l - 1)] : 0.0;
    }
    vhat[m] = wenorec1(p8, m8, bsc[m], ord, version);
  }
}

// WENO face flux (characteristic Roe/LLF or component-wise Rusanov); RUS selects the
// splitting at compile time, so that each kernel carries the registers of one path only
// WS > 0: weno_scheme known at compile time (the stencil slots outside it fold to zero);
// ORD > 0: the reconstruction order known as well (no order reduction at this face).
template <int NS, int DIR, bool RUS, int WS = 0, int ORD = 0>
__device__ void weno_flux(const Line &L, int f, const EulerArgs &A, int ord_rt, double *fh) {
  using X = Idx<NS>;
  constexpr int NV = NS + 4;
  constexpr int JN = DIR == 1 ? X::J_U : DIR == 2 ? X::J_V : X::J_W;
  constexpr int IN = DIR == 1 ? X::I_U : DIR == 2 ? X::I_V : X::I_W;
  const int ws = WS > 0 ? WS : A.weno_scheme, wsize = 2 * ws;
  const int ord = ORD > 0 ? ORD : ord_rt;
  if constexpr (RUS) {
    double gp[NV * 8], gm[NV * 8];
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
  // ---- characteristic projection of the split fluxes (the left eigenvectors of
  // eigenvectors_x/y/z, :2703-2984, applied through their structure): with
  // b1m_s = b1 + p_rho_s / c^2 and, per stencil cell, the sums S_b = sum_s b1m_s U_s,
  // S = sum_s U_s, M = u U_u + v U_v + w U_w,
  //   acoustic -/+:  (S_b +- un S / c) / 2 - (b2 M +- U_n / c) / 2 + b2 U_E / 2
  //   species s:     U_s + Y_s (b2 (M - U_E) - S_b)
  //   shear a / b:   U_a - ta S,  U_b - tb S
  // for the conserved variables U and, the same way, the Euler fluxes F.  The species
  // fields are projected one at a time right before their reconstruction.
  // LLF eigenvalue bounds over the stencil
  double ev1 = -1.0, ev2 = -1.0, ev4 = -1.0;
  for (int l = 1; l <= wsize; ++l) {
    int ll = f + l - ws;
    double unl = L(ll, JN), cl = L(ll, X::J_C);
    ev1 = fmax(fabs(unl - cl), ev1);
    ev2 = fmax(fabs(unl), ev2);
    ev4 = fmax(fabs(unl + cl), ev4);
  }
  double ap[8], am[8], zp[8], zm[8], sap[8], sam[8], sbp[8], sbm[8];   // acoustic -, +, shear a, b
  double psu[8], psf[8], rlv[8], ruv[8];
#pragma unroll
  for (int d = -3; d <= 4; ++d) {
    const int q = d + 3;
    ap[q] = am[q] = zp[q] = zm[q] = sap[q] = sam[q] = sbp[q] = sbm[q] = 0.0;
    psu[q] = psf[q] = rlv[q] = ruv[q] = 0.0;
    if (d < 1 - ws || d > ws) continue;
    const int ll = f + d;
    const double rl = L(ll, X::J_R), ul = L(ll, X::J_U), vl = L(ll, X::J_V), wl = L(ll, X::J_W);
    const double hl = L(ll, X::J_H), pl = L(ll, X::J_P), unl = L(ll, JN);
    const double rhou = rl * unl;
    double sbu = 0.0, su = 0.0, sbf = 0.0, sf = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double y = L(ll, s + 1);
      const double us = rl * y, fs = rhou * y;
      const double b1m = b1 + prhoi[s] * cci;
      sbu = sbu + b1m * us;
      su = su + us;
      sbf = sbf + b1m * fs;
      sf = sf + fs;
    }
    const double Uu = rl * ul, Uv = rl * vl, Uw = rl * wl, Ue = rl * hl - pl;
    double Fu = ul * rhou, Fv = vl * rhou, Fw = wl * rhou;
    const double Fe = hl * rhou;
    if (IN == X::I_U) Fu = Fu + pl;
    else if (IN == X::I_V) Fv = Fv + pl;
    else Fw = Fw + pl;
    const double mu_ = uu * Uu + vv * Uv + ww * Uw, mf = uu * Fu + vv * Fv + ww * Fw;
    const double Un = DIR == 1 ? Uu : DIR == 2 ? Uv : Uw, Fn = DIR == 1 ? Fu : DIR == 2 ? Fv : Fw;
    const double Ua = IA == X::I_U ? Uu : Uv, Fa = IA == X::I_U ? Fu : Fv;
    const double Ub = IB == X::I_V ? Uv : Uw, Fb = IB == X::I_V ? Fv : Fw;
    const double unci = un * ci;
    const double w1 = 0.5 * (sbu + unci * su) - 0.5 * (b2 * mu_ + ci * Un) + 0.5 * b2 * Ue;
    const double g1 = 0.5 * (sbf + unci * sf) - 0.5 * (b2 * mf + ci * Fn) + 0.5 * b2 * Fe;
    const double w4 = 0.5 * (sbu - unci * su) - 0.5 * (b2 * mu_ - ci * Un) + 0.5 * b2 * Ue;
    const double g4 = 0.5 * (sbf - unci * sf) - 0.5 * (b2 * mf - ci * Fn) + 0.5 * b2 * Fe;
    const double wa = Ua - ta * su, ga = Fa - ta * sf;
    const double wb = Ub - tb * su, gb = Fb - tb * sf;
    double cq;
    cq = 0.5 * (g1 + ev1 * w1); ap[q] = cq; am[q] = g1 - cq;
    cq = 0.5 * (g4 + ev4 * w4); zp[q] = cq; zm[q] = g4 - cq;
    cq = 0.5 * (ga + ev2 * wa); sap[q] = cq; sam[q] = ga - cq;
    cq = 0.5 * (gb + ev2 * wb); sbp[q] = cq; sbm[q] = gb - cq;
    psu[q] = b2 * (mu_ - Ue) - sbu;
    psf[q] = b2 * (mf - Fe) - sbf;
    rlv[q] = rl;
    ruv[q] = rhou;
  }
  // reconstruction field by field (wenorec's smoothness scales: 1/(rho0 u0)^2 for the
  // acoustic and species fields, 1/(rho0^2 u0^4) for the shear fields)
  const double u0_2 = A.u0 * A.u0, r2u2 = A.rho0 * A.rho0 * u0_2, r2u4 = r2u2 * u0_2;
  const double bsc_a = 1.0 / r2u2, bsc_t = 1.0 / r2u4;
  double fc[NV];
  fc[0] = wenorec1(ap, am, bsc_a, ord, A.weno_version);
  fc[NS + 1] = wenorec1(sap, sam, bsc_t, ord, A.weno_version);
  fc[NS + 2] = wenorec1(sbp, sbm, bsc_t, ord, A.weno_version);
  fc[NS + 3] = wenorec1(zp, zm, bsc_a, ord, A.weno_version);
#pragma unroll 1
  for (int s = 0; s < NS; ++s) {
    double p8[8], m8[8];
    const double ys = yyroe[s];
#pragma unroll
    for (int d = -3; d <= 4; ++d) {
      const int q = d + 3;
      p8[q] = m8[q] = 0.0;
      if (d < 1 - ws || d > ws) continue;
      const double y = L(f + d, s + 1);
      const double w = rlv[q] * y + ys * psu[q], g = ruv[q] * y + ys * psf[q];
      const double cq = 0.5 * (g + ev2 * w);
      p8[q] = cq;
      m8[q] = g - cq;
    }
    fc[1 + s] = wenorec1(p8, m8, bsc_a, ord, A.weno_version);
  }
  // ---- back to conserved fluxes with the right eigenvectors er(field, row)
  // er(1,.) acoustic-, er(1+s,.) species, er(NS+2/NS+3,.) shear, er(NS+4,.) acoustic+
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
