// visc_dev.h — device functions of the viscous / diffusive face flux, shared by the
// per-direction viscous kernels and the fused convective+viscous sweeps.
// Arithmetic follows visflx_x_cuf / _y / _z (kernels_gpu.F90:3658-4031) and
// get_species_h_from_temperature_dev (:7144-7182).
#pragma once
#include "common.h"

namespace streams {


struct ViscArgs {
  const double *w_aux, *coord, *rgas;
  double *fl;
  Thermo th;
  double r_univ;
  int n;
};

struct VLine {
  const double *a;
  long base, st, ps;
  __device__ inline double operator()(int p, int m) const { return a[base + (long)p * st + ps * (m - 1)]; }
};

template <int DIR>
__device__ inline void vline_base(const Box &b, int t1, int t2, long &base, long &st) {
  if (DIR == 1) { base = b.p(0, t1, t2); st = 1; }
  else if (DIR == 2) { base = b.p(t1, 0, t2); st = b.px(); }
  else { base = b.p(t1, t2, 0); st = b.px() * b.py(); }
}

// Species enthalpies of the two cells of a face from the reference's table layout
// (get_species_h_from_temperature_dev per species and cell).
template <int NS>
struct EnthalpyGlobal {
  const Thermo &th;
  double ta, tc;
  __device__ inline double operator()(int side, int s) const {
    return species_hcoef<NS>(side == 0 ? ta : tc, th.cp_coeff, th, s + 1);
  }
};

template <int NS, class LineT, class HFun>
__device__ void visc_face_h(const LineT &L, int f, const ViscArgs &A, int ng, double *fh, const HFun &H);

template <int NS, class LineT>
__device__ void visc_face(const LineT &L, int f, const ViscArgs &A, int ng, double *fh) {
  using X = Idx<NS>;
  const EnthalpyGlobal<NS> H{A.th, L(f, X::J_T), L(f + 1, X::J_T)};
  visc_face_h<NS>(L, f, A, ng, fh, H);
}

// Staggered viscous / diffusive flux of face f (between cells f and f+1); H(side, s) gives
// the enthalpy of species s in the left (side 0) or right (side 1) cell.
template <int NS, class LineT, class HFun>
__device__ void visc_face_h(const LineT &L, int f, const ViscArgs &A, int ng, double *fh, const HFun &H) {
  using X = Idx<NS>;
  const int a = f, c = f + 1;
  const double uu = L(a, X::J_U), uup = L(c, X::J_U);
  const double vv = L(a, X::J_V), vvp = L(c, X::J_V);
  const double ww = L(a, X::J_W), wwp = L(c, X::J_W);
  const double tt = L(a, X::J_T), ttp = L(c, X::J_T);
  const double mu = L(a, X::J_MU), mup = L(c, X::J_MU);
  const double kc = L(a, X::J_K_COND), kcp = L(c, X::J_K_COND);
  const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
  const double qqp = 0.5 * (uup * uup + vvp * vvp + wwp * wwp);
  const double dx = A.coord[f + 1 + ng - 1] - A.coord[f + ng - 1];
  double sigx = uup - uu, sigy = vvp - vv, sigz = wwp - ww;
  const double sigq_tt = ttp - tt, sigq_qq = qqp - qq;
  double muf = mu + mup;
  muf = 0.5 * muf / dx;
  double kcf = kc + kcp;
  kcf = 0.5 * kcf / dx;
  sigx = sigx * muf;
  sigy = sigy * muf;
  sigz = sigz * muf;
  const double sigq = sigq_tt * kcf + sigq_qq * muf;
  double ya[NS], yc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) { ya[s] = L(a, s + 1); yc[s] = L(c, s + 1); }
  double rm = A.rgas[0] * ya[0];
#pragma unroll
  for (int s = 1; s < NS; ++s) rm = rm + A.rgas[s] * ya[s];
  const double mwm = A.r_univ / rm;
  rm = A.rgas[0] * yc[0];
#pragma unroll
  for (int s = 1; s < NS; ++s) rm = rm + A.rgas[s] * yc[s];
  const double mwmp = A.r_univ / rm;
  double mwmif = 0.5 * (mwmp + mwm) / (mwmp * mwm);
  mwmif = mwmif / dx;
  double rd[NS];
  double fsum = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rd[s] = 0.5 * (L(a, X::J_D_START + s + 1) + L(c, X::J_D_START + s + 1));
    fsum = fsum + (yc[s] * mwmp - ya[s] * mwm) * rd[s];
  }
  fsum = fsum * mwmif;
  double et = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double yyf = 0.5 * (yc[s] + ya[s]);
    const double sig = -rd[s] * (yc[s] * mwmp - ya[s] * mwm) * mwmif + fsum * yyf;
    fh[s] = sig;
    const double hh = H(0, s);
    const double hhp = H(1, s);
    et = et + sig * (0.5 * (hh + hhp));
  }
  fh[X::I_U - 1] = -sigx;
  fh[X::I_V - 1] = -sigy;
  fh[X::I_W - 1] = -sigz;
  fh[X::I_E - 1] = -sigq + et;
}

}  // namespace streams
