// bc_nr.hip — non-reflecting, relaxation and reflective-wall treatment of a boundary face
// (bc_nr_lat_{x,y,z}_kernel, kernels_gpu.F90:4330-4793), one thread per boundary cell.
//
// The face cells are left out of the convective update (bc_preproc shifts eul_imin etc.);
// this kernel supplies their convective term from characteristics: a one-sided
// second-order derivative of the conserved variables normal to the face, projected on the
// left eigenvectors of the cell (the Roe-form matrices of eigenvectors_{x,y,z}), the
// incoming waves removed (nr_type 1), replaced by the difference to the ghost cell (2) or to
// the freestream state winf (3), or reflected (6), multiplied by the eigenvalues and
// projected back: fl += R (lambda L dw/dn) * metric.  The eigenvector matrices are private
// per thread; this is boundary work (one face), not worth tiling.
#include "common.h"

namespace streams {

template <int NS, int DIR>
__global__ void __launch_bounds__(128) bc_nr_kernel(Box b, int start_or_end, int nr_type,
                                                    const double *__restrict__ w_aux, const double *__restrict__ w,
                                                    double *__restrict__ fl, const double *__restrict__ metric,
                                                    const double *__restrict__ rgas, Thermo th,
                                                    const double *__restrict__ winf) {
  using X = Idx<NS>;
  constexpr int NV = NS + 4;
  const int n1 = DIR == 1 ? b.ny : b.nx, n2 = DIR == 3 ? b.ny : b.nz;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n1 * n2) return;
  const int c1 = 1 + (int)(t % n1), c2 = 1 + (int)(t / n1);
  const int sgn = start_or_end == 1 ? 1 : -1;
  const int nb = DIR == 1 ? b.nx : DIR == 2 ? b.ny : b.nz;
  const int bnd = start_or_end == 1 ? 1 : nb;
  int i, j, k;
  if (DIR == 1) { i = bnd; j = c1; k = c2; }
  else if (DIR == 2) { i = c1; j = bnd; k = c2; }
  else { i = c1; j = c2; k = bnd; }
  const long ps = b.pstride();
  const long pt = b.p(i, j, k);
  const long step = DIR == 1 ? 1 : DIR == 2 ? b.px() : b.px() * b.py();
  const double c_one[3] = {-1.5, 2.0, -0.5};
  double dw_dn[NV], dw_out[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    double d = 0.0;
#pragma unroll
    for (int l = 0; l < 3; ++l) d = d + sgn * c_one[l] * w[pt + sgn * l * step + ps * m];
    dw_dn[m] = d;
    double tgt = 0.0;
    if (nr_type == 2) tgt = w[pt - sgn * step + ps * m];
    else if (nr_type == 3) tgt = winf[m];
    dw_out[m] = sgn * (w[pt + ps * m] - tgt);
  }
  double yy[NS], prhoi[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) yy[s] = w_aux[pt + ps * s];
  const double rho = w_aux[pt + ps * (X::J_R - 1)], uu = w_aux[pt + ps * (X::J_U - 1)];
  const double vv = w_aux[pt + ps * (X::J_V - 1)], ww = w_aux[pt + ps * (X::J_W - 1)];
  const double h = w_aux[pt + ps * (X::J_H - 1)], tt = w_aux[pt + ps * (X::J_T - 1)];
  const double pp = w_aux[pt + ps * (X::J_P - 1)], c = w_aux[pt + ps * (X::J_C - 1)];
  const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
  const double cc = c * c, rmix_tt = pp / rho, gamloc = cc / rmix_tt, gm1 = gamloc - 1.0, ci = 1.0 / c;
  const double p_e = rho * gm1, etot = h - rmix_tt;
  const double b3 = etot, b2 = p_e / (rho * cc), b1 = -b2 * (etot - 2.0 * qq);
#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    const double eei = species_hcoef<NS>(tt, th.cv_coeff, th, s);
    prhoi[s - 1] = rgas[s - 1] * tt + gm1 * (etot - qq - eei);
  }
  // eigenvectors_{x,y,z}: el(row = conserved variable, column = field), er(field, variable)
  double el[NV * NV], er[NV * NV];
#pragma unroll
  for (int q = 0; q < NV * NV; ++q) { el[q] = 0.0; er[q] = 0.0; }
#define EL(r, q) el[((r) - 1) + NV * ((q) - 1)]
#define ER(r, q) er[((r) - 1) + NV * ((q) - 1)]
  const double cci = ci * ci;
  const double un = DIR == 1 ? uu : DIR == 2 ? vv : ww;
  const double ta = DIR == 1 ? vv : uu, tb = DIR == 3 ? vv : ww;
  constexpr int IA = DIR == 1 ? X::I_V : X::I_U;
  constexpr int IB = DIR == 3 ? X::I_V : X::I_W;
#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    const double b1i = b1 + prhoi[s - 1] * cci;
    EL(s, 1) = 0.5 * (b1i + un * ci);
    EL(s, NS + 2) = -ta;
    EL(s, NS + 3) = -tb;
    EL(s, NS + 4) = 0.5 * (b1i - un * ci);
    EL(s, s + 1) = 1.0;
    EL(X::I_U, 1 + s) = b2 * uu * yy[s - 1];
    EL(X::I_V, 1 + s) = b2 * vv * yy[s - 1];
    EL(X::I_W, 1 + s) = b2 * ww * yy[s - 1];
    EL(X::I_E, 1 + s) = -b2 * yy[s - 1];
#pragma unroll
    for (int m = 1; m <= NS; ++m) EL(m, 1 + s) = EL(m, 1 + s) - (b1 + prhoi[m - 1] * cci) * yy[s - 1];
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
#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    ER(1, s) = yy[s - 1];
    ER(1 + s, s) = 1.0;
    ER(NS + 4, s) = yy[s - 1];
    ER(1 + s, X::I_U) = uu;
    ER(1 + s, X::I_V) = vv;
    ER(1 + s, X::I_W) = ww;
    ER(1 + s, X::I_E) = b3 - prhoi[s - 1] / b2 * cci;
  }
  ER(1, X::I_U) = uu - (DIR == 1 ? c : 0.0);
  ER(1, X::I_V) = vv - (DIR == 2 ? c : 0.0);
  ER(1, X::I_W) = ww - (DIR == 3 ? c : 0.0);
  ER(1, X::I_E) = h - un * c;
  ER(NS + 2, IA) = 1.0;
  ER(NS + 3, IB) = 1.0;
  ER(NS + 4, X::I_U) = uu + (DIR == 1 ? c : 0.0);
  ER(NS + 4, X::I_V) = vv + (DIR == 2 ? c : 0.0);
  ER(NS + 4, X::I_W) = ww + (DIR == 3 ? c : 0.0);
  ER(NS + 2, X::I_E) = ta;
  ER(NS + 3, X::I_E) = tb;
  ER(NS + 4, X::I_E) = h + un * c;
  double wc[NV], wco[NV], ev[NV];
#pragma unroll
  for (int m = 1; m <= NV; ++m) {
    double a = 0.0, o = 0.0;
#pragma unroll
    for (int mm = 1; mm <= NV; ++mm) {
      a = a + EL(mm, m) * dw_dn[mm - 1];
      o = o + EL(mm, m) * dw_out[mm - 1];
    }
    wc[m - 1] = a;
    wco[m - 1] = o;
  }
  ev[0] = un - c;
#pragma unroll
  for (int s = 2; s <= NS + 3; ++s) ev[s - 1] = un;
  ev[NV - 1] = un + c;
  if (nr_type == 1) {
#pragma unroll
    for (int l = 0; l < NV; ++l) ev[l] = sgn * fmin(sgn * ev[l], 0.0);
  }
  if (nr_type == 2 || nr_type == 3) {
#pragma unroll
    for (int m = 0; m < NV; ++m)
      if (sgn * ev[m] > 0.0) wc[m] = wco[m];
  }
#pragma unroll
  for (int m = 0; m < NV; ++m) wc[m] = ev[m] * wc[m];
  if (nr_type == 6) {
#pragma unroll
    for (int s = 2; s <= NS + 3; ++s) wc[s - 1] = 0.0;
    if (start_or_end == 1) wc[NV - 1] = wc[0];
    else wc[0] = wc[NV - 1];
  }
  const double met = metric[(DIR == 1 ? i : DIR == 2 ? j : k) - 1];
  const long fidx = b.f(i, j, k), fs = b.fstride();
#pragma unroll
  for (int m = 1; m <= NV; ++m) {
    double df = 0.0;
#pragma unroll
    for (int mm = 1; mm <= NV; ++mm) df = df + ER(mm, m) * wc[mm - 1];
    fl[fidx + fs * (m - 1)] = fl[fidx + fs * (m - 1)] + df * met;
  }
#undef EL
#undef ER
}

template <int NS>
int bc_nr_launch(int dir, const Box &b, int soe, int nr_type, const double *w_aux, const double *w, double *fl,
                 const double *metric, const double *rgas, const Thermo &th, const double *winf, hipStream_t st) {
  const long n = dir == 1 ? (long)b.ny * b.nz : dir == 2 ? (long)b.nx * b.nz : (long)b.nx * b.ny;
  if (n == 0) return 0;
  const unsigned g = blocks_for(n, 128);
  if (dir == 1) bc_nr_kernel<NS, 1><<<g, 128, 0, st>>>(b, soe, nr_type, w_aux, w, fl, metric, rgas, th, winf);
  else if (dir == 2) bc_nr_kernel<NS, 2><<<g, 128, 0, st>>>(b, soe, nr_type, w_aux, w, fl, metric, rgas, th, winf);
  else bc_nr_kernel<NS, 3><<<g, 128, 0, st>>>(b, soe, nr_type, w_aux, w, fl, metric, rgas, th, winf);
  return hip_status();
}

}  // namespace streams

using namespace streams;

extern "C" int streams_bc_nr_lat(int dir, int start_or_end, int nr_type, int n_s, int nx, int ny, int nz, int ng,
                                 int nv, const double *w_aux, const double *w, double *fl, const double *metric,
                                 int indx_cp_l, int indx_cp_r, const double *rgas, const double *cv_coeff,
                                 const double *cp_coeff, int nsetcv, const double *trange, const double *winf,
                                 void *stream) {
  if (dir < 1 || dir > 3 || (start_or_end != 1 && start_or_end != 2)) return STREAMS_E_DIMS;
  if (nv != n_s + 4 || ng < 1) return STREAMS_E_DIMS;
  if (nr_type != 1 && nr_type != 2 && nr_type != 3 && nr_type != 6) return STREAMS_E_OPTION;
  if (nr_type == 3 && winf == nullptr) return STREAMS_E_DIMS;
  const int n = dir == 1 ? nx : dir == 2 ? ny : nz;
  if (n < 3) return STREAMS_E_DIMS;                 // the one-sided derivative takes three cells
  Box b{nx, ny, nz, ng};
  Thermo th{cv_coeff, cp_coeff, trange, nsetcv, indx_cp_l, indx_cp_r};
  STREAMS_DISPATCH_NS(n_s, return (bc_nr_launch<NS>(dir, b, start_or_end, nr_type, w_aux, w, fl, metric, rgas, th,
                                                     winf, as_stream(stream))));
  return 0;
}
