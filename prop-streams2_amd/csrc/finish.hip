// finish.hip — the pointwise tail of an RK substep in one pass over the interior:
//   visflx_reduced_cuf (kernels_gpu.F90:3539-3656)   fl(mom, E) -= reduced stress terms
//   compute_chemistry_cuf (:9503-9674)               fl(species) -= omega_s W_s
//   update_flux_cuf (:3297-3314)                     fln -= gamdt * fl
//   update_field_cuf (:3336-3355)                    w += fln (fluid cells)
// fl, fln and w are each read and written once.  The chemistry takes rho, Y and T from
// w_aux as written by eval_aux in the same substep: w is unchanged since then, so rho and
// Y are the identical numbers compute_chemistry_cuf recomputes; its extra Newton iteration
// moves T by less than the 1e-12 K stopping tolerance.
#include "chem_common.h"

namespace streams {

struct FinishArgs {
  double *w;
  const double *w_aux;
  double *fl, *fln;
  const double *cd1, *dcsidx, *detady, *dzitdz;
  const int *mask;
  Kin K;
  int lmax, reduced, chem;
  double gamdt;
};

template <int NS>
__global__ void __launch_bounds__(256) finish_kernel(Box b, FinishArgs F) {
  using X = Idx<NS>;
  constexpr int NV = NS + 4;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.fstride()) return;
  const int i = 1 + (int)(t % b.nx);
  const long r = t / b.nx;
  const int j = 1 + (int)(r % b.ny);
  const int k = 1 + (int)(r / b.ny);
  const long pt = b.p(i, j, k), ps = b.pstride(), fs_ = b.fstride();
  const double *wa = F.w_aux;
  double acc[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) acc[m] = F.fl[t + fs_ * m];
  if (F.reduced) {
    const long sx = 1, sy = b.px(), sz = b.px() * b.py();
    const double *U = wa + ps * (X::J_U - 1), *V = wa + ps * (X::J_V - 1), *Wv = wa + ps * (X::J_W - 1);
    const double *MU = wa + ps * (X::J_MU - 1), *DV = wa + ps * (X::J_DIV - 1);
    const double uu = U[pt], vv = V[pt], ww = Wv[pt], mu = MU[pt], div3l = DV[pt];
    double ux = 0, vx = 0, wx = 0, mux = 0, dvx = 0, uy = 0, vy = 0, wy = 0, muy = 0, dvy = 0;
    double uz = 0, vz = 0, wz = 0, muz = 0, dvz = 0;
    const int lmax = F.lmax;
    for (int l = 1; l <= lmax; ++l) {
      const double ccl = F.cd1[(l - 1) + 4 * (lmax - 1)];
      ux = ux + ccl * (U[pt + l * sx] - U[pt - l * sx]);
      vx = vx + ccl * (V[pt + l * sx] - V[pt - l * sx]);
      wx = wx + ccl * (Wv[pt + l * sx] - Wv[pt - l * sx]);
      mux = mux + ccl * (MU[pt + l * sx] - MU[pt - l * sx]);
      dvx = dvx + ccl * (DV[pt + l * sx] - DV[pt - l * sx]);
      uy = uy + ccl * (U[pt + l * sy] - U[pt - l * sy]);
      vy = vy + ccl * (V[pt + l * sy] - V[pt - l * sy]);
      wy = wy + ccl * (Wv[pt + l * sy] - Wv[pt - l * sy]);
      muy = muy + ccl * (MU[pt + l * sy] - MU[pt - l * sy]);
      dvy = dvy + ccl * (DV[pt + l * sy] - DV[pt - l * sy]);
      uz = uz + ccl * (U[pt + l * sz] - U[pt - l * sz]);
      vz = vz + ccl * (V[pt + l * sz] - V[pt - l * sz]);
      wz = wz + ccl * (Wv[pt + l * sz] - Wv[pt - l * sz]);
      muz = muz + ccl * (MU[pt + l * sz] - MU[pt - l * sz]);
      dvz = dvz + ccl * (DV[pt + l * sz] - DV[pt - l * sz]);
    }
    const double dx = F.dcsidx[i - 1], dy = F.detady[j - 1], dz = F.dzitdz[k - 1];
    ux *= dx; vx *= dx; wx *= dx; mux *= dx; dvx *= dx;
    uy *= dy; vy *= dy; wy *= dy; muy *= dy; dvy *= dy;
    uz *= dz; vz *= dz; wz *= dz; muz *= dz; dvz *= dz;
    const double s11 = ux - 2.0 * div3l, s12 = vx, s13 = wx;
    const double s21 = uy, s22 = vy - 2.0 * div3l, s23 = wy;
    const double s31 = uz, s32 = vz, s33 = wz - 2.0 * div3l;
    const double sigx = mux * s11 + muy * s12 + muz * s13 + mu * dvx;
    const double sigy = mux * s21 + muy * s22 + muz * s23 + mu * dvy;
    const double sigz = mux * s31 + muy * s32 + muz * s33 + mu * dvz;
    const double sigah = (s11 * ux + s12 * uy + s13 * uz + s21 * vx + s22 * vy + s23 * vz + s31 * wx + s32 * wy +
                          s33 * wz) * mu;
    const double sigq = sigx * uu + sigy * vv + sigz * ww + sigah;
    acc[X::I_U - 1] = acc[X::I_U - 1] - sigx;
    acc[X::I_V - 1] = acc[X::I_V - 1] - sigy;
    acc[X::I_W - 1] = acc[X::I_W - 1] - sigz;
    acc[X::I_E - 1] = acc[X::I_E - 1] - sigq;
  }
  if (F.chem) {
    const double rho = wa[pt + ps * (X::J_R - 1)];
    const double tt = wa[pt + ps * (X::J_T - 1)];
    double conc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) conc[s] = rho * wa[pt + ps * s] / F.K.mw[s];
    chemistry_source<NS>(F.K, tt, conc, acc);
  }
  const bool fluid = F.mask == nullptr || F.mask[pt] == 0;
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    F.fl[t + fs_ * m] = acc[m];
    const double fn = F.fln[t + fs_ * m] - F.gamdt * acc[m];
    F.fln[t + fs_ * m] = fn;
    if (fluid) F.w[pt + ps * m] = F.w[pt + ps * m] + fn;
  }
}

}  // namespace streams

using namespace streams;

// Tail of one RK substep: reduced viscous terms (reduced=1), explicit chemistry (chem=1),
// update_flux and update_field, in one pass.
extern "C" int streams_rhs_finish(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order, double *w,
                                  const double *w_aux, double *fl, double *fln, const double *coeff_deriv1,
                                  const double *dcsidx, const double *detady, const double *dzitdz, int reduced,
                                  int chem, int nreactions, const double *mw, const double *arr_a,
                                  const double *arr_b, const double *arr_ea, const double *tb_eff,
                                  const double *falloff_coeffs, const int *reac_ty, const int *isrev,
                                  const double *r_coeffs, const double *p_coeffs, const double *kc_tab,
                                  int num_t_tab, double t_min_tab, double dt_tab, double r_univ, double gamdt,
                                  const int *fluid_mask, void *stream) {
  if (nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  const int lmax = visc_order / 2;
  if (reduced && (lmax < 1 || lmax > 4 || lmax > ng)) return STREAMS_E_OPTION;
  Box b{nx, ny, nz, ng};
  const long n = b.fstride();
  if (n == 0) return 0;
  FinishArgs F{w, w_aux, fl, fln, coeff_deriv1, dcsidx, detady, dzitdz, fluid_mask,
               Kin{mw, arr_a, arr_b, arr_ea, tb_eff, falloff_coeffs, r_coeffs, p_coeffs, kc_tab, reac_ty, isrev,
                   nreactions, num_t_tab, t_min_tab, dt_tab, r_univ},
               lmax, reduced, chem && nreactions > 0, gamdt};
  STREAMS_DISPATCH_NS(n_s, (finish_kernel<NS><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(b, F)));
  return hip_status();
}
