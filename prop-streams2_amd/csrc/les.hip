// les.hip — the LES auxiliaries of compute_aux_les (multideal_gpu.F90:445-498) and the
// output-time chemistry diagnostics (compute_chem_aux, :500-534).
//
// eval_velaux: one thread per point of the padded box (the box is one contiguous plane
// per variable, so the point index is the flat index).  les_wale_mut: one thread per
// interior point, i fastest; the 3 x lmax x 2 neighbour loads of each velocity component
// hit the L2 that the previous rows of the wave brought in.  eval_chem_aux: one thread per
// interior point, the reaction loop of the right-hand side over the state in w_aux.
#include "chem_common.h"

namespace streams {

template <int NS>
__global__ void __launch_bounds__(256) velaux_kernel(long npts, long ps, double *__restrict__ w,
                                                     double *__restrict__ w_aux) {
  using X = Idx<NS>;
  const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= npts) return;
  double yy[NS];
  double rho = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double v = w[pt + ps * s];
    if (!(v > 0.0)) {
      v = 0.0;
      w[pt + ps * s] = 0.0;
    }
    yy[s] = v;
    rho = (s == 0) ? v : rho + v;
  }
  const double ri = 1.0 / rho;
#pragma unroll
  for (int s = 0; s < NS; ++s) w_aux[pt + ps * s] = yy[s] * ri;
  w_aux[pt + ps * (X::J_R - 1)] = rho;
  w_aux[pt + ps * (X::J_U - 1)] = w[pt + ps * (X::I_U - 1)] * ri;
  w_aux[pt + ps * (X::J_V - 1)] = w[pt + ps * (X::I_V - 1)] * ri;
  w_aux[pt + ps * (X::J_W - 1)] = w[pt + ps * (X::I_W - 1)] * ri;
  w_aux[pt + ps * (X::J_LES1 - 1)] = 0.0;
}

// les_wale_mut_cuf (kernels_gpu.F90:9921-10047)
template <int NS>
__global__ void __launch_bounds__(256) wale_kernel(Box b, int lmax, double *__restrict__ w_aux,
                                                   const double *__restrict__ cd1, const double *__restrict__ dcsidx,
                                                   const double *__restrict__ detady,
                                                   const double *__restrict__ dzitdz, double c_wale,
                                                   double eps_sensor, int sensor_type) {
  using X = Idx<NS>;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.fstride()) return;
  const int i = 1 + (int)(t % b.nx);
  const long r = t / b.nx;
  const int j = 1 + (int)(r % b.ny);
  const int k = 1 + (int)(r / b.ny);
  const long pt = b.p(i, j, k), ps = b.pstride();
  const long sy = b.px(), sz = b.px() * b.py();
  const double *U = w_aux + ps * (X::J_U - 1), *V = w_aux + ps * (X::J_V - 1), *Wv = w_aux + ps * (X::J_W - 1);
  const double dx = dcsidx[i - 1], dy = detady[j - 1], dz = dzitdz[k - 1];
  const double delta2 = pow(dx * dy * dz, -2.0 / 3.0);
  double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0, wx = 0, wy = 0, wz = 0;
  for (int l = 1; l <= lmax; ++l) {
    const double ccl = cd1[(l - 1) + 4 * (lmax - 1)];
    ux = ux + ccl * (U[pt + l] - U[pt - l]);
    uy = uy + ccl * (U[pt + l * sy] - U[pt - l * sy]);
    uz = uz + ccl * (U[pt + l * sz] - U[pt - l * sz]);
    vx = vx + ccl * (V[pt + l] - V[pt - l]);
    vy = vy + ccl * (V[pt + l * sy] - V[pt - l * sy]);
    vz = vz + ccl * (V[pt + l * sz] - V[pt - l * sz]);
    wx = wx + ccl * (Wv[pt + l] - Wv[pt - l]);
    wy = wy + ccl * (Wv[pt + l * sy] - Wv[pt - l * sy]);
    wz = wz + ccl * (Wv[pt + l * sz] - Wv[pt - l * sz]);
  }
  ux *= dx; uy *= dy; uz *= dz;
  vx *= dx; vy *= dy; vz *= dz;
  wx *= dx; wy *= dy; wz *= dz;
  const double div = ux + vy + wz;
  const double ox = wy - vz, oy = uz - wx, oz = vx - uy;
  const double omod2 = ox * ox + oy * oy + oz * oz;
  double sen = -div / sqrt(omod2 + div * div + eps_sensor);
  if (sensor_type != 0 && sen < 0.0) sen = 0.0;
  w_aux[pt + ps * (X::J_DUC - 1)] = sen * sen;
  // strain rate S and the traceless symmetric part S^d of the squared velocity gradient
  const double sxx = ux, sxy = 0.5 * (uy + vx), sxz = 0.5 * (uz + wx);
  const double syy = vy, syz = 0.5 * (vz + wy), szz = wz;
  const double gxx = ux * ux + uy * vx + uz * wx, gxy = ux * uy + uy * vy + uz * wy, gxz = ux * uz + uy * vz + uz * wz;
  const double gyx = vx * ux + vy * vx + vz * wx, gyy = vx * uy + vy * vy + vz * wy, gyz = vx * uz + vy * vz + vz * wz;
  const double gzx = wx * ux + wy * vx + wz * wx, gzy = wx * uy + wy * vy + wz * wy, gzz = wx * uz + wy * vz + wz * wz;
  const double tce = (gxx + gyy + gzz) / 3.0;
  const double dxx = gxx - tce, dxy = 0.5 * (gxy + gyx), dxz = 0.5 * (gxz + gzx);
  const double dyy = gyy - tce, dyz = 0.5 * (gyz + gzy), dzz = gzz - tce;
  const double nS = sqrt(sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz));
  const double nSd = sqrt(dxx * dxx + dyy * dyy + dzz * dzz + 2.0 * (dxy * dxy + dxz * dxz + dyz * dyz));
  const double rho = w_aux[pt + ps * (X::J_R - 1)];
  const double nS2 = nS * nS, nSd2 = nSd * nSd;
  const double mu_sgs =
      rho * c_wale * c_wale * delta2 * (nSd2 * nSd) / (nS2 * nS2 * nS + sqrt(nSd2 * nSd2 * nSd) + DBL_MIN);
  w_aux[pt + ps * (X::J_LES1 - 1)] = mu_sgs;
  w_aux[pt + ps * (X::J_DIV - 1)] = div / 3.0;
}

// eval_chem_aux_cuf (kernels_gpu.F90:9676-9877)
template <int NS>
__global__ void __launch_bounds__(128) chem_aux_kernel(Box b, double *__restrict__ w_aux, Kin K,
                                                       const double *__restrict__ h298, Pasr pasr) {
  using X = Idx<NS>;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.fstride()) return;
  const int i = 1 + (int)(t % b.nx);
  const long r = t / b.nx;
  const int j = 1 + (int)(r % b.ny);
  const int k = 1 + (int)(r / b.ny);
  const long pt = b.p(i, j, k), ps = b.pstride();
  const double rho = w_aux[pt + ps * (X::J_R - 1)], ri = 1.0 / rho;
  const double tt = w_aux[pt + ps * (X::J_T - 1)];
  const double Rtti = 1.0 / (K.r_univ * tt);
  double yy[NS], conc[NS], wdot[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    yy[s] = w_aux[pt + ps * s];
    conc[s] = rho * yy[s] / K.mw[s];
    wdot[s] = 0.0;
  }
  const double tau_k = pasr.on ? pasr.tau_k<NS>(w_aux, pt, ps, i, j, k) : 0.0;
  int itt = (int)((tt - K.t_min_tab) / K.dt_tab) + 1;
  itt = itt < 1 ? 1 : itt;
  itt = itt > K.num_t_tab ? K.num_t_tab : itt;
  const double ttleft = (itt - 1) * K.dt_tab + K.t_min_tab;
  const double dtt = (tt - ttleft) / K.dt_tab;
  for (int lr = 0; lr < K.nr; ++lr) {
    const double q = progress_rate<NS>(K, lr, tt, Rtti, itt, dtt, conc);
#pragma unroll
    for (int s = 0; s < NS; ++s) wdot[s] = wdot[s] + (K.p_co[lr + K.nr * s] - K.r_co[lr + K.nr * s]) * q * K.mw[s];
  }
  // the PaSR factor goes to J_GAM_PASR in the reference, the slot that the heat release
  // rate then overwrites: only the scaled rates remain
  if (pasr.on && tau_k > DBL_MIN) pasr_scale<NS>(tau_k, ri, yy, wdot);
  double hrr = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    w_aux[pt + ps * (X::J_WDOT_START + s)] = wdot[s];
    hrr = hrr - h298[s] * wdot[s] / K.mw[s];
  }
  w_aux[pt + ps * (X::J_HRR - 1)] = hrr;
}

}  // namespace streams

using namespace streams;

extern "C" int streams_eval_velaux(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux, double *w,
                                   double *w_aux, void *stream) {
  if (nv != n_s + 4 || nv_aux < 2 * n_s + 14) return STREAMS_E_DIMS;
  if (nx < 1 || ny < 1 || nz < 1 || ng < 0) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  const long n = b.pstride();
  STREAMS_DISPATCH_NS(n_s, (velaux_kernel<NS><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(n, n, w, w_aux)));
  return hip_status();
}

extern "C" int streams_les_wale_mut(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int ep_order,
                                    double *w_aux, const double *coeff_deriv1, const double *dcsidx,
                                    const double *detady, const double *dzitdz, double les_c_wale,
                                    double eps_sensor, int sensor_type, void *stream) {
  if (nv_aux < 2 * n_s + 14) return STREAMS_E_DIMS;
  const int lmax = ep_order / 2;
  if (lmax < 1 || lmax > 4 || lmax > ng) return STREAMS_E_OPTION;
  if (nx < 1 || ny < 1 || nz < 1) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  const long n = b.fstride();
  STREAMS_DISPATCH_NS(n_s, (wale_kernel<NS><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
                                b, lmax, w_aux, coeff_deriv1, dcsidx, detady, dzitdz, les_c_wale, eps_sensor,
                                sensor_type)));
  return hip_status();
}

extern "C" int streams_eval_chem_aux(int n_s, int nx, int ny, int nz, int nv_aux, int ng, int nreactions,
                                     double *w_aux, const double *mw, const double *arr_a, const double *arr_b,
                                     const double *arr_ea, const double *tb_eff, const double *falloff_coeffs,
                                     const int *reac_ty, const int *isrev, const double *r_coeffs,
                                     const double *p_coeffs, const double *kc_tab, int num_t_tab, double t_min_tab,
                                     double dt_tab, double r_univ, const double *dcsidx, const double *detady,
                                     const double *dzitdz, const double *h298, int enable_pasr,
                                     double les_c_yoshi, double les_c_mix, double les_c_eps, void *stream) {
  if (nv_aux < 3 * n_s + 15 || nreactions < 0 || num_t_tab < 1) return STREAMS_E_DIMS;
  if (enable_pasr > 0 && (nv_aux < 2 * n_s + 14 || !dcsidx || !detady || !dzitdz)) return STREAMS_E_DIMS;
  if (nx < 1 || ny < 1 || nz < 1 || ng < 0) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  const long n = b.fstride();
  Kin K{mw, arr_a, arr_b, arr_ea, tb_eff, falloff_coeffs, r_coeffs, p_coeffs, kc_tab, reac_ty, isrev,
        nreactions, num_t_tab, t_min_tab, dt_tab, r_univ};
  const Pasr pasr{enable_pasr > 0, les_c_yoshi, les_c_mix, les_c_eps, dcsidx, detady, dzitdz};
  STREAMS_DISPATCH_NS(n_s, (chem_aux_kernel<NS><<<blocks_for(n, 128), 128, 0, as_stream(stream)>>>(
                                b, w_aux, K, h298, pasr)));
  return hip_status();
}
