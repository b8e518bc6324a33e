// aux.hip — auxiliary variables: thermo state, mixture transport, shock sensor.
//
// eval_aux: one thread per point of the requested box; the box is flattened with i
// fastest so that every load and store of a wave touches 64 consecutive doubles of one
// variable plane (coalesced 512-byte segments).  Pairwise Wilke constants that depend on
// the molar masses only are built once per workgroup in LDS.
#include "thermo_dev.h"

namespace streams {

struct AuxTables {
  const double *rgas, *mw, *mwinv, *visc, *lambda, *diffbin;
  int num_t_tab;
  double t_min_tab, dt_tab, p0, r_univ;
  int n_eoi;
  const double *aw_eoi, *beta0, *coeff_eoi;
  const int *nainsp;
};

// The mixture-rule divisions of the transport kernel use rcp_nr (reciprocal estimate + two
// Newton steps): the transport coefficients then differ from the reference's in the last
// bits only.
// The thermodynamic state (eval_aux_cuf up to the mixture fraction, kernels_gpu.F90:6154-6251)
// and the mixture transport (:6252-6300) run as two kernels over the same points: apart,
// each holds far fewer registers than the pair would, for the price of re-reading Y, T, p
// and rho (12 of the ~45 doubles a point moves).
// LES = true is eval_aux_les_cuf (kernels_gpu.F90:10049-10161): rho, u, v, w and Y come
// from w_aux (left there by eval_velaux), nothing of them is written back, the mixture
// fraction is not touched, and c_p is left in the J_K_COND slot for the transport kernel,
// which overwrites it with the conductivity.
template <int NS, int LO, bool LES = false>
__global__ void __launch_bounds__(256) eval_aux_kernel(Box b, int istart, int jstart, int kstart,
                                                       int ni, int nj, long npts, double *__restrict__ w,
                                                       double *__restrict__ w_aux, Thermo th,
                                                       AuxTables tb, double tol_iter_nr) {
  using X = Idx<NS>;
  __shared__ double tlds[thermo_lds_doubles<NS>(TH_CP | TH_CV | TH_CVH)];
  const ThermoLds TL = thermo_stage<NS, LO>(th, tlds, TH_CP | TH_CV | TH_CVH);
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  int i = istart + (int)(t % ni);
  long r = t / ni;
  int j = jstart + (int)(r % nj);
  int k = kstart + (int)(r / nj);
  const long pt = b.p(i, j, k), ps = b.pstride();

  double yy[NS];
  double rho = 0.0, uu, vv, ww;
  if (LES) {
#pragma unroll
    for (int s = 0; s < NS; ++s) yy[s] = w_aux[pt + ps * s];
    rho = w_aux[pt + ps * (X::J_R - 1)];
    uu = w_aux[pt + ps * (X::J_U - 1)];
    vv = w_aux[pt + ps * (X::J_V - 1)];
    ww = w_aux[pt + ps * (X::J_W - 1)];
  } else {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double v = w[pt + ps * s];
      if (!(v > 0.0)) {               // max(rho*Y, 0): written back only when clipped
        v = 0.0;
        w[pt + ps * s] = 0.0;
      }
      yy[s] = v;
      rho = (s == 0) ? v : rho + v;
    }
  }
  const double ri = 1.0 / rho;
  if (!LES) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      yy[s] = yy[s] * ri;
      w_aux[pt + ps * s] = yy[s];
    }
    uu = w[pt + ps * (X::I_U - 1)] * ri;
    vv = w[pt + ps * (X::I_V - 1)] * ri;
    ww = w[pt + ps * (X::I_W - 1)] * ri;
  }
  const double rhoe = w[pt + ps * (X::I_E - 1)];
  const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
  const double ee = rhoe * ri - qq;
  int nrange[NS];
  const double tt = temperature_from_e_lds<NS, LO>(ee, w_aux[pt + ps * (X::J_T - 1)], TL, tol_iter_nr, yy, nrange);
  double rmix = tb.rgas[0] * yy[0];
#pragma unroll
  for (int s = 1; s < NS; ++s) rmix = rmix + tb.rgas[s] * yy[s];
#pragma unroll
  for (int s = 0; s < NS; ++s) nrange[s] = range_of<NS>(TL, s, tt);
  const double cploc = mixture_cp_lds<NS, LO>(TL, nrange, yy, TPow<LO>(tt));
  const double gamloc = cploc / (cploc - rmix);
  const double pp = rho * tt * rmix;
  const double c = sqrt(gamloc * pp * ri);
  if (LES) {
    w_aux[pt + ps * (X::J_H - 1)] = (rhoe + pp) * ri;
    w_aux[pt + ps * (X::J_T - 1)] = tt;
    w_aux[pt + ps * (X::J_P - 1)] = pp;
    w_aux[pt + ps * (X::J_C - 1)] = c;
    w_aux[pt + ps * (X::J_K_COND - 1)] = cploc;
    return;
  }
  w_aux[pt + ps * (X::J_R - 1)] = rho;
  w_aux[pt + ps * (X::J_U - 1)] = uu;
  w_aux[pt + ps * (X::J_V - 1)] = vv;
  w_aux[pt + ps * (X::J_W - 1)] = ww;
  w_aux[pt + ps * (X::J_H - 1)] = (rhoe + pp) * ri;
  w_aux[pt + ps * (X::J_T - 1)] = tt;
  w_aux[pt + ps * (X::J_P - 1)] = pp;
  w_aux[pt + ps * (X::J_C - 1)] = c;
  // Bilger mixture fraction (kernels_gpu.F90:6228-6243)
  double zz = 0.5;
  if (tb.n_eoi != 1) {
    double beta = 0.0;
    for (int le = 0; le < tb.n_eoi; ++le) {
      double yatm = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (yy[s] > 1e-14) yatm = yatm + tb.nainsp[s + NS * le] * tb.aw_eoi[le] * yy[s] * tb.mwinv[s];
      beta = beta + tb.coeff_eoi[le] * yatm / tb.aw_eoi[le];
    }
    zz = (beta - tb.beta0[1]) / (tb.beta0[0] - tb.beta0[1]);
  }
  w_aux[pt + ps * (X::J_Z - 1)] = zz;
}

// LES = true adds the eddy viscosity mu_sgs (J_LES1) as eval_aux_les_cuf does: mu + mu_sgs,
// k + mu_sgs c_p / Pr_sgs (c_p left in J_K_COND by the thermo kernel), rho D_s + mu_sgs / Sc_sgs.
template <int NS, bool LES = false>
__global__ void __launch_bounds__(256) eval_transport_kernel(Box b, int istart, int jstart, int kstart, int ni,
                                                             int nj, long npts, double *__restrict__ w_aux,
                                                             AuxTables tb, double les_pr = 0.0,
                                                             double les_sc = 0.0) {
  using X = Idx<NS>;
  __shared__ double c_inv[NS * NS];  // 1/sqrt(1 + W_l/W_m)
  __shared__ double c_pow[NS * NS];  // (W_m/W_l)^(1/4)
  for (int q = threadIdx.x; q < NS * NS; q += blockDim.x) {
    int l = q % NS, m = q / NS;
    c_inv[q] = 1.0 / sqrt(1.0 + tb.mw[l] * tb.mwinv[m]);
    c_pow[q] = pow(tb.mw[m] * tb.mwinv[l], 0.25);
  }
  __syncthreads();
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  int i = istart + (int)(t % ni);
  long r = t / ni;
  int j = jstart + (int)(r % nj);
  int k = kstart + (int)(r / nj);
  const long pt = b.p(i, j, k), ps = b.pstride();
  double yy[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) yy[s] = w_aux[pt + ps * s];
  const double tt = w_aux[pt + ps * (X::J_T - 1)];
  const double rho = w_aux[pt + ps * (X::J_R - 1)];
  const double pp = w_aux[pt + ps * (X::J_P - 1)];
  double rmix = tb.rgas[0] * yy[0];
#pragma unroll
  for (int s = 1; s < NS; ++s) rmix = rmix + tb.rgas[s] * yy[s];
  // table row
  int itt = (int)((tt - tb.t_min_tab) / tb.dt_tab) + 1;
  itt = itt < 1 ? 1 : itt;
  itt = itt > tb.num_t_tab ? tb.num_t_tab : itt;
  const double tloc = tb.t_min_tab + (itt - 1) * tb.dt_tab;
  const double dtt = (tt - tloc) / tb.dt_tab;
  const long nt = tb.num_t_tab + 1;
  // Mathur conductivity
  const double mwmixt = tb.r_univ / rmix;
  double k1 = 0.0, k2 = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double *lp = tb.lambda + (itt - 1) + nt * s;
    double l0 = lp[0], l1 = lp[1];
    double klsp = l0 + (l1 - l0) * dtt;
    const double xs = yy[s] * mwmixt * tb.mwinv[s];
    k1 = k1 + xs * klsp;
    k2 = k2 + xs * rcp_nr(klsp);
  }
  const double k_cond = 0.5 * (k1 + rcp_nr(k2));
  // Wilke viscosity: sqrt(mu_m/mu_l) evaluated as sqrt(mu_m) * (1/sqrt(mu_l))
  double mus[NS], smu[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double *vp = tb.visc + (itt - 1) + nt * s;
    double v0 = vp[0], v1 = vp[1];
    mus[s] = v0 + (v1 - v0) * dtt;
    smu[s] = sqrt(mus[s]);
  }
  __builtin_amdgcn_sched_barrier(0);
  double mu = 0.0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {      // unrolled (no register indexing), one l at a time
    const double isl = rcp_nr(smu[l]);
    double den = 0.0;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      double t1 = 1.0 + (smu[m] * isl) * c_pow[l + NS * m];
      double phi = c_inv[l + NS * m] * (t1 * t1);
      den = den + yy[m] * phi * tb.mwinv[m];
    }
    mu = mu + mus[l] * yy[l] * tb.mwinv[l] * rcp_nr(den);
    __builtin_amdgcn_sched_barrier(0);
  }
  mu = sqrt(8.0) * mu;
  double mu_sgs = 0.0;
  if (LES) {
    mu_sgs = w_aux[pt + ps * (X::J_LES1 - 1)];
    const double cploc = w_aux[pt + ps * (X::J_K_COND - 1)];
    w_aux[pt + ps * (X::J_MU - 1)] = mu + mu_sgs;
    w_aux[pt + ps * (X::J_K_COND - 1)] = k_cond + mu_sgs * cploc / les_pr;
  } else {
    w_aux[pt + ps * (X::J_MU - 1)] = mu;
    w_aux[pt + ps * (X::J_K_COND - 1)] = k_cond;
  }
  // Bird species diffusion into the mixture (mole fractions rebuilt here rather than kept
  // live through the Wilke sums; the fence keeps the table reads below from being hoisted
  // into the Wilke loop)
  __builtin_amdgcn_sched_barrier(0);
  double mwx = mwmixt;
  opaque(mwx);
  double xs[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) xs[s] = yy[s] * mwx * tb.mwinv[s];
  const double fac = tb.p0 / pp * rho;
#pragma unroll 1
  for (int l = 0; l < NS; ++l) {
    double dden = 0.0;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      if (m != l) {
        const double *dp = tb.diffbin + (itt - 1) + nt * (l + NS * m);
        double d0 = dp[0], d1 = dp[1];
        double dij = d0 + (d1 - d0) * dtt;
        dden = dden + xs[m] * rcp_nr(dij);
      }
    }
    const double rd = (dden > 1.0e-15) ? (1.0 - yy[l]) / dden * fac : 0.0;
    w_aux[pt + ps * (X::J_D_START + l)] = LES ? rd + mu_sgs / les_sc : rd;
  }
}

// DUC = false: div/3 only (the sensor is not wanted), from the three diagonal derivatives
template <int NS, int LB, bool DUC>
__global__ void __launch_bounds__(256) eval_aux2_kernel(Box b, int lmax, double *__restrict__ w_aux,
                                                        const double *__restrict__ cd1,
                                                        const double *__restrict__ dcsidx,
                                                        const double *__restrict__ detady,
                                                        const double *__restrict__ dzitdz,
                                                        double eps_sensor, int sensor_type, StencilRows SR) {
  using X = Idx<NS>;
  int i, j, k;
  if (!stencil_point(b, SR, i, j, k)) return;
  const long pt = b.p(i, j, k), ps = b.pstride();
  const long sx = 1, sy = b.px(), sz = b.px() * b.py();
  const double *U = w_aux + ps * (X::J_U - 1), *V = w_aux + ps * (X::J_V - 1), *Wv = w_aux + ps * (X::J_W - 1);
  double ux = 0, vx = 0, wx = 0, uy = 0, vy = 0, wy = 0, uz = 0, vz = 0, wz = 0;
  // one stencil offset of one direction in flight at a time (see finish_mom_kernel)
#pragma unroll
  for (int l = 1; l <= LB; ++l) {
    if (l > lmax) break;
    double ccl = cd1[(l - 1) + 4 * (lmax - 1)];
    long q = pt;
    opaque(q);
    ux = ux + ccl * (U[q + l * sx] - U[q - l * sx]);
    if (DUC) {
      vx = vx + ccl * (V[q + l * sx] - V[q - l * sx]);
      wx = wx + ccl * (Wv[q + l * sx] - Wv[q - l * sx]);
    }
    __builtin_amdgcn_sched_barrier(0);
    q = pt + l * sy;
    opaque(q);
    long qm = pt - l * sy;
    opaque(qm);
    if (DUC) uy = uy + ccl * (U[q] - U[qm]);
    vy = vy + ccl * (V[q] - V[qm]);
    if (DUC) wy = wy + ccl * (Wv[q] - Wv[qm]);
    __builtin_amdgcn_sched_barrier(0);
    q = pt + l * sz;
    opaque(q);
    qm = pt - l * sz;
    opaque(qm);
    if (DUC) {
      uz = uz + ccl * (U[q] - U[qm]);
      vz = vz + ccl * (V[q] - V[qm]);
    }
    wz = wz + ccl * (Wv[q] - Wv[qm]);
    __builtin_amdgcn_sched_barrier(0);
  }
  const double dx = dcsidx[i - 1], dy = detady[j - 1], dz = dzitdz[k - 1];
  ux *= dx; vx *= dx; wx *= dx;
  uy *= dy; vy *= dy; wy *= dy;
  uz *= dz; vz *= dz; wz *= dz;
  const double div = ux + vy + wz;
  if (DUC) {
    const double ox = wy - vz, oy = uz - wx, oz = vx - uy;
    const double omod2 = ox * ox + oy * oy + oz * oz;
    double s = -div / sqrt(omod2 + div * div + eps_sensor);
    if (sensor_type != 0 && s < 0.0) s = 0.0;
    w_aux[pt + ps * (X::J_DUC - 1)] = s * s;
  }
  w_aux[pt + ps * (X::J_DIV - 1)] = div / 3.0;
}

}  // namespace streams

using namespace streams;

extern "C" int streams_eval_aux(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                                int istart, int iend, int jstart, int jend, int kstart, int kend,
                                double *w, double *w_aux, double p0, double t_min_tab, double dt_tab,
                                double r_univ, const double *rgas, const double *cv_coeff,
                                const double *cp_coeff, int nsetcv, const double *trange,
                                int indx_cp_l, int indx_cp_r, double tol_iter_nr, void *stream,
                                const double *mw, const double *mwinv, const double *visc_species,
                                const double *lambda_species, const double *diffbin_species,
                                int num_t_tab, int n_eoi, const double *aw_eoi, const int *nainsp,
                                const double *beta0, const double *coeff_eoi) {
  if (nv != n_s + 4 || nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  if (iend < istart || jend < jstart || kend < kstart) return 0;
  if (istart < 1 - ng || iend > nx + ng || jstart < 1 - ng || jend > ny + ng || kstart < 1 - ng ||
      kend > nz + ng)
    return STREAMS_E_DIMS;
  if (num_t_tab < 1) return STREAMS_E_DIMS;
  if (nsetcv < 1 || nsetcv > TMAXSET) return STREAMS_E_OPTION;
  Box b{nx, ny, nz, ng};
  Thermo th{cv_coeff, cp_coeff, trange, nsetcv, indx_cp_l, indx_cp_r};
  AuxTables tb{rgas, mw, mwinv, visc_species, lambda_species, diffbin_species, num_t_tab, t_min_tab,
               dt_tab, p0, r_univ, n_eoi, aw_eoi, beta0, coeff_eoi, nainsp};
  int ni = iend - istart + 1, nj = jend - jstart + 1, nk = kend - kstart + 1;
  long npts = (long)ni * nj * nk;
  STREAMS_DISPATCH_NS(n_s, STREAMS_DISPATCH_NASA(indx_cp_l, indx_cp_r,
                               (eval_aux_kernel<NS, LO><<<blocks_for(npts, 256), 256, 0, as_stream(stream)>>>(
                                   b, istart, jstart, kstart, ni, nj, npts, w, w_aux, th, tb, tol_iter_nr),
                                eval_transport_kernel<NS><<<blocks_for(npts, 256), 256, 0, as_stream(stream)>>>(
                                   b, istart, jstart, kstart, ni, nj, npts, w_aux, tb))));
  return hip_status();
}

extern "C" int streams_eval_aux2(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order,
                                 double *w_aux, const double *coeff_deriv1, const double *dcsidx,
                                 const double *detady, const double *dzitdz, double eps_sensor,
                                 int sensor_type, void *stream) {
  if (nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  int lmax = visc_order / 2;
  if (lmax < 1 || lmax > 4 || lmax > ng) return STREAMS_E_OPTION;
  Box b{nx, ny, nz, ng};
  long n = (long)nx * ny * nz;
  if (n == 0) return 0;
  const StencilRows SR = stencil_rows(b);
  if (SR.tiles(nz) > 0xffffff00u) return STREAMS_E_DIMS;
  const unsigned gm = xcd_grid(SR.tiles(nz));
  hipStream_t st = as_stream(stream);
  if (sensor_type >= 0) {
    STREAMS_DISPATCH_NS(n_s, (lmax == 3 ? eval_aux2_kernel<NS, 3, true><<<gm, 256, 0, st>>>(
                                              b, lmax, w_aux, coeff_deriv1, dcsidx, detady, dzitdz, eps_sensor,
                                              sensor_type, SR)
                                        : eval_aux2_kernel<NS, 4, true><<<gm, 256, 0, st>>>(
                                              b, lmax, w_aux, coeff_deriv1, dcsidx, detady, dzitdz, eps_sensor,
                                              sensor_type, SR)));
  } else {
    STREAMS_DISPATCH_NS(n_s, (lmax == 3 ? eval_aux2_kernel<NS, 3, false><<<gm, 256, 0, st>>>(
                                              b, lmax, w_aux, coeff_deriv1, dcsidx, detady, dzitdz, eps_sensor,
                                              sensor_type, SR)
                                        : eval_aux2_kernel<NS, 4, false><<<gm, 256, 0, st>>>(
                                              b, lmax, w_aux, coeff_deriv1, dcsidx, detady, dzitdz, eps_sensor,
                                              sensor_type, SR)));
  }
  return hip_status();
}

extern "C" int streams_eval_aux_les(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux, double *w,
                                    double *w_aux, double p0, double t_min_tab, double dt_tab, double r_univ,
                                    const double *rgas, const double *cv_coeff, const double *cp_coeff, int nsetcv,
                                    const double *trange, int indx_cp_l, int indx_cp_r, double tol_iter_nr,
                                    const double *mw, const double *mwinv, const double *visc_species,
                                    const double *lambda_species, const double *diffbin_species, int num_t_tab,
                                    double les_pr, double les_sc, void *stream) {
  if (nv != n_s + 4 || nv_aux < 2 * n_s + 14) return STREAMS_E_DIMS;
  if (nx < 1 || ny < 1 || nz < 1 || ng < 0 || num_t_tab < 1) return STREAMS_E_DIMS;
  if (nsetcv < 1 || nsetcv > TMAXSET) return STREAMS_E_OPTION;
  Box b{nx, ny, nz, ng};
  Thermo th{cv_coeff, cp_coeff, trange, nsetcv, indx_cp_l, indx_cp_r};
  AuxTables tb{rgas, mw, mwinv, visc_species, lambda_species, diffbin_species, num_t_tab, t_min_tab,
               dt_tab, p0, r_univ, 1, nullptr, nullptr, nullptr, nullptr};
  const int ni = nx + 2 * ng, nj = ny + 2 * ng;
  const long npts = b.pstride();
  STREAMS_DISPATCH_NS(n_s, STREAMS_DISPATCH_NASA(indx_cp_l, indx_cp_r,
                               (eval_aux_kernel<NS, LO, true><<<blocks_for(npts, 256), 256, 0, as_stream(stream)>>>(
                                   b, 1 - ng, 1 - ng, 1 - ng, ni, nj, npts, w, w_aux, th, tb, tol_iter_nr),
                                eval_transport_kernel<NS, true><<<blocks_for(npts, 256), 256, 0, as_stream(stream)>>>(
                                   b, 1 - ng, 1 - ng, 1 - ng, ni, nj, npts, w_aux, tb, les_pr, les_sc))));
  return hip_status();
}
