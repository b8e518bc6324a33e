// finish.hip — the pointwise tail of an RK substep in one pass over the interior:
//   visflx_reduced_cuf (kernels_gpu.F90:3539-3656)   fl(mom, E) -= reduced stress terms
//   compute_chemistry_cuf (:9503-9674)               fl(species) -= omega_s W_s
//   update_flux_cuf (:3297-3314)                     fln -= gamdt * fl
//   update_field_cuf (:3336-3355)                    w += fln (fluid cells)
// fl, fln and w are each read and written once.  The chemistry takes rho, Y and T from
// w_aux as written by eval_aux in the same substep: w is unchanged since then, so rho and
// Y are the identical numbers compute_chemistry_cuf recomputes; its extra Newton iteration
// moves T by less than the 1e-12 K stopping tolerance.
#include "kinpack.h"

namespace streams {

struct FinishArgs {
  const double *cd1, *dcsidx, *detady, *dzitdz;
  const int *mask;
  const double *mw;
  KinPacked K;
  int lmax, reduced, chem;
  double gamdt;
  // fold = 1: the next substep's init_flux is done here (fln = -rhodt_next * fl, written in
  // place of fl, which then needs no store: the next x sweep overwrites it with init = 2)
  int fold;
  double rhodt_next;
};

// One interior cell per thread, i fastest, so that every access of a wave is a coalesced
// 512-byte segment of one variable plane.  The species equations (chemistry, then their RK
// update) and the momentum / energy equations (reduced viscous terms, then their RK update)
// touch disjoint data, so they run as two kernels: each holds fewer registers than the
// pair would, which buys occupancy, and the pair moves exactly the bytes of one fused pass.
template <int NS>
__device__ inline bool cell_of(const Box &b, long t, int &i, int &j, int &k, long &pt) {
  if (t >= b.fstride()) return false;
  i = 1 + (int)(t % b.nx);
  const long r = t / b.nx;
  j = 1 + (int)(r % b.ny);
  k = 1 + (int)(r / b.ny);
  pt = b.p(i, j, k);
  return true;
}

// 6 waves per SIMD for up to 9 species (80 VGPRs, 12 B/lane of spills: faster than the
// compiler's own 5); beyond 9 species those 80 VGPRs spill 144-224 B/lane, and 4 waves
// (about 100 VGPRs, no spills) is the fastest measured.
template <int NS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NS <= 9 ? 6 : 4))) finish_species_kernel(Box b, double *__restrict__ w,
                                                             const double *__restrict__ wa, double *__restrict__ fl,
                                                             double *__restrict__ fln, FinishArgs F) {
  using X = Idx<NS>;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  long pt;
  if (!cell_of<NS>(b, t, i, j, k, pt)) return;
  const long ps = b.pstride(), fs_ = b.fstride();
  const double gamdt = F.gamdt;
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = fl[t + fs_ * s];
  if (F.chem) {
    const double rho = wa[pt + ps * (X::J_R - 1)];
    const double tt = wa[pt + ps * (X::J_T - 1)];
    double conc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) conc[s] = rho * wa[pt + ps * s] / F.mw[s];
    chemistry_source_packed<NS>(F.K, tt, conc, acc);
  }
  opaque(t);
  opaque(pt);
  const bool fluid = F.mask == nullptr || F.mask[pt] == 0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double fn = fln[t + fs_ * s] - gamdt * acc[s];
    if (F.fold) {
      fln[t + fs_ * s] = -F.rhodt_next * acc[s];
    } else {
      fl[t + fs_ * s] = acc[s];
      fln[t + fs_ * s] = fn;
    }
    if (fluid) w[pt + ps * s] = w[pt + ps * s] + fn;
  }
}

template <int NS, int LB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) finish_mom_kernel(Box b, double *__restrict__ w, const double *__restrict__ wa,
                                                         double *__restrict__ fl, double *__restrict__ fln,
                                                         FinishArgs F, StencilRows SR) {
  using X = Idx<NS>;
  int i, j, k;
  if (!stencil_point(b, SR, i, j, k)) return;
  const long t = b.f(i, j, k), pt = b.p(i, j, k);
  const long ps = b.pstride(), fs_ = b.fstride();
  const bool fluid = F.mask == nullptr || F.mask[pt] == 0;
  const double gamdt = F.gamdt;
  double red[4] = {0.0, 0.0, 0.0, 0.0};   // reduced viscous terms, applied after the stencil
  if (F.reduced) {
    const long sx = 1, sy = b.px(), sz = b.px() * b.py();
    const double *U = wa + ps * (X::J_U - 1), *V = wa + ps * (X::J_V - 1), *Wv = wa + ps * (X::J_W - 1);
    const double *MU = wa + ps * (X::J_MU - 1), *DV = wa + ps * (X::J_DIV - 1);
    const double uu = U[pt], vv = V[pt], ww = Wv[pt], mu = MU[pt], div3l = DV[pt];
    double ux = 0, vx = 0, wx = 0, mux = 0, dvx = 0, uy = 0, vy = 0, wy = 0, muy = 0, dvy = 0;
    double uz = 0, vz = 0, wz = 0, muz = 0, dvz = 0;
    const int lmax = F.lmax;
    // unrolled to the compile-time bound LB >= lmax, one stencil offset of one direction
    // (10 loads) in flight at a time: the kernel then fits 6 waves per SIMD
#pragma unroll
    for (int l = 1; l <= LB; ++l) {
      if (l > lmax) break;
      const double ccl = F.cd1[(l - 1) + 4 * (lmax - 1)];
      // a fresh copy of the point index per group makes the compiler rebuild the ten
      // addresses from it rather than keep dozens of 64-bit addresses live
      long q = pt;
      opaque(q);
      ux = ux + ccl * (U[q + l * sx] - U[q - l * sx]);
      vx = vx + ccl * (V[q + l * sx] - V[q - l * sx]);
      wx = wx + ccl * (Wv[q + l * sx] - Wv[q - l * sx]);
      mux = mux + ccl * (MU[q + l * sx] - MU[q - l * sx]);
      dvx = dvx + ccl * (DV[q + l * sx] - DV[q - l * sx]);
      __builtin_amdgcn_sched_barrier(0);
      q = pt + l * sy;
      opaque(q);
      long qm = pt - l * sy;
      opaque(qm);
      uy = uy + ccl * (U[q] - U[qm]);
      vy = vy + ccl * (V[q] - V[qm]);
      wy = wy + ccl * (Wv[q] - Wv[qm]);
      muy = muy + ccl * (MU[q] - MU[qm]);
      dvy = dvy + ccl * (DV[q] - DV[qm]);
      __builtin_amdgcn_sched_barrier(0);
      q = pt + l * sz;
      opaque(q);
      qm = pt - l * sz;
      opaque(qm);
      uz = uz + ccl * (U[q] - U[qm]);
      vz = vz + ccl * (V[q] - V[qm]);
      wz = wz + ccl * (Wv[q] - Wv[qm]);
      muz = muz + ccl * (MU[q] - MU[qm]);
      dvz = dvz + ccl * (DV[q] - DV[qm]);
      __builtin_amdgcn_sched_barrier(0);
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
    red[0] = sigx;
    red[1] = sigy;
    red[2] = sigz;
    red[3] = sigq;
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const long q = t + fs_ * (NS + m);
    const double acc = F.reduced ? fl[q] - red[m] : fl[q];
    const double fn = fln[q] - gamdt * acc;
    if (F.fold) {
      fln[q] = -F.rhodt_next * acc;
    } else {
      fl[q] = acc;
      fln[q] = fn;
    }
    if (fluid) w[pt + ps * (NS + m)] = w[pt + ps * (NS + m)] + fn;
  }
}

}  // namespace streams

using namespace streams;

// Host packer of the kinetics records read by the fused chemistry (layout in kinpack.h).
extern "C" int streams_kinetics_pack(int n_s, int nreactions, const double *arr_a, const double *arr_b,
                                     const double *arr_ea, const double *tb_eff, const double *falloff_coeffs,
                                     const int *reac_ty, const int *isrev, const double *r_coeffs,
                                     const double *p_coeffs, const double *mw, double *out, int cap) {
  if (n_s < 1 || n_s > KTB) return STREAMS_E_NSPECIES;
  if (nreactions < 0) return STREAMS_E_DIMS;
  const long need = (long)nreactions * KREC;
  if (out == nullptr) return (int)need;
  if (cap < need) return STREAMS_E_DIMS;
  const int nr = nreactions;
  for (int r = 0; r < nr; ++r) {
    double *R = out + (long)KREC * r;
    for (int q = 0; q < KREC; ++q) R[q] = 0.0;
    int *I = reinterpret_cast<int *>(R);
    I[KI_TY] = reac_ty[r];
    I[KI_REV] = isrev[r];
    int nre = 0, npr = 0;
    for (int s = 0; s < n_s; ++s) {
      const double rr = r_coeffs[r + (long)nr * s], pp = p_coeffs[r + (long)nr * s];
      if (rr != 0.0) {
        if (nre == KSIDE) return STREAMS_E_OPTION;
        I[KI_RE + nre] = s;
        R[KD_REX + nre] = rr;
        ++nre;
      }
      if (pp != 0.0) {
        if (npr == KSIDE) return STREAMS_E_OPTION;
        I[KI_PR + npr] = s;
        R[KD_PEX + npr] = pp;
        ++npr;
      }
      R[KD_NET + s] = (pp - rr) * mw[s];
      R[KD_TB + s] = tb_eff[r + (long)nr * s];
    }
    I[KI_NRE] = nre;
    I[KI_NPR] = npr;
    I[KI_ARR] = (arr_b[r] != 0.0 || arr_ea[r] != 0.0) ? 1 : 0;
    R[KD_A] = arr_a[r];
    R[KD_A + 1] = arr_b[r];
    R[KD_A + 2] = arr_ea[r];
    R[KD_AINF] = arr_a[r + nr];
    R[KD_AINF + 1] = arr_b[r + nr];
    R[KD_AINF + 2] = arr_ea[r + nr];
    for (int q = 0; q < 5; ++q) R[KD_FC + q] = falloff_coeffs[r + (long)nr * q];
    I[KI_T2] = falloff_coeffs[r + 3L * nr] != -3.14 ? 1 : 0;
  }
  return (int)need;
}

// Tail of one RK substep: reduced viscous terms (reduced=1), explicit chemistry (chem=1),
// update_flux and update_field, in one pass; fold=1 also does the next substep's init_flux.
static int rhs_finish_impl(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order, double *w,
                           const double *w_aux, double *fl, double *fln, const double *coeff_deriv1,
                           const double *dcsidx, const double *detady, const double *dzitdz, int reduced, int chem,
                           int nreactions, const double *mw, const double *kin_pack, const double *kc_tab,
                           int num_t_tab, double t_min_tab, double dt_tab, double r_univ, double gamdt, int fold,
                           double rhodt_next, const int *fluid_mask, void *stream) {
  if (nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  const int lmax = visc_order / 2;
  if (reduced && (lmax < 1 || lmax > 4 || lmax > ng)) return STREAMS_E_OPTION;
  if (n_s > KTB) return STREAMS_E_NSPECIES;
  Box b{nx, ny, nz, ng};
  const long n = b.fstride();
  if (n == 0) return 0;
  FinishArgs F{coeff_deriv1, dcsidx, detady, dzitdz, fluid_mask, mw,
               KinPacked{kin_pack, kc_tab, nreactions, num_t_tab, t_min_tab, dt_tab, r_univ},
               lmax, reduced, chem && nreactions > 0, gamdt, fold, rhodt_next};
  const StencilRows SR = stencil_rows(b);
  if (SR.tiles(nz) > 0xffffff00u) return STREAMS_E_DIMS;
  const unsigned gm = xcd_grid(SR.tiles(nz));
  STREAMS_DISPATCH_NS(n_s, (finish_species_kernel<NS><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(b, w, w_aux, fl, fln, F),
                            (lmax == 3 ? finish_mom_kernel<NS, 3><<<gm, 256, 0, as_stream(stream)>>>(b, w, w_aux, fl, fln, F, SR)
                                       : finish_mom_kernel<NS, 4><<<gm, 256, 0, as_stream(stream)>>>(b, w, w_aux, fl, fln, F, SR))));
  return hip_status();
}

extern "C" int streams_rhs_finish(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order, double *w,
                                  const double *w_aux, double *fl, double *fln, const double *coeff_deriv1,
                                  const double *dcsidx, const double *detady, const double *dzitdz, int reduced,
                                  int chem, int nreactions, const double *mw, const double *kin_pack,
                                  const double *kc_tab, int num_t_tab, double t_min_tab, double dt_tab,
                                  double r_univ, double gamdt, const int *fluid_mask, void *stream) {
  return rhs_finish_impl(n_s, nx, ny, nz, ng, nv_aux, visc_order, w, w_aux, fl, fln, coeff_deriv1, dcsidx, detady,
                         dzitdz, reduced, chem, nreactions, mw, kin_pack, kc_tab, num_t_tab, t_min_tab, dt_tab,
                         r_univ, gamdt, 0, 0.0, fluid_mask, stream);
}

extern "C" int streams_rhs_finish_next(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order,
                                       double *w, const double *w_aux, double *fl, double *fln,
                                       const double *coeff_deriv1, const double *dcsidx, const double *detady,
                                       const double *dzitdz, int reduced, int chem, int nreactions,
                                       const double *mw, const double *kin_pack, const double *kc_tab,
                                       int num_t_tab, double t_min_tab, double dt_tab, double r_univ, double gamdt,
                                       double rhodt_next, const int *fluid_mask, void *stream) {
  return rhs_finish_impl(n_s, nx, ny, nz, ng, nv_aux, visc_order, w, w_aux, fl, fln, coeff_deriv1, dcsidx, detady,
                         dzitdz, reduced, chem, nreactions, mw, kin_pack, kc_tab, num_t_tab, t_min_tab, dt_tab,
                         r_univ, gamdt, 1, rhodt_next, fluid_mask, stream);
}
