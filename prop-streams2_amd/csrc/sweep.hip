// sweep.hip — fused convective + viscous sweep along one direction, LDS-tiled, in rounds.
//
// A workgroup of 256 threads owns a tile of C columns x S1 faces (S1-1 cells updated) of
// one direction, one thread per face.  The KEEP split-form central flux (the central branch
// of euler_{x,y,z}_hybrid, kernels_gpu.F90:150-337) is evaluated in rounds that each stage
// part of the stencil variables in LDS with coalesced loads:
//   round 1: rho, u, v, w, h, p -> the pair factors (rho_a+rho_b)(un_a+un_b)*KEEP(n), kept in
//            registers, and the momentum / energy fluxes;
//   round 2..: the mass fractions, KCH species at a time -> the species fluxes.
// Each round's face fluxes meet in a small LDS exchange buffer and are differenced into the
// thread's cell right away, so only the cell's accumulator stays live between rounds.  The
// viscous / diffusive flux of the face (visflx_{x,y,z}_cuf, :3658-4031) then reads its two
// cells from HBM (cache hits: each cell is read by two faces) with the species enthalpies
// from an LDS copy of the cp tables, and is exchanged the same way.  Faces flagged by the
// shock sensor take the WENO flux that the pre-pass parked in fhat.  One read-modify-write
// of fl per cell:
//     fl += (Fc(c) - Fc(c-1)) * dxi(c) + 2/(x(c+1)-x(c-1)) * (Fv(c) - Fv(c-1))
// which is euler_{x,y,z}_update_cuf followed by the update of visflx_{x,y,z}_cuf
// (:732-793, :3775-3788).  In the first sweep of a substep the RK bookkeeping of
// init_flux_cuf (:28-45) is folded in: fln = -rhodt * fl_old and fl = this sweep's part.
//
// Tiles: x (faces along the lanes): C = 8 lines x S1 = 32 faces; y, z (lanes along i, the
// contiguous index): C = 16 columns x S1 = 16 faces.  LDS layout keeps the lane index
// fastest.
#include "euler_dev.h"
#include "thermo_dev.h"
#include "visc_dev.h"

namespace streams {

template <int DIR> struct TileCfg { static constexpr int C = 16, S1 = 16; };
template <> struct TileCfg<1> { static constexpr int C = 8, S1 = 32; };
constexpr int LMX = 4;                      // largest half stencil (ep_order 8)
constexpr int KCH = 6;                      // LDS slots per staging round

struct SweepArgs {
  EulerArgs E;
  ViscArgs V;
  double *fln;
  unsigned ncb, nseg, ntiles;   // column tiles, segments along the sweep, their product with the third extent
  double rhodt;
  int init;          // 1: fold init_flux (fln = -rhodt*fl_old, fl = contribution); 2: fl = contribution
  int visc;          // add the viscous flux
  int check_shock;   // sensor threshold below 1: consult the pre-pass results
};

template <int DIR>
struct Tile {
  static constexpr int C = TileCfg<DIR>::C, S1 = TileCfg<DIR>::S1, R = S1 + 2 * LMX - 1;
  static constexpr int SLOT = C * R;        // doubles per staged variable
  static constexpr int XSLOT = C * S1;      // doubles per exchanged variable
  __device__ static inline int at(int c, int r) { return DIR == 1 ? c * R + r : r * C + c; }
  __device__ static inline int xat(int c, int fl) { return DIR == 1 ? c * S1 + fl : fl * C + c; }
};

// Pair factors and the momentum / energy fluxes of face f from the round-1 slots
// (0 rho, 1 u, 2 v, 3 w, 4 h -- or the internal energy e when nkeep >= 0 --, 5 p); rows are
// cell positions minus p0.
template <int DIR, int LB>
__device__ inline void keep_round1(const double *sb, int c, int f, int p0, int lmax, int nkeep, const double *cd1,
                                   bool zero, double *pf, double *fme) {
  using T = Tile<DIR>;
  auto V = [&](int slot, int p) { return sb[slot * T::SLOT + T::at(c, p - p0)]; };
  constexpr int SN = DIR == 1 ? 1 : DIR == 2 ? 2 : 3;   // slot of the normal velocity
  double ft2 = 0, ft3 = 0, ft4 = 0, ft5 = 0, ft6 = 0;
#pragma unroll
  for (int l = 1; l <= LB; ++l) {
    if (l > lmax) break;
    double uvs2 = 0, uvs3 = 0, uvs4 = 0, uvs5 = 0, uvs5_i = 0, uvs5_k = 0, uvs5_p = 0, uvs6 = 0;
#pragma unroll
    for (int m = 0; m <= l - 1; ++m) {
      const int q = l * (l - 1) / 2 + m;
      const int a = f - m, b = f - m + l;
      const double rhoi = V(0, a), uui = V(1, a), vvi = V(2, a), wwi = V(3, a), enti = V(4, a), ppi = V(5, a);
      const double rhoip = V(0, b), uuip = V(1, b), vvip = V(2, b), wwip = V(3, b), entip = V(4, b), ppip = V(5, b);
      const double uni = SN == 1 ? uui : SN == 2 ? vvi : wwi;
      const double unip = SN == 1 ? uuip : SN == 2 ? vvip : wwip;
      const double rhom = rhoi + rhoip;
      if (nkeep >= 0) {
        double drhof = 1.0;
        if (nkeep > 0) {
          const double drho = 2.0 * (rhoip - rhoi) / rhom;
          double ts = (0.5 * drho) * (0.5 * drho);
          const double t2 = ts;
          double sden = 1.0 + ts / 3.0;
          for (int n = 2; n <= nkeep; ++n) { ts = t2 * ts; sden = sden + ts / (1.0 + 2 * n); }
          drhof = 1.0 / sden;
        }
        const double uv_part = (uni + unip) * rhom * drhof;
        pf[q] = uv_part;
        const double eei = enti, eeip = entip;    // slot 4 holds e for nkeep >= 0
        const double eem = eei + eeip;
        double deef = 1.0;
        if (nkeep > 0) {
          const double dee = 2.0 * (eeip - eei) / eem;
          double te = (0.5 * dee) * (0.5 * dee);
          const double te2 = te;
          double sde = 1.0 + te, sne = 1.0 + te / 3.0;
          for (int n = 2; n <= nkeep; ++n) { te = te2 * te; sde = sde + te; sne = sne + te / (1.0 + 2 * n); }
          deef = sne / sde;
        }
        uvs2 = uvs2 + uv_part * (uui + uuip);
        uvs3 = uvs3 + uv_part * (vvi + vvip);
        uvs4 = uvs4 + uv_part * (wwi + wwip);
        uvs5_i = uvs5_i + uv_part * eem * deef;
        uvs5_k = uvs5_k + uv_part * (uui * uuip + vvi * vvip + wwi * wwip);
        uvs5_p = uvs5_p + 4.0 * (uni * ppip + unip * ppi);
        uvs6 = uvs6 + 2.0 * (ppi + ppip);
      } else {
        const double uv_part = (uni + unip) * rhom;
        pf[q] = uv_part;
        uvs2 = uvs2 + uv_part * (uui + uuip);
        uvs3 = uvs3 + uv_part * (vvi + vvip);
        uvs4 = uvs4 + uv_part * (wwi + wwip);
        uvs5 = uvs5 + uv_part * (enti + entip);
        uvs6 = uvs6 + 2.0 * (ppi + ppip);
      }
    }
    const double cl = cd1[(l - 1) + 4 * (lmax - 1)];
    __builtin_amdgcn_sched_barrier(0);     // one stencil width at a time in registers
    ft2 = ft2 + cl * uvs2;
    ft3 = ft3 + cl * uvs3;
    ft4 = ft4 + cl * uvs4;
    ft5 = ft5 + cl * (nkeep >= 0 ? (uvs5_i + uvs5_k + uvs5_p) : uvs5);
    ft6 = ft6 + cl * uvs6;
  }
  fme[0] = zero ? 0.0 : 0.25 * ft2;
  fme[1] = zero ? 0.0 : 0.25 * ft3;
  fme[2] = zero ? 0.0 : 0.25 * ft4;
  fme[3] = zero ? 0.0 : 0.25 * ft5;
  fme[SN - 1] = fme[SN - 1] + 0.5 * ft6;
}

// species flux of face f from one staged mass-fraction slot
template <int DIR, int LB>
__device__ inline double keep_species(const double *slot, int c, int f, int p0, int lmax, const double *cd1,
                                      const double *pf, bool zero) {
  using T = Tile<DIR>;
  double ft1 = 0.0;
#pragma unroll
  for (int l = 1; l <= LB; ++l) {
    if (l > lmax) break;
    double uvs1 = 0.0;
#pragma unroll
    for (int m = 0; m <= l - 1; ++m)
      uvs1 = uvs1 + pf[l * (l - 1) / 2 + m] * (slot[T::at(c, f - m - p0)] + slot[T::at(c, f - m + l - p0)]);
    ft1 = ft1 + cd1[(l - 1) + 4 * (lmax - 1)] * uvs1;
  }
  return zero ? 0.0 : 0.25 * ft1;
}

template <int NS, int DIR, int LO, int LB>
__global__ void __launch_bounds__(256) sweep_kernel(Box b, const double *__restrict__ wa, double *__restrict__ fl,
                                                    double *__restrict__ fln, const double *__restrict__ fhat,
                                                    SweepArgs A) {
  using X = Idx<NS>;
  using T = Tile<DIR>;
  constexpr int NV = NS + 4, C = T::C, S1 = T::S1, R = T::R;
  __shared__ double sb[KCH * T::SLOT];
  __shared__ double xb[KCH * T::XSLOT];
  __shared__ double tlds[thermo_lds_doubles<NS>(TH_CPH)];
  const EulerArgs &E = A.E;
  const int t = threadIdx.x;
  int c, fl_;
  if (DIR == 1) { fl_ = t % S1; c = t / S1; }
  else { c = t % C; fl_ = t / C; }
  const ThermoLds TL = thermo_stage<NS, LO>(A.V.th, tlds, TH_CPH);
  // tile coordinates
  const int us = E.fs + 1, ue = E.fe;
  // tile of this workgroup (XCD-aware order; tiles beyond the last are idle)
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  if (tile >= A.ntiles) return;
  const unsigned nfast = DIR == 1 ? A.nseg : A.ncb, nmid = DIR == 1 ? A.ncb : A.nseg;
  const unsigned bfast = tile % nfast, bmid = (tile / nfast) % nmid, bslow = tile / (nfast * nmid);
  const int seg = DIR == 1 ? bfast : bmid;
  const int s0 = us + seg * (S1 - 1);
  const int cbase = 1 + (DIR == 1 ? bmid : bfast) * C;
  const int t2 = 1 + bslow;
  const int ncol = DIR == 1 ? b.ny : b.nx;
  const int lmax_b = E.lmax_base;
  const int p0 = s0 - LMX;
  const long ps = b.pstride();
  // rows of the tile that any face reads: [plo, phi]
  const int plo = s0 - lmax_b;
  int phi = s0 + S1 - 2;
  phi = (phi > ue ? ue : phi) + lmax_b;
  // staging of up to KCH variables (reference indices m[k]) into sb
  auto stage = [&](const int *vars, int nvar) {
    for (int idx = t; idx < C * R; idx += 256) {
      int cc, r;
      if (DIR == 1) { cc = idx / R; r = idx % R; }
      else { r = idx / C; cc = idx % C; }
      const int p = p0 + r;
      const int col = cbase + cc;
      if (col > ncol || p < plo || p > phi) continue;
      long base, st;
      line_base<DIR>(b, col, t2, base, st);
      const long g = base + (long)p * st;
      const int li = T::at(cc, r);
      for (int k = 0; k < nvar; ++k) sb[k * T::SLOT + li] = wa[g + ps * (vars[k] - 1)];
    }
  };
  // round 1 with the KEEP energy split: slot 4 holds the cell's internal energy
  // e = h - p/rho - (u^2+v^2+w^2)/2 (the expression the pair loop would otherwise evaluate
  // for every pair a cell takes part in) instead of h
  auto stage_round1_e = [&]() {
    for (int idx = t; idx < C * R; idx += 256) {
      int cc, r;
      if (DIR == 1) { cc = idx / R; r = idx % R; }
      else { r = idx / C; cc = idx % C; }
      const int p = p0 + r;
      const int col = cbase + cc;
      if (col > ncol || p < plo || p > phi) continue;
      long base, st;
      line_base<DIR>(b, col, t2, base, st);
      const long g = base + (long)p * st;
      const int li = T::at(cc, r);
      const double rho = wa[g + ps * (X::J_R - 1)], uu = wa[g + ps * (X::J_U - 1)];
      const double vv = wa[g + ps * (X::J_V - 1)], ww = wa[g + ps * (X::J_W - 1)];
      const double ent = wa[g + ps * (X::J_H - 1)], pp = wa[g + ps * (X::J_P - 1)];
      sb[0 * T::SLOT + li] = rho;
      sb[1 * T::SLOT + li] = uu;
      sb[2 * T::SLOT + li] = vv;
      sb[3 * T::SLOT + li] = ww;
      sb[4 * T::SLOT + li] = ent - pp / rho - 0.5 * (uu * uu + vv * vv + ww * ww);
      sb[5 * T::SLOT + li] = pp;
    }
  };
  const int col = cbase + c;
  const int f = s0 - 1 + fl_;
  const bool valid = col <= ncol && f <= ue;
  const bool upd = valid && fl_ >= 1;
  const int me = T::xat(c, fl_);
  const int left = DIR == 1 ? me - 1 : me - C;
  Line G{wa, 0, 0, ps};
  line_base<DIR>(b, col <= ncol ? col : ncol, t2, G.base, G.st);
  const bool shock = valid && E.check_shock_flag && shocked<NS>(G, f, E.weno_scheme, E.sensor_threshold);
  const long fidx = G.base + (long)f * G.st;     // face f in fhat
  int lmax = 1;
  bool zero = false;
  if (valid) {
    lmax = lmax_b + epo_at<DIR>(b, E.epo, col, t2, f);
    lmax = lmax < 1 ? 1 : lmax;
    zero = (f == 0 && E.zmin == 1) || (f == E.n && E.zmax == 1);
  }
  const double met = upd ? E.metric[f - 1] : 0.0;
  double acc[NV];
  double pf[LB * (LB + 1) / 2];
  // ---- round 1: rho, u, v, w, h, p
  {
    if (E.nkeep >= 0) {
      stage_round1_e();
    } else {
      const int vars[6] = {X::J_R, X::J_U, X::J_V, X::J_W, X::J_H, X::J_P};
      stage(vars, 6);
    }
    __syncthreads();
    double fme[4] = {0.0, 0.0, 0.0, 0.0};
    if (valid) {
      if (shock) {
#pragma unroll
        for (int k = 0; k < 4; ++k) fme[k] = fhat[fidx + ps * (NS + k)];
      } else {
        keep_round1<DIR, LB>(sb, c, f, p0, lmax, E.nkeep, E.cd1, zero, pf, fme);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) xb[k * T::XSLOT + me] = fme[k];
    }
    __syncthreads();
    if (upd) {
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[NS + k] = (fme[k] - xb[k * T::XSLOT + left]) * met;
    }
  }
  // ---- species rounds
#pragma unroll
  for (int s0k = 0; s0k < NS; s0k += KCH) {
    const int nk = NS - s0k < KCH ? NS - s0k : KCH;
    int vars[KCH];
#pragma unroll
    for (int k = 0; k < KCH; ++k) vars[k] = s0k + k + 1;
    __syncthreads();                        // sb and xb free
    stage(vars, nk);
    __syncthreads();
    if (valid) {
#pragma unroll 1
      for (int k = 0; k < nk; ++k)
        xb[k * T::XSLOT + me] = shock ? fhat[fidx + ps * (s0k + k)]
                                      : keep_species<DIR, LB>(sb + k * T::SLOT, c, f, p0, lmax, E.cd1, pf, zero);
    }
    __syncthreads();
    if (upd) {
#pragma unroll
      for (int k = 0; k < KCH; ++k)
        if (k < nk) acc[s0k + k] = (xb[k * T::XSLOT + me] - xb[k * T::XSLOT + left]) * met;
    }
  }
  // ---- viscous / diffusive flux (visflx_{x,y,z}_cuf), two cells a = f, c = f + 1
  if (A.visc) {
    const ViscArgs &V = A.V;
    const int a = f, cc = f + 1;
    double sigx = 0, sigy = 0, sigz = 0, sigq = 0, mwm = 1, mwmp = 1, mwmif = 0, fsum = 0, ta = 0, tc = 0;
    // the two cells' offsets in w_aux; opaque() before each phase makes the compiler
    // rebuild the per-variable addresses (one add per load) instead of keeping ~20 of them
    // live, and the fences bound the loads in flight
    long ga = G.base + (long)a * G.st, gc = ga + G.st;
    auto CA = [&](int m) { return wa[ga + ps * (m - 1)]; };
    auto CC = [&](int m) { return wa[gc + ps * (m - 1)]; };
    if (valid) {
      const double uu = CA(X::J_U), uup = CC(X::J_U);
      const double vv = CA(X::J_V), vvp = CC(X::J_V);
      const double ww = CA(X::J_W), wwp = CC(X::J_W);
      ta = CA(X::J_T);
      tc = CC(X::J_T);
      const double mu = CA(X::J_MU), mup = CC(X::J_MU);
      const double kc = CA(X::J_K_COND), kcp = CC(X::J_K_COND);
      const double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
      const double qqp = 0.5 * (uup * uup + vvp * vvp + wwp * wwp);
      const double dx = V.coord[f + 1 + b.ng - 1] - V.coord[f + b.ng - 1];
      double muf = mu + mup;
      muf = 0.5 * muf / dx;
      double kcf = kc + kcp;
      kcf = 0.5 * kcf / dx;
      sigx = (uup - uu) * muf;
      sigy = (vvp - vv) * muf;
      sigz = (wwp - ww) * muf;
      sigq = (tc - ta) * kcf + (qqp - qq) * muf;
      __builtin_amdgcn_sched_barrier(0);
      opaque(ga);
      opaque(gc);
      double rm = V.rgas[0] * CA(1), rmp = V.rgas[0] * CC(1);
#pragma unroll
      for (int s = 1; s < NS; ++s) {
        rm = rm + V.rgas[s] * CA(s + 1);
        rmp = rmp + V.rgas[s] * CC(s + 1);
        if (s % 3 == 0) __builtin_amdgcn_sched_barrier(0);
      }
      __builtin_amdgcn_sched_barrier(0);
      mwm = V.r_univ / rm;
      mwmp = V.r_univ / rmp;
      mwmif = 0.5 * (mwmp + mwm) / (mwmp * mwm);
      mwmif = mwmif / dx;
      opaque(ga);
      opaque(gc);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const double rd = 0.5 * (CA(X::J_D_START + s + 1) + CC(X::J_D_START + s + 1));
        fsum = fsum + (CC(s + 1) * mwmp - CA(s + 1) * mwm) * rd;
        if (s % 3 == 2) __builtin_amdgcn_sched_barrier(0);
      }
      fsum = fsum * mwmif;
    }
    const double dxhl = upd ? 2.0 / (V.coord[f + 1 + b.ng - 1] - V.coord[f - 1 + b.ng - 1]) : 0.0;
    double et = 0.0;
#pragma unroll
    for (int s0k = 0; s0k < NS; s0k += KCH) {
      const int nk = NS - s0k < KCH ? NS - s0k : KCH;
      __syncthreads();                      // xb free
      if (valid) {
        opaque(ga);
        opaque(gc);
#pragma unroll 1
        for (int k = 0; k < nk; ++k) {
          const int s = s0k + k;
          const double ya = CA(s + 1), yc = CC(s + 1);
          const double rd = 0.5 * (CA(X::J_D_START + s + 1) + CC(X::J_D_START + s + 1));
          const double sig = -rd * (yc * mwmp - ya * mwm) * mwmif + fsum * (0.5 * (yc + ya));
          const double hh = species_h<NS, LO>(TL.cph, s, range_of<NS>(TL, s, ta), TPow<LO>(ta));
          const double hhp = species_h<NS, LO>(TL.cph, s, range_of<NS>(TL, s, tc), TPow<LO>(tc));
          et = et + sig * (0.5 * (hh + hhp));
          xb[k * T::XSLOT + me] = sig;
        }
      }
      __syncthreads();
      if (upd) {
#pragma unroll
        for (int k = 0; k < KCH; ++k)
          if (k < nk) acc[s0k + k] = acc[s0k + k] + dxhl * (xb[k * T::XSLOT + me] - xb[k * T::XSLOT + left]);
      }
    }
    double fme[4] = {-sigx, -sigy, -sigz, -sigq + et};
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int k = 0; k < 4; ++k) xb[k * T::XSLOT + me] = fme[k];
    }
    __syncthreads();
    if (upd) {
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[NS + k] = acc[NS + k] + dxhl * (fme[k] - xb[k * T::XSLOT + left]);
    }
  }
  if (!upd) return;
  const int cell = f;
  long fcell;
  if (DIR == 1) fcell = b.f(cell, col, t2);
  else if (DIR == 2) fcell = b.f(col, cell, t2);
  else fcell = b.f(col, t2, cell);
  const long fs_ = b.fstride();
  if (A.init == 1) {
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      const double old = fl[fcell + fs_ * m];
      fln[fcell + fs_ * m] = -A.rhodt * old;
      fl[fcell + fs_ * m] = acc[m];
    }
  } else if (A.init == 2) {   // fln set by the previous substep's streams_rhs_finish_next
#pragma unroll
    for (int m = 0; m < NV; ++m) fl[fcell + fs_ * m] = acc[m];
  } else {
#pragma unroll
    for (int m = 0; m < NV; ++m) fl[fcell + fs_ * m] = fl[fcell + fs_ * m] + acc[m];
  }
}

template <int NS, int DIR, int LO>
int sweep_launch(const Box &b, SweepArgs A, hipStream_t st) {
  constexpr int C = TileCfg<DIR>::C, S1 = TileCfg<DIR>::S1;
  const int ncell = A.E.fe - A.E.fs;
  if (ncell < 1) return 0;
  const int nseg = (ncell + S1 - 2) / (S1 - 1);
  const int ncol = DIR == 1 ? b.ny : b.nx;
  const int ncb = (ncol + C - 1) / C;
  const int nt2 = DIR == 3 ? b.ny : b.nz;
  const unsigned long ntiles = (unsigned long)nseg * ncb * nt2;
  if (ntiles > 0xffffff00ul) return STREAMS_E_DIMS;
  A.ncb = ncb;
  A.nseg = nseg;
  A.ntiles = (unsigned)ntiles;
  // the KEEP pair loops are unrolled up to a compile-time bound on lmax (ep_ord_change only
  // lowers the order): 3 for the sixth-order scheme, LMX for the others
  if (A.E.lmax_base == 3)
    sweep_kernel<NS, DIR, LO, 3><<<xcd_grid(A.ntiles), 256, 0, st>>>(b, A.E.w_aux, A.E.fl, A.fln, A.E.fhat, A);
  else
    sweep_kernel<NS, DIR, LO, LMX><<<xcd_grid(A.ntiles), 256, 0, st>>>(b, A.E.w_aux, A.E.fl, A.fln, A.E.fhat, A);
  return hip_status();
}

}  // namespace streams

using namespace streams;

// Fused sweep of direction dir (1..3): KEEP/WENO convective flux of faces 0..n plus the
// viscous flux (visc=1) and the fl update of cells 1..n; init=1 also performs init_flux,
// init=2 overwrites fl without it (fln already set by streams_rhs_finish_next).
extern "C" int streams_rhs_sweep(int dir, int n_s, int nv_aux, int nx, int ny, int nz, int ng, int lmax_base,
                                 int nkeep, const double *rgas, const double *w_aux, const double *coeff_deriv1,
                                 const double *metric, const double *coord, double *fhat,
                                 int force_zero_flux_min, int force_zero_flux_max, int weno_scheme,
                                 int weno_version, double sensor_threshold, const double *cv_coeff,
                                 const double *cp_coeff, int nsetcv, const double *trange, int indx_cp_l,
                                 int indx_cp_r, const int *ep_ord_change, double tol_iter_nr, double rho0, double u0,
                                 int flux_splitting, int visc, double r_univ, int init, double rhodt, double *fl,
                                 double *fln, void *stream) {
  if (dir < 1 || dir > 3) return STREAMS_E_DIMS;
  if (init < 0 || init > 2) return STREAMS_E_OPTION;
  if (nv_aux < 2 * n_s + 13 || ng < lmax_base || ng < weno_scheme) return STREAMS_E_DIMS;
  if (lmax_base < 1 || lmax_base > LMX || weno_scheme < 1 || weno_scheme > 4) return STREAMS_E_OPTION;
  if (nsetcv < 1 || nsetcv > TMAXSET) return STREAMS_E_OPTION;
  const int n = dir == 1 ? nx : dir == 2 ? ny : nz;
  if (n < 1) return 0;
  Box b{nx, ny, nz, ng};
  Thermo th{cv_coeff, cp_coeff, trange, nsetcv, indx_cp_l, indx_cp_r};
  SweepArgs A{};
  A.E = EulerArgs{w_aux, coeff_deriv1, metric, rgas, fhat, fl, ep_ord_change, th, lmax_base, nkeep,
                  force_zero_flux_min, force_zero_flux_max, weno_scheme, weno_version, flux_splitting,
                  sensor_threshold, tol_iter_nr, rho0, u0, 0, n, n};
  A.E.check_shock_flag = sensor_threshold < 1.0;
  A.V = ViscArgs{w_aux, coord, rgas, fl, th, r_univ, n};
  A.fln = fln;
  A.rhodt = rhodt;
  A.init = init;
  A.visc = visc;
  A.check_shock = A.E.check_shock_flag;
  hipStream_t st = as_stream(stream);
  if (A.check_shock) {
    int rc = weno_prepass_launch(dir, n_s, b, A.E, st);
    if (rc) return rc;
  }
  if (dir == 1) {
    STREAMS_DISPATCH_NS(n_s, STREAMS_DISPATCH_NASA(indx_cp_l, indx_cp_r, return (sweep_launch<NS, 1, LO>(b, A, st))));
  }
  if (dir == 2) {
    STREAMS_DISPATCH_NS(n_s, STREAMS_DISPATCH_NASA(indx_cp_l, indx_cp_r, return (sweep_launch<NS, 2, LO>(b, A, st))));
  }
  STREAMS_DISPATCH_NS(n_s, STREAMS_DISPATCH_NASA(indx_cp_l, indx_cp_r, return (sweep_launch<NS, 3, LO>(b, A, st))));
  return 0;
}
