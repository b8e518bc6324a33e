// sweep.hip — fused convective + viscous sweep along one direction, LDS-tiled.
//
// One workgroup of 256 threads owns a tile of C columns x S1 faces (S1-1 cells) of one
// direction.  It stages the convective stencil variables (Y_s, u, v, w, h, p, rho) of the
// tile and its halo in LDS with coalesced loads, then each thread evaluates one face:
// the KEEP split-form flux from LDS (or the WENO flux parked in fhat by the shock
// pre-pass) and the staggered viscous / diffusive flux (the two cells of the face, the
// transport-only variables read straight from HBM).  Face fluxes meet in LDS and each
// thread updates its cell:
//     fl += (Fc(c) - Fc(c-1)) * dxi(c) + 2/(x(c+1)-x(c-1)) * (Fv(c) - Fv(c-1))
// which is euler_{x,y,z}_update_cuf followed by the update of visflx_{x,y,z}_cuf
// (kernels_gpu.F90:732-793, :3775-3788) in one read-modify-write of fl.  In the first
// sweep of a substep the RK bookkeeping of init_flux_cuf (:28-45) is folded in:
// fln = -rhodt * fl_old and fl starts from this sweep's contribution.
//
// Tiles: x (lanes along the sweep): C = 8 lines x S1 = 32 faces; y, z (lanes along i):
// C = 16 columns x S1 = 16 faces.  LDS layout keeps the lane index fastest.
#include "euler_dev.h"
#include "visc_dev.h"

namespace streams {

template <int DIR> struct TileCfg { static constexpr int C = 16, S1 = 16; };
template <> struct TileCfg<1> { static constexpr int C = 8, S1 = 32; };
constexpr int LMX = 4;  // largest half stencil (ep_order 8)

struct SweepArgs {
  EulerArgs E;
  ViscArgs V;
  double *fln;
  double rhodt;
  int init;          // fold init_flux: fln = -rhodt*fl_old, fl = contribution
  int visc;          // add the viscous flux
  int check_shock;   // sensor threshold below 1: consult the pre-pass results
};

// staged slots: Y_1..Y_NS, u, v, w, h, p, rho
template <int NS>
__device__ constexpr int slot_of(int m) {
  return m <= NS + 4 ? m - 1 : (m == NS + 6 ? NS + 4 : NS + 5);
}

template <int NS, int DIR>
struct LdsLine {
  static constexpr int C = TileCfg<DIR>::C, S1 = TileCfg<DIR>::S1, R = S1 + 2 * LMX - 1;
  const double *lds;
  int c, p0;
  __device__ inline int at(int slot, int r) const {
    return DIR == 1 ? slot * C * R + c * R + r : slot * C * R + r * C + c;
  }
  __device__ inline double operator()(int p, int m) const { return lds[at(slot_of<NS>(m), p - p0)]; }
};

// viscous stencil access: staged variables from LDS, transport-only ones from HBM
template <int NS, int DIR>
struct MixLine {
  LdsLine<NS, DIR> s;
  Line g;
  __device__ inline double operator()(int p, int m) const {
    using X = Idx<NS>;
    if (m <= NS + 3) return s(p, m);   // Y, u, v, w
    return g(p, m);                    // T, mu, k, rho*D
  }
};

template <int NS, int DIR>
__global__ void __launch_bounds__(256) sweep_kernel(Box b, SweepArgs A) {
  using X = Idx<NS>;
  constexpr int NV = NS + 4, NSLOT = NS + 6;
  constexpr int C = TileCfg<DIR>::C, S1 = TileCfg<DIR>::S1, R = S1 + 2 * LMX - 1;
  constexpr int STAGE = NSLOT * C * R;
  constexpr int EXCH = NV * C * S1;
  __shared__ double lds[STAGE > EXCH ? STAGE : EXCH];
  const EulerArgs &E = A.E;
  const int t = threadIdx.x;
  int c, fl_;
  if (DIR == 1) { fl_ = t % S1; c = t / S1; }
  else { c = t % C; fl_ = t / C; }
  // tile coordinates
  const int us = E.fs + 1, ue = E.fe;
  const int seg = DIR == 1 ? blockIdx.x : blockIdx.y;
  const int s0 = us + seg * (S1 - 1);
  const int cbase = 1 + (DIR == 1 ? blockIdx.y : blockIdx.x) * C;
  const int t2 = 1 + blockIdx.z;
  const int ncol = DIR == 1 ? b.ny : b.nx;
  const int lmax_b = E.lmax_base;
  const int p0 = s0 - LMX;
  // ---- stage the convective variables of the tile
  {
    const int plo = s0 - lmax_b;
    int phi = s0 + S1 - 2;
    phi = (phi > ue ? ue : phi) + lmax_b;
    const long ps = b.pstride();
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
      const int li = DIR == 1 ? cc * R + r : r * C + cc;
#pragma unroll
      for (int sl = 0; sl < NSLOT; ++sl) {
        const int m = sl < NS + 4 ? sl + 1 : (sl == NS + 4 ? X::J_P : X::J_R);
        lds[sl * C * R + li] = E.w_aux[g + ps * (m - 1)];
      }
    }
  }
  __syncthreads();
  // ---- one face per thread
  const int col = cbase + c;
  const int f = s0 - 1 + fl_;
  const bool valid = col <= ncol && f <= ue;
  double fc[NV], fv[NV];
  if (valid) {
    Line G{E.w_aux, 0, 0, b.pstride()};
    line_base<DIR>(b, col, t2, G.base, G.st);
    LdsLine<NS, DIR> L{lds, c, p0};
    if (E.check_shock_flag && shocked<NS>(G, f, E.weno_scheme, E.sensor_threshold)) {
#pragma unroll
      for (int m = 0; m < NV; ++m) fc[m] = E.fhat[G.base + (long)f * G.st + b.pstride() * m];
    } else {
      int lmax = lmax_b + epo_at<DIR>(b, E.epo, col, t2, f);
      lmax = lmax < 1 ? 1 : lmax;
      const bool zero = (f == 0 && E.zmin == 1) || (f == E.n && E.zmax == 1);
      keep_flux<NS, DIR>(L, f, lmax, E.nkeep, E.cd1, zero, fc);
    }
    if (A.visc) {
      MixLine<NS, DIR> M{L, G};
      visc_face<NS>(M, f, A.V, b.ng, fv);
    } else {
#pragma unroll
      for (int m = 0; m < NV; ++m) fv[m] = 0.0;
    }
  }
  __syncthreads();
  // ---- convective flux exchange
  const int me = DIR == 1 ? c * S1 + fl_ : fl_ * C + c;
  const int left = DIR == 1 ? me - 1 : me - C;
  if (valid) {
#pragma unroll
    for (int m = 0; m < NV; ++m) lds[m * C * S1 + me] = fc[m];
  }
  __syncthreads();
  const bool upd = valid && fl_ >= 1;
  double acc[NV];
  const int cell = f;
  if (upd) {
    const double met = E.metric[cell - 1];
#pragma unroll
    for (int m = 0; m < NV; ++m) acc[m] = (fc[m] - lds[m * C * S1 + left]) * met;
  }
  if (A.visc) {
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int m = 0; m < NV; ++m) lds[m * C * S1 + me] = fv[m];
    }
    __syncthreads();
    if (upd) {
      const double *x = A.V.coord;
      const double dxhl = 2.0 / (x[cell + 1 + b.ng - 1] - x[cell - 1 + b.ng - 1]);
#pragma unroll
      for (int m = 0; m < NV; ++m) acc[m] = acc[m] + dxhl * (fv[m] - lds[m * C * S1 + left]);
    }
  }
  if (!upd) return;
  long fcell;
  if (DIR == 1) fcell = b.f(cell, col, t2);
  else if (DIR == 2) fcell = b.f(col, cell, t2);
  else fcell = b.f(col, t2, cell);
  const long fs_ = b.fstride();
  if (A.init) {
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      const double old = E.fl[fcell + fs_ * m];
      A.fln[fcell + fs_ * m] = -A.rhodt * old;
      E.fl[fcell + fs_ * m] = acc[m];
    }
  } else {
#pragma unroll
    for (int m = 0; m < NV; ++m) E.fl[fcell + fs_ * m] = E.fl[fcell + fs_ * m] + acc[m];
  }
}

template <int NS, int DIR>
int sweep_launch(const Box &b, const SweepArgs &A, hipStream_t st) {
  constexpr int C = TileCfg<DIR>::C, S1 = TileCfg<DIR>::S1;
  const int ncell = A.E.fe - A.E.fs;
  if (ncell < 1) return 0;
  const int nseg = (ncell + S1 - 2) / (S1 - 1);
  const int ncol = DIR == 1 ? b.ny : b.nx;
  const int ncb = (ncol + C - 1) / C;
  const int nt2 = DIR == 3 ? b.ny : b.nz;
  dim3 grid = DIR == 1 ? dim3(nseg, ncb, nt2) : dim3(ncb, nseg, nt2);
  sweep_kernel<NS, DIR><<<grid, 256, 0, st>>>(b, A);
  return hip_status();
}

}  // namespace streams

using namespace streams;

// Fused sweep of direction dir (1..3): KEEP/WENO convective flux of faces 0..n plus the
// viscous flux (visc=1) and the fl update of cells 1..n; init=1 also performs init_flux.
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
  if (nv_aux < 2 * n_s + 13 || ng < lmax_base || ng < weno_scheme) return STREAMS_E_DIMS;
  if (lmax_base < 1 || lmax_base > LMX || weno_scheme < 1 || weno_scheme > 4) return STREAMS_E_OPTION;
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
  if (dir == 1) { STREAMS_DISPATCH_NS(n_s, return (sweep_launch<NS, 1>(b, A, st))); }
  if (dir == 2) { STREAMS_DISPATCH_NS(n_s, return (sweep_launch<NS, 2>(b, A, st))); }
  STREAMS_DISPATCH_NS(n_s, return (sweep_launch<NS, 3>(b, A, st)));
  return 0;
}
