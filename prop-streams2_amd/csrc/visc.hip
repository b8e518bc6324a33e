// visc.hip — viscous, conductive and multicomponent diffusive fluxes.
//
// visflx_{x,y,z}_cuf (kernels_gpu.F90:3658-4031) compute a staggered face flux into fhat
// and then difference it in a second kernel.  Here each direction is one kernel with the
// same traversal as the convective sweeps (x: flattened faces, LDS exchange; y/z: column
// march with the previous face in registers), so face fluxes stay on chip.
// visflx_reduced_cuf (:3539-3656) and the Laplacian form visflx_cuf (:3357-3537) are
// pointwise 3-D stencils, one thread per cell.
#include "visc_dev.h"

namespace streams {

constexpr int VXT = 256;

template <int NS>
__global__ void __launch_bounds__(VXT) visc_x_kernel(Box b, ViscArgs A, long nfaces) {
  constexpr int NV = NS + 4;
  __shared__ double sf[NV][VXT];
  const int nf = b.nx + 1;  // faces 0..nx
  const long q = (long)blockIdx.x * (VXT - 1) + threadIdx.x;
  const bool valid = q < nfaces;
  int f = 0, j = 1, k = 1;
  if (valid) {
    f = (int)(q % nf);
    long r = q / nf;
    j = 1 + (int)(r % b.ny);
    k = 1 + (int)(r / b.ny);
    VLine L{A.w_aux, 0, 0, b.pstride()};
    vline_base<1>(b, j, k, L.base, L.st);
    double fh[NV];
    visc_face<NS>(L, f, A, b.ng, fh);
#pragma unroll
    for (int m = 0; m < NV; ++m) sf[m][threadIdx.x] = fh[m];
  }
  __syncthreads();
  if (!valid || threadIdx.x == 0 || f == 0) return;
  const int c = f;
  const double dxhl = 2.0 / (A.coord[c + 1 + b.ng - 1] - A.coord[c - 1 + b.ng - 1]);
  const long fc = b.f(c, j, k), fs_ = b.fstride();
#pragma unroll
  for (int m = 0; m < NV; ++m)
    A.fl[fc + fs_ * m] = A.fl[fc + fs_ * m] + dxhl * (sf[m][threadIdx.x] - sf[m][threadIdx.x - 1]);
}

template <int NS, int DIR>
__global__ void __launch_bounds__(128) visc_yz_kernel(Box b, ViscArgs A, int seg, long nthreads) {
  constexpr int NV = NS + 4;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nthreads) return;
  const int t2n = DIR == 2 ? b.nz : b.ny;
  int t1 = 1 + (int)(t % b.nx);
  long r = t / b.nx;
  int t2 = 1 + (int)(r % t2n);
  int sg = (int)(r / t2n);
  int c0 = 1 + sg * seg, c1 = c0 + seg - 1;
  if (c1 > A.n) c1 = A.n;
  if (c0 > c1) return;
  VLine L{A.w_aux, 0, 0, b.pstride()};
  vline_base<DIR>(b, t1, t2, L.base, L.st);
  double prev[NV], cur[NV];
  visc_face<NS>(L, c0 - 1, A, b.ng, prev);
  const long fs_ = b.fstride();
  for (int c = c0; c <= c1; ++c) {
    visc_face<NS>(L, c, A, b.ng, cur);
    const double dxhl = 2.0 / (A.coord[c + 1 + b.ng - 1] - A.coord[c - 1 + b.ng - 1]);
    long fc = DIR == 2 ? b.f(t1, c, t2) : b.f(t1, t2, c);
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      A.fl[fc + fs_ * m] = A.fl[fc + fs_ * m] + dxhl * (cur[m] - prev[m]);
      prev[m] = cur[m];
    }
  }
}

template <int NS>
__global__ void __launch_bounds__(256) visc_reduced_kernel(Box b, int lmax, const double *__restrict__ w_aux,
                                                           double *__restrict__ fl, const double *__restrict__ cd1,
                                                           const double *__restrict__ dcsidx,
                                                           const double *__restrict__ detady,
                                                           const double *__restrict__ dzitdz) {
  using X = Idx<NS>;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.fstride()) return;
  int i = 1 + (int)(t % b.nx);
  long r = t / b.nx;
  int j = 1 + (int)(r % b.ny);
  int k = 1 + (int)(r / b.ny);
  const long pt = b.p(i, j, k), ps = b.pstride();
  const long sx = 1, sy = b.px(), sz = b.px() * b.py();
  const double *U = w_aux + ps * (X::J_U - 1), *V = w_aux + ps * (X::J_V - 1), *Wv = w_aux + ps * (X::J_W - 1);
  const double *MU = w_aux + ps * (X::J_MU - 1), *DV = w_aux + ps * (X::J_DIV - 1);
  const double uu = U[pt], vv = V[pt], ww = Wv[pt], mu = MU[pt], div3l = DV[pt];
  double ux = 0, vx = 0, wx = 0, mux = 0, dvx = 0, uy = 0, vy = 0, wy = 0, muy = 0, dvy = 0;
  double uz = 0, vz = 0, wz = 0, muz = 0, dvz = 0;
  for (int l = 1; l <= lmax; ++l) {
    const double ccl = cd1[(l - 1) + 4 * (lmax - 1)];
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
  const double dx = dcsidx[i - 1], dy = detady[j - 1], dz = dzitdz[k - 1];
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
  const long fs_ = b.fstride();
  fl[t + fs_ * (X::I_U - 1)] = fl[t + fs_ * (X::I_U - 1)] - sigx;
  fl[t + fs_ * (X::I_V - 1)] = fl[t + fs_ * (X::I_V - 1)] - sigy;
  fl[t + fs_ * (X::I_W - 1)] = fl[t + fs_ * (X::I_W - 1)] - sigz;
  fl[t + fs_ * (X::I_E - 1)] = fl[t + fs_ * (X::I_E - 1)] - sigq;
}

// visflx_cuf (conservative_viscous = 0): the whole stress divergence and conduction from
// cell-centred first and second derivatives (Laplacians with the grid-stretching term
// d2xi/dx2), viscous work and aerodynamic heating; momentum and energy only.
struct LaplArgs {
  const double *cd1, *cd2;                       // coeff_deriv1(4,4), coeff_deriv2(0:4,4)
  const double *dcsidx, *detady, *dzitdz, *dcsidxs, *detadys, *dzitdzs, *dcsidx2, *detady2, *dzitdz2;
};

template <int NS>
__global__ void __launch_bounds__(256) visc_lapl_kernel(Box b, int lmax, const double *__restrict__ w_aux,
                                                        double *__restrict__ fl, LaplArgs L) {
  using X = Idx<NS>;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.fstride()) return;
  int i = 1 + (int)(t % b.nx);
  long r = t / b.nx;
  int j = 1 + (int)(r % b.ny);
  int k = 1 + (int)(r / b.ny);
  const long pt = b.p(i, j, k), ps = b.pstride();
  const long sx = 1, sy = b.px(), sz = b.px() * b.py();
  const double *U = w_aux + ps * (X::J_U - 1), *V = w_aux + ps * (X::J_V - 1), *Wv = w_aux + ps * (X::J_W - 1);
  const double *TT = w_aux + ps * (X::J_T - 1), *MU = w_aux + ps * (X::J_MU - 1);
  const double *DV = w_aux + ps * (X::J_DIV - 1);
  const double uu = U[pt], vv = V[pt], ww = Wv[pt], tt = TT[pt], mu = MU[pt];
  const double k_cond = w_aux[pt + ps * (X::J_K_COND - 1)], div3l = DV[pt];
  const double c0 = L.cd2[0 + 5 * (lmax - 1)];
  double ulapx = c0 * uu, vlapx = c0 * vv, wlapx = c0 * ww, tlapx = c0 * tt;
  double ulapy = ulapx, vlapy = vlapx, wlapy = wlapx, tlapy = tlapx;
  double ulapz = ulapx, vlapz = vlapx, wlapz = wlapx, tlapz = tlapx;
  for (int l = 1; l <= lmax; ++l) {
    const double clapl = L.cd2[l + 5 * (lmax - 1)];
    ulapx = ulapx + clapl * (U[pt + l * sx] + U[pt - l * sx]);
    ulapy = ulapy + clapl * (U[pt + l * sy] + U[pt - l * sy]);
    ulapz = ulapz + clapl * (U[pt + l * sz] + U[pt - l * sz]);
    vlapx = vlapx + clapl * (V[pt + l * sx] + V[pt - l * sx]);
    vlapy = vlapy + clapl * (V[pt + l * sy] + V[pt - l * sy]);
    vlapz = vlapz + clapl * (V[pt + l * sz] + V[pt - l * sz]);
    wlapx = wlapx + clapl * (Wv[pt + l * sx] + Wv[pt - l * sx]);
    wlapy = wlapy + clapl * (Wv[pt + l * sy] + Wv[pt - l * sy]);
    wlapz = wlapz + clapl * (Wv[pt + l * sz] + Wv[pt - l * sz]);
    tlapx = tlapx + clapl * (TT[pt + l * sx] + TT[pt - l * sx]);
    tlapy = tlapy + clapl * (TT[pt + l * sy] + TT[pt - l * sy]);
    tlapz = tlapz + clapl * (TT[pt + l * sz] + TT[pt - l * sz]);
  }
  double ux = 0, vx = 0, wx = 0, tx = 0, mux = 0, dvx = 0, uy = 0, vy = 0, wy = 0, ty = 0, muy = 0, dvy = 0;
  double uz = 0, vz = 0, wz = 0, tz = 0, muz = 0, dvz = 0;
  for (int l = 1; l <= lmax; ++l) {
    const double ccl = L.cd1[(l - 1) + 4 * (lmax - 1)];
    ux = ux + ccl * (U[pt + l * sx] - U[pt - l * sx]);
    vx = vx + ccl * (V[pt + l * sx] - V[pt - l * sx]);
    wx = wx + ccl * (Wv[pt + l * sx] - Wv[pt - l * sx]);
    tx = tx + ccl * (TT[pt + l * sx] - TT[pt - l * sx]);
    mux = mux + ccl * (MU[pt + l * sx] - MU[pt - l * sx]);
    dvx = dvx + ccl * (DV[pt + l * sx] - DV[pt - l * sx]);
    uy = uy + ccl * (U[pt + l * sy] - U[pt - l * sy]);
    vy = vy + ccl * (V[pt + l * sy] - V[pt - l * sy]);
    wy = wy + ccl * (Wv[pt + l * sy] - Wv[pt - l * sy]);
    ty = ty + ccl * (TT[pt + l * sy] - TT[pt - l * sy]);
    muy = muy + ccl * (MU[pt + l * sy] - MU[pt - l * sy]);
    dvy = dvy + ccl * (DV[pt + l * sy] - DV[pt - l * sy]);
    uz = uz + ccl * (U[pt + l * sz] - U[pt - l * sz]);
    vz = vz + ccl * (V[pt + l * sz] - V[pt - l * sz]);
    wz = wz + ccl * (Wv[pt + l * sz] - Wv[pt - l * sz]);
    tz = tz + ccl * (TT[pt + l * sz] - TT[pt - l * sz]);
    muz = muz + ccl * (MU[pt + l * sz] - MU[pt - l * sz]);
    dvz = dvz + ccl * (DV[pt + l * sz] - DV[pt - l * sz]);
  }
  const double dx = L.dcsidx[i - 1], dy = L.detady[j - 1], dz = L.dzitdz[k - 1];
  ux *= dx; vx *= dx; wx *= dx; tx *= dx; mux *= dx; dvx *= dx;
  uy *= dy; vy *= dy; wy *= dy; ty *= dy; muy *= dy; dvy *= dy;
  uz *= dz; vz *= dz; wz *= dz; tz *= dz; muz *= dz; dvz *= dz;
  const double dxs = L.dcsidxs[i - 1], dys = L.detadys[j - 1], dzs = L.dzitdzs[k - 1];
  const double dx2 = L.dcsidx2[i - 1], dy2 = L.detady2[j - 1], dz2 = L.dzitdz2[k - 1];
  ulapx = ulapx * dxs + ux * dx2;
  vlapx = vlapx * dxs + vx * dx2;
  wlapx = wlapx * dxs + wx * dx2;
  tlapx = tlapx * dxs + tx * dx2;
  ulapy = ulapy * dys + uy * dy2;
  vlapy = vlapy * dys + vy * dy2;
  wlapy = wlapy * dys + wy * dy2;
  tlapy = tlapy * dys + ty * dy2;
  ulapz = ulapz * dzs + uz * dz2;
  vlapz = vlapz * dzs + vz * dz2;
  wlapz = wlapz * dzs + wz * dz2;
  tlapz = tlapz * dzs + tz * dz2;
  const double ulap = ulapx + ulapy + ulapz, vlap = vlapx + vlapy + vlapz;
  const double wlap = wlapx + wlapy + wlapz, tlap = tlapx + tlapy + tlapz;
  const double sig11 = 2.0 * (ux - div3l), sig12 = uy + vx, sig13 = uz + wx;
  const double sig22 = 2.0 * (vy - div3l), sig23 = vz + wy, sig33 = 2.0 * (wz - div3l);
  const double sigx = mux * sig11 + muy * sig12 + muz * sig13 + mu * (ulap + dvx);
  const double sigy = mux * sig12 + muy * sig22 + muz * sig23 + mu * (vlap + dvy);
  const double sigz = mux * sig13 + muy * sig23 + muz * sig33 + mu * (wlap + dvz);
  const double sigqt = (mux * tx + muy * ty + muz * tz + mu * tlap) * k_cond / mu;
  const double sigah = (sig11 * ux + sig12 * uy + sig13 * uz + sig12 * vx + sig22 * vy + sig23 * vz + sig13 * wx +
                        sig23 * wy + sig33 * wz) * mu;
  const double sigq = sigx * uu + sigy * vv + sigz * ww + sigah + sigqt;
  const long fs_ = b.fstride();
  fl[t + fs_ * (X::I_U - 1)] = fl[t + fs_ * (X::I_U - 1)] - sigx;
  fl[t + fs_ * (X::I_V - 1)] = fl[t + fs_ * (X::I_V - 1)] - sigy;
  fl[t + fs_ * (X::I_W - 1)] = fl[t + fs_ * (X::I_W - 1)] - sigz;
  fl[t + fs_ * (X::I_E - 1)] = fl[t + fs_ * (X::I_E - 1)] - sigq;
}

template <int NS, int DIR>
int visc_launch(const Box &b, ViscArgs A, hipStream_t st) {
  if (A.n < 1) return 0;
  if (DIR == 1) {
    long nfaces = (long)(b.nx + 1) * b.ny * b.nz;
    long nblk = (nfaces - 1 + (VXT - 2)) / (VXT - 1);
    visc_x_kernel<NS><<<(unsigned)nblk, VXT, 0, st>>>(b, A, nfaces);
  } else {
    const int seg = 16;
    const int nseg = (A.n + seg - 1) / seg;
    const int t2n = DIR == 2 ? b.nz : b.ny;
    long nth = (long)b.nx * t2n * nseg;
    visc_yz_kernel<NS, DIR><<<blocks_for(nth, 128), 128, 0, st>>>(b, A, seg, nth);
  }
  return hip_status();
}

}  // namespace streams

using namespace streams;

static int visflx_common(int dir, int n_s, int nx, int ny, int nz, int nv, int nv_aux, int ng, const double *coord,
                         const double *w_aux, double *fl, int lo, int hi, const double *cp_coeff, int nsetcv,
                         const double *trange, const double *rgas, double r_univ, void *stream) {
  if (nv != n_s + 4 || nv_aux < 2 * n_s + 13 || ng < 1) return STREAMS_E_DIMS;
  Box b{nx, ny, nz, ng};
  int n = dir == 1 ? nx : dir == 2 ? ny : nz;
  ViscArgs A{w_aux, coord, rgas, fl, Thermo{nullptr, cp_coeff, trange, nsetcv, lo, hi}, r_univ, n};
  hipStream_t st = as_stream(stream);
  if (dir == 1) { STREAMS_DISPATCH_NS(n_s, return (visc_launch<NS, 1>(b, A, st))); }
  if (dir == 2) { STREAMS_DISPATCH_NS(n_s, return (visc_launch<NS, 2>(b, A, st))); }
  STREAMS_DISPATCH_NS(n_s, return (visc_launch<NS, 3>(b, A, st)));
  return 0;
}

extern "C" int streams_visflx_x(int n_s, int nx, int ny, int nz, int nv, int nv_aux, int ng, const double *x,
                                const double *w_aux, double *fl, double *fhat, int indx_cp_l, int indx_cp_r,
                                const double *cp_coeff, int nsetcv, const double *trange, const double *rgas,
                                double r_univ, void *stream) {
  (void)fhat;
  return visflx_common(1, n_s, nx, ny, nz, nv, nv_aux, ng, x, w_aux, fl, indx_cp_l, indx_cp_r, cp_coeff, nsetcv,
                       trange, rgas, r_univ, stream);
}

extern "C" int streams_visflx_y(int n_s, int nx, int ny, int nz, int nv, int nv_aux, int ng, const double *y,
                                const double *w_aux, double *fl, double *fhat, int indx_cp_l, int indx_cp_r,
                                const double *cp_coeff, int nsetcv, const double *trange, const double *rgas,
                                double r_univ, void *stream) {
  (void)fhat;
  return visflx_common(2, n_s, nx, ny, nz, nv, nv_aux, ng, y, w_aux, fl, indx_cp_l, indx_cp_r, cp_coeff, nsetcv,
                       trange, rgas, r_univ, stream);
}

extern "C" int streams_visflx_z(int n_s, int nx, int ny, int nz, int nv, int nv_aux, int ng, const double *z,
                                const double *w_aux, double *fl, double *fhat, int indx_cp_l, int indx_cp_r,
                                const double *cp_coeff, int nsetcv, const double *trange, const double *rgas,
                                double r_univ, void *stream) {
  (void)fhat;
  return visflx_common(3, n_s, nx, ny, nz, nv, nv_aux, ng, z, w_aux, fl, indx_cp_l, indx_cp_r, cp_coeff, nsetcv,
                       trange, rgas, r_univ, stream);
}

extern "C" int streams_visflx_reduced(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order,
                                      const double *w_aux, double *fl, const double *coeff_deriv1,
                                      const double *dcsidx, const double *detady, const double *dzitdz,
                                      void *stream) {
  if (nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  int lmax = visc_order / 2;
  if (lmax < 1 || lmax > 4 || lmax > ng) return STREAMS_E_OPTION;
  Box b{nx, ny, nz, ng};
  long n = b.fstride();
  if (n == 0) return 0;
  STREAMS_DISPATCH_NS(n_s, (visc_reduced_kernel<NS><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
                                b, lmax, w_aux, fl, coeff_deriv1, dcsidx, detady, dzitdz)));
  return hip_status();
}

extern "C" int streams_visflx(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order, double *w_aux,
                             double *fl, const double *coeff_deriv1, const double *coeff_deriv2,
                             const double *dcsidx, const double *detady, const double *dzitdz,
                             const double *dcsidxs, const double *detadys, const double *dzitdzs,
                             const double *dcsidx2, const double *detady2, const double *dzitdz2, void *stream) {
  if (nv_aux < 2 * n_s + 13) return STREAMS_E_DIMS;
  int lmax = visc_order / 2;
  if (lmax < 1 || lmax > 4 || lmax > ng) return STREAMS_E_OPTION;
  Box b{nx, ny, nz, ng};
  long n = b.fstride();
  if (n == 0) return 0;
  LaplArgs L{coeff_deriv1, coeff_deriv2, dcsidx, detady, dzitdz, dcsidxs, detadys, dzitdzs, dcsidx2, detady2, dzitdz2};
  STREAMS_DISPATCH_NS(n_s, (visc_lapl_kernel<NS><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
                                b, lmax, w_aux, fl, L)));
  return hip_status();
}
