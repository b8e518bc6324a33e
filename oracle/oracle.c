/*
 * oracle.c — CPU FP64 restatement of the STREAmS-2 multideal hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU kernels are compared
 * against (and the timed CPU baseline of bench.py).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load liboracle.so; the product never does.
 *
 * Every function restates one routine of the reference statement by statement, in the
 * serial loop order of the reference's generated CPU backend (k outermost, i innermost,
 * tools/sutils/backend_generation/backend_config.toml), in FP64, and carries the same
 * argument list as the matching streams_* entry point of include/streams_hip.h with the
 * prefix orc_ (the stream argument is accepted and ignored).  Citations are
 * path:line in the reference tree.
 *
 * Pinning (which check fixes each function) is listed in the test files; functions whose
 * only check is agreement with the GPU path are marked "parity unpinned" there.
 */
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define NSMAX 64
#define NVMAX (NSMAX + 4)

/* liboracle_omp.so (CPU-baseline timing only) splits the outermost k / column loops over
 * threads, one z-slab per thread as the reference's CPU backend runs one rank per core;
 * every cell is still computed by the same serial statements. */
#ifdef ORACLE_OMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#else
#define OMP_FOR
#endif

/* padded field (1-ng:nx+ng, 1-ng:ny+ng, 1-ng:nz+ng, 1:nv) */
#define PX (nx + 2 * ng)
#define PY (ny + 2 * ng)
#define PZ (nz + 2 * ng)
#define PIDX(i, j, k) ((size_t)((i) + ng - 1) + (size_t)PX * ((size_t)((j) + ng - 1) + (size_t)PY * (size_t)((k) + ng - 1)))
#define PSTRIDE ((size_t)PX * PY * PZ)
#define W(a, i, j, k, m) (a)[PIDX(i, j, k) + PSTRIDE * (size_t)((m) - 1)]
/* interior field (1:nx, 1:ny, 1:nz, 1:nv) */
#define FIDX(i, j, k) ((size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1)))
#define FSTRIDE ((size_t)nx * ny * nz)
#define F(a, i, j, k, m) (a)[FIDX(i, j, k) + FSTRIDE * (size_t)((m) - 1)]

/* variable indices, src/multideal/index_define.h */
#define I_U (ns + 1)
#define I_V (ns + 2)
#define I_W (ns + 3)
#define I_E (ns + 4)
#define J_U (ns + 1)
#define J_V (ns + 2)
#define J_W (ns + 3)
#define J_H (ns + 4)
#define J_T (ns + 5)
#define J_P (ns + 6)
#define J_C (ns + 7)
#define J_MU (ns + 8)
#define J_DUC (ns + 9)
#define J_DIV (ns + 10)
#define J_R (ns + 11)
#define J_Z (ns + 12)
#define J_K_COND (ns + 13)
#define J_D_START (ns + 13)
#define J_LES1 (14 + 2 * ns)
#define J_WDOT_START (14 + 2 * ns)
#define J_HRR (15 + 3 * ns)
#define J_GAM_PASR (15 + 3 * ns)

/* tables */
#define NCOL (indx_cp_r + 2 - indx_cp_l + 1)
#define CPC(tab, l, s, m) (tab)[(size_t)((l) - indx_cp_l) + (size_t)NCOL * ((size_t)((s) - 1) + (size_t)ns * (size_t)((m) - 1))]
#define TRANGE(s, m) trange[(size_t)((s) - 1) + (size_t)ns * (size_t)((m) - 1)]
#define TAB2(tab, itt, s) (tab)[(size_t)((itt) - 1) + (size_t)(num_t_tab + 1) * (size_t)((s) - 1)]
#define TAB3(tab, itt, l, m) (tab)[(size_t)((itt) - 1) + (size_t)(num_t_tab + 1) * ((size_t)((l) - 1) + (size_t)ns * (size_t)((m) - 1))]
#define RX(tab, lr, c) (tab)[(size_t)((lr) - 1) + (size_t)nreactions * (size_t)((c) - 1)]
#define KC(itt, lr) kc_tab[(size_t)((itt) - 1) + (size_t)(num_t_tab + 1) * (size_t)((lr) - 1)]
#define CD1(l, lmax) coeff_deriv1[(size_t)((l) - 1) + 4 * (size_t)((lmax) - 1)]
#define EPO(i, j, k, d) ep_ord_change[(size_t)(i) + (size_t)(nx + 1) * ((size_t)(j) + (size_t)(ny + 1) * ((size_t)(k) + (size_t)(nz + 1) * (size_t)((d) - 1)))]

static inline double ipow(double x, int n)
{
    /* Fortran x**n for an integer exponent */
    if (n == 0) return 1.0;
    if (n < 0) return 1.0 / ipow(x, -n);
    double r = 1.0;
    for (int q = 0; q < n; ++q) r *= x;
    return r;
}

/* Numerical-Recipes locate on the species range table, clamped to [1,nsetcv]
 * (kernels_gpu.F90:6400-6418). */
static inline int locate_range(double tt, const double *trange, int ns, int nsetcv, int lsp)
{
    if (nsetcv <= 1) return 1;
    int jl = 0, ju = nsetcv + 2;
    while (ju - jl > 1) {
        int jm = (ju + jl) / 2;
        if (tt >= TRANGE(lsp, jm)) jl = jm; else ju = jm;
    }
    if (jl < 1) jl = 1;
    if (jl > nsetcv) jl = nsetcv;
    return jl;
}

/* get_mixture_temperature_from_e_dev, kernels_gpu.F90:6532-6622 (yy = mass fractions) */
static double temperature_from_e(int ns, double ee, double t_start, const double *cv_coeff,
                                 int nsetcv, const double *trange, int indx_cp_l, int indx_cp_r,
                                 double tol_iter_nr, const double *yy)
{
    int nrange[NSMAX];
    for (int lsp = 1; lsp <= ns; ++lsp) nrange[lsp - 1] = locate_range(t_start, trange, ns, nsetcv, lsp);
    double t_old = t_start, tt = t_start;
    for (int iter = 1; iter <= 10; ++iter) {
        if (nsetcv > 1) {
            for (int lsp = 1; lsp <= ns; ++lsp) {
                int nr = nrange[lsp - 1];
                if (t_old < TRANGE(lsp, nr)) nrange[lsp - 1] = nr - 1;
                else if (t_old > TRANGE(lsp, nr + 1)) nrange[lsp - 1] = nr + 1;
                if (nrange[lsp - 1] < 1) nrange[lsp - 1] = 1;
                if (nrange[lsp - 1] > nsetcv) nrange[lsp - 1] = nsetcv;
            }
        }
        double sumb = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) sumb = sumb + CPC(cv_coeff, indx_cp_r + 1, lsp, nrange[lsp - 1]) * yy[lsp - 1];
        double ebar = ee - sumb;
        double den = 0.0, num = 0.0;
        for (int l = indx_cp_l; l <= indx_cp_r; ++l) {
            double t_pow = ipow(t_old, l);
            double tden = t_pow, tnum;
            if (l == -1) tnum = log(t_old);
            else tnum = (t_old * t_pow) / (l + 1.0);
            double cv_l = 0.0;
            for (int lsp = 1; lsp <= ns; ++lsp) cv_l = cv_l + CPC(cv_coeff, l, lsp, nrange[lsp - 1]) * yy[lsp - 1];
            num = num + cv_l * tnum;
            den = den + cv_l * tden;
        }
        tt = t_old + (ebar - num) / den;
        if (fabs(tt - t_old) < tol_iter_nr) break;
        t_old = tt;
    }
    return tt;
}

/* get_cp_dev, kernels_gpu.F90:6384-6432 */
static double mixture_cp(int ns, double tt, int indx_cp_l, int indx_cp_r, const double *cp_coeff,
                         int nsetcv, const double *trange, const double *yy)
{
    int nrange[NSMAX];
    for (int lsp = 1; lsp <= ns; ++lsp) nrange[lsp - 1] = locate_range(tt, trange, ns, nsetcv, lsp);
    double cploc = 0.0;
    for (int l = indx_cp_l; l <= indx_cp_r; ++l) {
        double tpow = ipow(tt, l);
        double cp_l = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) cp_l = cp_l + CPC(cp_coeff, l, lsp, nrange[lsp - 1]) * yy[lsp - 1];
        cploc = cploc + cp_l * tpow;
    }
    return cploc;
}

/* get_species_h_from_temperature_dev (kernels_gpu.F90:7144-7182), and with cv_coeff
 * get_species_e_from_temperature_dev (:7104-7142) */
static double species_h(int ns, double tt, int indx_cp_l, int indx_cp_r, const double *cp_coeff,
                        int nsetcv, const double *trange, int isp)
{
    int nr = locate_range(tt, trange, ns, nsetcv, isp);
    double hh = CPC(cp_coeff, indx_cp_r + 1, isp, nr);
    for (int l = indx_cp_l; l <= indx_cp_r; ++l) {
        if (l == -1) hh = hh + CPC(cp_coeff, l, isp, nr) * log(tt);
        else hh = hh + CPC(cp_coeff, l, isp, nr) / (l + 1.0) * ipow(tt, l + 1);
    }
    return hh;
}

/* ===================================================================================== */
/* a1: eval_aux_cuf, kernels_gpu.F90:6154-6308                                            */
/* ===================================================================================== */
int orc_eval_aux(int ns, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                 int istart, int iend, int jstart, int jend, int kstart, int kend,
                 double *w, double *w_aux, double p0, double t_min_tab, double dt_tab,
                 double r_univ, const double *rgas, const double *cv_coeff,
                 const double *cp_coeff, int nsetcv, const double *trange,
                 int indx_cp_l, int indx_cp_r, double tol_iter_nr, void *stream,
                 const double *mw, const double *mwinv, const double *visc_species,
                 const double *lambda_species, const double *diffbin_species,
                 int num_t_tab, int n_eoi, const double *aw_eoi, const int *nainsp,
                 const double *beta0, const double *coeff_eoi)
{
    (void)nv; (void)nv_aux; (void)stream;
    OMP_FOR
    for (int k = kstart; k <= kend; ++k)
    for (int j = jstart; j <= jend; ++j)
    for (int i = istart; i <= iend; ++i) {
        double yy[NSMAX];
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double v = W(w, i, j, k, lsp);
            W(w, i, j, k, lsp) = v > 0.0 ? v : 0.0;   /* max(w,0) */
        }
        double rho = W(w, i, j, k, 1);
        for (int lsp = 2; lsp <= ns; ++lsp) rho = rho + W(w, i, j, k, lsp);
        double ri = 1.0 / rho;
        for (int lsp = 1; lsp <= ns; ++lsp) {
            W(w_aux, i, j, k, lsp) = W(w, i, j, k, lsp) * ri;
            yy[lsp - 1] = W(w_aux, i, j, k, lsp);
        }
        double rhou = W(w, i, j, k, I_U), rhov = W(w, i, j, k, I_V);
        double rhow = W(w, i, j, k, I_W), rhoe = W(w, i, j, k, I_E);
        double uu = rhou * ri, vv = rhov * ri, ww = rhow * ri;
        double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
        double ee = rhoe * ri - qq;
        double tt = temperature_from_e(ns, ee, W(w_aux, i, j, k, J_T), cv_coeff, nsetcv, trange,
                                       indx_cp_l, indx_cp_r, tol_iter_nr, yy);
        double rmixtloc = rgas[0] * yy[0];
        for (int lsp = 2; lsp <= ns; ++lsp) rmixtloc = rmixtloc + rgas[lsp - 1] * yy[lsp - 1];
        double cploc = mixture_cp(ns, tt, indx_cp_l, indx_cp_r, cp_coeff, nsetcv, trange, yy);
        double gamloc = cploc / (cploc - rmixtloc);
        double pp = rho * tt * rmixtloc;
        double c = sqrt(gamloc * pp * ri);
        W(w_aux, i, j, k, J_R) = rho;
        W(w_aux, i, j, k, J_U) = uu;
        W(w_aux, i, j, k, J_V) = vv;
        W(w_aux, i, j, k, J_W) = ww;
        W(w_aux, i, j, k, J_H) = (rhoe + pp) * ri;
        W(w_aux, i, j, k, J_T) = tt;
        W(w_aux, i, j, k, J_P) = pp;
        W(w_aux, i, j, k, J_C) = c;
        /* mixture fraction (Bilger) */
        double beta = 0.0;
        if (n_eoi != 1) {
            for (int le = 1; le <= n_eoi; ++le) {
                double yatm = 0.0;
                for (int lsp = 1; lsp <= ns; ++lsp)
                    if (yy[lsp - 1] > 1e-14)
                        yatm = yatm + nainsp[(lsp - 1) + (size_t)ns * (le - 1)] * aw_eoi[le - 1] * yy[lsp - 1] * mwinv[lsp - 1];
                beta = beta + coeff_eoi[le - 1] * yatm / aw_eoi[le - 1];
            }
            W(w_aux, i, j, k, J_Z) = (beta - beta0[1]) / (beta0[0] - beta0[1]);
        } else {
            W(w_aux, i, j, k, J_Z) = 0.5;
        }
        int itt = (int)((tt - t_min_tab) / dt_tab) + 1;
        if (itt < 1) itt = 1;
        if (itt > num_t_tab) itt = num_t_tab;
        double tloc = t_min_tab + (itt - 1) * dt_tab;
        double dtt = (tt - tloc) / dt_tab;
        /* Mathur mixture conductivity */
        double mwmixt = r_univ / rmixtloc;
        double k_cond = 0.0, k_cond2 = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double dlam = TAB2(lambda_species, itt + 1, lsp) - TAB2(lambda_species, itt, lsp);
            double klsp = TAB2(lambda_species, itt, lsp) + dlam * dtt;
            double xlsp = yy[lsp - 1] * mwmixt * mwinv[lsp - 1];
            k_cond = k_cond + xlsp * klsp;
            k_cond2 = k_cond2 + xlsp / klsp;
        }
        k_cond = 0.5 * (k_cond + 1.0 / k_cond2);
        /* Wilke mixture viscosity */
        double mu = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double dmu = TAB2(visc_species, itt + 1, lsp) - TAB2(visc_species, itt, lsp);
            double mulsp = TAB2(visc_species, itt, lsp) + dmu * dtt;
            double mu_den = 0.0;
            for (int msp = 1; msp <= ns; ++msp) {
                dmu = TAB2(visc_species, itt + 1, msp) - TAB2(visc_species, itt, msp);
                double mumsp = TAB2(visc_species, itt, msp) + dmu * dtt;
                double t1 = 1.0 + sqrt(mumsp / mulsp) * pow(mw[msp - 1] * mwinv[lsp - 1], 0.25);
                double phi_lm = 1.0 / sqrt(1.0 + mw[lsp - 1] * mwinv[msp - 1]) * (t1 * t1);
                mu_den = mu_den + yy[msp - 1] * phi_lm * mwinv[msp - 1];
            }
            mu = mu + mulsp * yy[lsp - 1] * mwinv[lsp - 1] / mu_den;
        }
        mu = sqrt(8.0) * mu;
        W(w_aux, i, j, k, J_MU) = mu;
        W(w_aux, i, j, k, J_K_COND) = k_cond;
        /* Bird species diffusion into the mixture */
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double diff_den = 0.0;
            for (int msp = 1; msp <= ns; ++msp) {
                if (msp != lsp) {
                    double ddiff = TAB3(diffbin_species, itt + 1, lsp, msp) - TAB3(diffbin_species, itt, lsp, msp);
                    double diff_ij = TAB3(diffbin_species, itt, lsp, msp) + ddiff * dtt;
                    double xmsp = yy[msp - 1] * mwmixt * mwinv[msp - 1];
                    diff_den = diff_den + xmsp / diff_ij;
                }
            }
            if (diff_den > 1.0e-15)
                W(w_aux, i, j, k, J_D_START + lsp) = (1.0 - yy[lsp - 1]) / diff_den * p0 / pp * rho;
            else
                W(w_aux, i, j, k, J_D_START + lsp) = 0.0;
        }
    }
    return 0;
}

/* ===================================================================================== */
/* a2: eval_aux2_cuf, kernels_gpu.F90:6310-6382                                           */
/* ===================================================================================== */
int orc_eval_aux2(int ns, int nx, int ny, int nz, int ng, int nv_aux, int visc_order,
                  double *w_aux, const double *coeff_deriv1, const double *dcsidx,
                  const double *detady, const double *dzitdz, double eps_sensor,
                  int sensor_type, void *stream)
{
    (void)nv_aux; (void)stream;
    int lmax = visc_order / 2;
    OMP_FOR
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        double ux = 0, vx = 0, wx = 0, uy = 0, vy = 0, wy = 0, uz = 0, vz = 0, wz = 0;
        for (int l = 1; l <= lmax; ++l) {
            double ccl = CD1(l, lmax);
            ux = ux + ccl * (W(w_aux, i + l, j, k, J_U) - W(w_aux, i - l, j, k, J_U));
            vx = vx + ccl * (W(w_aux, i + l, j, k, J_V) - W(w_aux, i - l, j, k, J_V));
            wx = wx + ccl * (W(w_aux, i + l, j, k, J_W) - W(w_aux, i - l, j, k, J_W));
            uy = uy + ccl * (W(w_aux, i, j + l, k, J_U) - W(w_aux, i, j - l, k, J_U));
            vy = vy + ccl * (W(w_aux, i, j + l, k, J_V) - W(w_aux, i, j - l, k, J_V));
            wy = wy + ccl * (W(w_aux, i, j + l, k, J_W) - W(w_aux, i, j - l, k, J_W));
            uz = uz + ccl * (W(w_aux, i, j, k + l, J_U) - W(w_aux, i, j, k - l, J_U));
            vz = vz + ccl * (W(w_aux, i, j, k + l, J_V) - W(w_aux, i, j, k - l, J_V));
            wz = wz + ccl * (W(w_aux, i, j, k + l, J_W) - W(w_aux, i, j, k - l, J_W));
        }
        ux *= dcsidx[i - 1]; vx *= dcsidx[i - 1]; wx *= dcsidx[i - 1];
        uy *= detady[j - 1]; vy *= detady[j - 1]; wy *= detady[j - 1];
        uz *= dzitdz[k - 1]; vz *= dzitdz[k - 1]; wz *= dzitdz[k - 1];
        double div = ux + vy + wz;
        double div3l = div / 3.0;
        double omegax = wy - vz, omegay = uz - wx, omegaz = vx - uy;
        double omod2 = omegax * omegax + omegay * omegay + omegaz * omegaz;
        double s = -div / sqrt(omod2 + div * div + eps_sensor);
        if (sensor_type != 0 && s < 0.0) s = 0.0;
        W(w_aux, i, j, k, J_DUC) = s * s;
        W(w_aux, i, j, k, J_DIV) = div3l;
    }
    return 0;
}

/* ===================================================================================== */
/* a3/a11: ghost cells                                                                    */
/* ===================================================================================== */
int orc_bcextr_var(int nx, int ny, int nz, int ng, double *w_var, void *stream)
{
    (void)stream;
    double *w = w_var;
#define V(i, j, k) w[PIDX(i, j, k)]
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int l = 1; l <= ng; ++l) V(1 - l, j, k) = V(1, j, k);
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int l = 1; l <= ng; ++l) V(nx + l, j, k) = V(nx, j, k);
    for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) for (int l = 1; l <= ng; ++l) V(i, 1 - l, k) = V(i, 1, k);
    for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) for (int l = 1; l <= ng; ++l) V(i, ny + l, k) = V(i, ny, k);
    for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) for (int l = 1; l <= ng; ++l) V(i, j, 1 - l) = V(i, j, 1);
    for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) for (int l = 1; l <= ng; ++l) V(i, j, nz + l) = V(i, j, nz);
#undef V
    return 0;
}

/* base_gpu%bcswap for a single rank (base_gpu.F90:1132-1196 with bcswap_step_1_cuf
 * :155-207 and bcswap_step_3_cuf :297-383): pack all six face slabs, copy between the
 * buffers of a periodic axis, unpack. */
int orc_bcswap(int nx, int ny, int nz, int ng, int nv, double *w, const int *periodic, void *stream)
{
    (void)stream;
    size_t nxb = (size_t)ng * ny * nz * nv, nyb = (size_t)nx * ng * nz * nv, nzb = (size_t)nx * ny * ng * nv;
    double *b1 = malloc(nxb * sizeof(double)), *b2 = malloc(nxb * sizeof(double));
    double *b3 = malloc(nyb * sizeof(double)), *b4 = malloc(nyb * sizeof(double));
    double *b5 = malloc(nzb * sizeof(double)), *b6 = malloc(nzb * sizeof(double));
    if (!b1 || !b2 || !b3 || !b4 || !b5 || !b6) { free(b1); free(b2); free(b3); free(b4); free(b5); free(b6); return -4; }
    size_t q;
    /* step 1: pack */
    q = 0;
    for (int m = 1; m <= nv; ++m) for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= ng; ++i, ++q) {
        b1[q] = W(w, i, j, k, m); b2[q] = W(w, nx - ng + i, j, k, m);
    }
    q = 0;
    for (int m = 1; m <= nv; ++m) for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ng; ++j) for (int i = 1; i <= nx; ++i, ++q) {
        b3[q] = W(w, i, j, k, m); b4[q] = W(w, i, ny - ng + j, k, m);
    }
    q = 0;
    for (int m = 1; m <= nv; ++m) for (int k = 1; k <= ng; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i, ++q) {
        b5[q] = W(w, i, j, k, m); b6[q] = W(w, i, j, nz - ng + k, m);
    }
    /* step 2 (self neighbour: wbuf2r = wbuf1s, wbuf1r = wbuf2s) and step 3 unpack */
    if (periodic[0]) {
        q = 0;
        for (int m = 1; m <= nv; ++m) for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= ng; ++i, ++q) {
            W(w, i - ng, j, k, m) = b2[q]; W(w, nx + i, j, k, m) = b1[q];
        }
    }
    if (periodic[1]) {
        q = 0;
        for (int m = 1; m <= nv; ++m) for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ng; ++j) for (int i = 1; i <= nx; ++i, ++q) {
            W(w, i, j - ng, k, m) = b4[q]; W(w, i, ny + j, k, m) = b3[q];
        }
    }
    if (periodic[2]) {
        q = 0;
        for (int m = 1; m <= nv; ++m) for (int k = 1; k <= ng; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i, ++q) {
            W(w, i, j, k - ng, m) = b6[q]; W(w, i, j, nz + k, m) = b5[q];
        }
    }
    free(b1); free(b2); free(b3); free(b4); free(b5); free(b6);
    return 0;
}

/* bcextr_cuf, kernels_gpu.F90:5209-5288 */
int orc_bcextr(int ilat, int nx, int ny, int nz, int ng, int nv, double *w, void *stream)
{
    (void)stream;
    for (int m = 1; m <= nv; ++m) {
        if (ilat == 1) { for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int l = 1; l <= ng; ++l) W(w, 1 - l, j, k, m) = W(w, 1, j, k, m); }
        else if (ilat == 2) { for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int l = 1; l <= ng; ++l) W(w, nx + l, j, k, m) = W(w, nx, j, k, m); }
        else if (ilat == 3) { for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) for (int l = 1; l <= ng; ++l) W(w, i, 1 - l, k, m) = W(w, i, 1, k, m); }
        else if (ilat == 4) { for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) for (int l = 1; l <= ng; ++l) W(w, i, ny + l, k, m) = W(w, i, ny, k, m); }
        else if (ilat == 5) { for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) for (int l = 1; l <= ng; ++l) W(w, i, j, 1 - l, m) = W(w, i, j, 1, m); }
        else if (ilat == 6) { for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) for (int l = 1; l <= ng; ++l) W(w, i, j, nz + l, m) = W(w, i, j, nz, m); }
    }
    return 0;
}

/* bcwall_ad_cuf, kernels_gpu.F90:5749-5852 */
int orc_bcwall_ad(int ns, int ilat, int nx, int ny, int nz, int ng, double *w, void *stream)
{
    (void)stream;
    int nv = ns + 4;
    for (int m = 1; m <= nv; ++m) {
        double sgn = (m == I_U || m == I_V || m == I_W) ? -1.0 : 1.0;
        if (ilat == 1) { for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int l = 1; l <= ng; ++l) W(w, 1 - l, j, k, m) = sgn * W(w, l, j, k, m); }
        else if (ilat == 2) { for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int l = 1; l <= ng; ++l) W(w, nx + l, j, k, m) = sgn * W(w, nx + 1 - l, j, k, m); }
        else if (ilat == 3) { for (int k = 1; k <= nz; ++k) for (int l = 1; l <= ng; ++l) for (int i = 1; i <= nx; ++i) W(w, i, 1 - l, k, m) = sgn * W(w, i, l, k, m); }
        else if (ilat == 4) { for (int k = 1; k <= nz; ++k) for (int l = 1; l <= ng; ++l) for (int i = 1; i <= nx; ++i) W(w, i, ny + l, k, m) = sgn * W(w, i, ny + 1 - l, k, m); }
        else if (ilat == 5) { for (int l = 1; l <= ng; ++l) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) W(w, i, j, 1 - l, m) = sgn * W(w, i, j, l, m); }
        else if (ilat == 6) { for (int l = 1; l <= ng; ++l) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) W(w, i, j, nz + l, m) = sgn * W(w, i, j, nz + 1 - l, m); }
    }
    return 0;
}

/* ===================================================================================== */
/* a4/a5: hybrid KEEP / WENO convective flux                                              */
/* ===================================================================================== */

/* Access to a line of w_aux along direction dir through the face index: the cell at
 * line position p (the face between p and p+1 is face p). */
typedef struct {
    const double *w_aux;
    int nx, ny, nz, ng, ns;
    int dir;           /* 1=x, 2=y, 3=z */
    int i, j, k;       /* the two transverse indices are used, the sweep one ignored */
} line_t;

static inline double LA(const line_t *L, int p, int m)
{
    int nx = L->nx, ny = L->ny, nz = L->nz, ng = L->ng;
    (void)nz;
    if (L->dir == 1) return W(L->w_aux, p, L->j, L->k, m);
    if (L->dir == 2) return W(L->w_aux, L->i, p, L->k, m);
    return W(L->w_aux, L->i, L->j, p, m);
}

/* KEEP split-form central face flux (kernels_gpu.F90:150-337 for x; :1490-1690 y;
 * :858-1055 z).  Velocity roles: un = normal component. */
static void keep_flux(const line_t *L, int f, int lmax, int nkeep, const double *coeff_deriv1,
                      int zero_flux, double *fh)
{
    int ns = L->ns;
    int jn = (L->dir == 1) ? J_U : (L->dir == 2) ? J_V : J_W;   /* normal velocity */
    int in_mom = (L->dir == 1) ? I_U : (L->dir == 2) ? I_V : I_W;
    for (int lsp = 1; lsp <= ns; ++lsp) {
        double ft1 = 0.0;
        for (int l = 1; l <= lmax; ++l) {
            double uvs1 = 0.0;
            for (int m = 0; m <= l - 1; ++m) {
                double rhoi = LA(L, f - m, J_R), uni = LA(L, f - m, jn), yyi = LA(L, f - m, lsp);
                double rhoip = LA(L, f - m + l, J_R), unip = LA(L, f - m + l, jn), yyip = LA(L, f - m + l, lsp);
                double rhom = rhoi + rhoip;
                double drhof;
                if (nkeep <= 0) drhof = 1.0;
                else {
                    double drho = 2.0 * (rhoip - rhoi) / rhom;
                    double t_s = (0.5 * drho) * (0.5 * drho);
                    double t2 = t_s;
                    double sden = 1.0 + t_s / 3.0;
                    for (int n = 2; n <= nkeep; ++n) { t_s = t2 * t_s; sden = sden + t_s / (1.0 + 2 * n); }
                    drhof = 1.0 / sden;
                }
                double uv_part = (uni + unip) * rhom * drhof;
                uvs1 = uvs1 + uv_part * (yyi + yyip);
            }
            ft1 = ft1 + CD1(l, lmax) * uvs1;
        }
        double fh1 = 0.25 * ft1;
        if (zero_flux) fh1 = 0.0;
        fh[lsp - 1] = fh1;
    }
    double ft2 = 0, ft3 = 0, ft4 = 0, ft5 = 0, ft6 = 0;
    for (int l = 1; l <= lmax; ++l) {
        double uvs2 = 0, uvs3 = 0, uvs4 = 0, uvs5 = 0, uvs5_i = 0, uvs5_k = 0, uvs5_p = 0, uvs6 = 0;
        for (int m = 0; m <= l - 1; ++m) {
            double rhoi = LA(L, f - m, J_R), uui = LA(L, f - m, J_U), vvi = LA(L, f - m, J_V), wwi = LA(L, f - m, J_W);
            double enti = LA(L, f - m, J_H), ppi = LA(L, f - m, J_P);
            double rhoip = LA(L, f - m + l, J_R), uuip = LA(L, f - m + l, J_U), vvip = LA(L, f - m + l, J_V), wwip = LA(L, f - m + l, J_W);
            double entip = LA(L, f - m + l, J_H), ppip = LA(L, f - m + l, J_P);
            double uni = LA(L, f - m, jn), unip = LA(L, f - m + l, jn);
            double rhom = rhoi + rhoip;
            if (nkeep >= 0) {
                double eei = enti - ppi / rhoi - 0.5 * (uui * uui + vvi * vvi + wwi * wwi);
                double eeip = entip - ppip / rhoip - 0.5 * (uuip * uuip + vvip * vvip + wwip * wwip);
                double eem = eei + eeip;
                double drhof, deef;
                if (nkeep == 0) { drhof = 1.0; deef = 1.0; }
                else {
                    double drho = 2.0 * (rhoip - rhoi) / rhom;
                    double dee = 2.0 * (eeip - eei) / eem;
                    double tr = (0.5 * drho) * (0.5 * drho);
                    double te = (0.5 * dee) * (0.5 * dee);
                    double tr2 = tr, te2 = te;
                    double sdr = 1.0 + tr / 3.0, sde = 1.0 + te, sne = 1.0 + te / 3.0;
                    for (int n = 2; n <= nkeep; ++n) {
                        tr = tr2 * tr; te = te2 * te;
                        sdr = sdr + tr / (1.0 + 2 * n);
                        sde = sde + te;
                        sne = sne + te / (1.0 + 2 * n);
                    }
                    drhof = 1.0 / sdr;
                    deef = sne / sde;
                }
                double uv_part = (uni + unip) * rhom * drhof;
                uvs2 = uvs2 + uv_part * (uui + uuip);
                uvs3 = uvs3 + uv_part * (vvi + vvip);
                uvs4 = uvs4 + uv_part * (wwi + wwip);
                uvs5_i = uvs5_i + uv_part * eem * deef;
                uvs5_k = uvs5_k + uv_part * (uui * uuip + vvi * vvip + wwi * wwip);
                uvs5_p = uvs5_p + 4.0 * (uni * ppip + unip * ppi);
                uvs6 = uvs6 + 2.0 * (ppi + ppip);
            } else {
                double uv_part = (uni + unip) * rhom;
                uvs2 = uvs2 + uv_part * (uui + uuip);
                uvs3 = uvs3 + uv_part * (vvi + vvip);
                uvs4 = uvs4 + uv_part * (wwi + wwip);
                uvs5 = uvs5 + uv_part * (enti + entip);
                uvs6 = uvs6 + 2.0 * (ppi + ppip);
            }
        }
        double cl = CD1(l, lmax);
        ft2 = ft2 + cl * uvs2;
        ft3 = ft3 + cl * uvs3;
        ft4 = ft4 + cl * uvs4;
        if (nkeep >= 0) ft5 = ft5 + cl * (uvs5_i + uvs5_k + uvs5_p);
        else ft5 = ft5 + cl * uvs5;
        ft6 = ft6 + cl * uvs6;
    }
    double fh2 = 0.25 * ft2, fh3 = 0.25 * ft3, fh4 = 0.25 * ft4, fh5 = 0.25 * ft5;
    if (zero_flux) { fh2 = 0; fh3 = 0; fh4 = 0; fh5 = 0; }
    /* pressure term on the normal momentum */
    double pres = 0.5 * ft6;
    fh[I_U - 1] = fh2; fh[I_V - 1] = fh3; fh[I_W - 1] = fh4; fh[I_E - 1] = fh5;
    fh[in_mom - 1] += pres;
}


/* compute_roe_average, kernels_gpu.F90:2986-3295: Roe state between line cells f and f+1
 * with the inline Newton temperature, c_p, species energies, generalised pressure
 * derivatives and the pressure-jump consistency correction. */
static void roe_average(const line_t *L, int f, const double *rgas, const double *cv_coeff,
                        const double *cp_coeff, int nsetcv, const double *trange, int indx_cp_l,
                        int indx_cp_r, double tol_iter_nr, double *b1, double *b2, double *b3,
                        double *rho_o, double *c_o, double *ci_o, double *h_o, double *uu_o,
                        double *vv_o, double *ww_o, double *yyroe, double *prhoi)
{
    int ns = L->ns;
    double drho_slv[NSMAX];
    double rho = LA(L, f, J_R), uu = LA(L, f, J_U), vv = LA(L, f, J_V), ww = LA(L, f, J_W);
    double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
    double h = LA(L, f, J_H), pp = LA(L, f, J_P);
    double ee = h - pp / rho - qq;
    double rho_l = rho, ee_l = ee;
    double rhop = LA(L, f + 1, J_R), up = LA(L, f + 1, J_U), vp = LA(L, f + 1, J_V), wp = LA(L, f + 1, J_W);
    double qqp = 0.5 * (up * up + vp * vp + wp * wp);
    double hp = LA(L, f + 1, J_H), ppp = LA(L, f + 1, J_P);
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
    double p_over_rho = h - ee - qq;
    double rmixtloc = 0.0;
    for (int lsp = 1; lsp <= ns; ++lsp) {
        double yy = LA(L, f, lsp), yyp = LA(L, f + 1, lsp);
        drho_slv[lsp - 1] = rhop * yyp - rho_l * yy;
        yy = (r * yyp + yy) * rp1;
        rmixtloc = rmixtloc + rgas[lsp - 1] * yy;
        yyroe[lsp - 1] = yy;
    }
    double told = p_over_rho / rmixtloc;
    double tt = temperature_from_e(ns, ee, told, cv_coeff, nsetcv, trange, indx_cp_l, indx_cp_r, tol_iter_nr, yyroe);
    double cploc = mixture_cp(ns, tt, indx_cp_l, indx_cp_r, cp_coeff, nsetcv, trange, yyroe);
    double gamloc = cploc / (cploc - rmixtloc);
    double gm1loc = gamloc - 1.0;
    double p_e = rho * gm1loc;
    for (int lsp = 1; lsp <= ns; ++lsp)
        prhoi[lsp - 1] = species_h(ns, tt, indx_cp_l, indx_cp_r, cv_coeff, nsetcv, trange, lsp);
    for (int lsp = 1; lsp <= ns; ++lsp) {
        double eei = prhoi[lsp - 1];
        prhoi[lsp - 1] = rgas[lsp - 1] * tt + gm1loc * (ee - eei);
    }
    double etot = ee + qq;
    double dp_slv = ppp - pp;
    double de_slv = eep - ee_l;
    double pede = p_e * de_slv;
    double dp_err = dp_slv - pede;
    double den_correction = pede * pede;
    for (int lsp = 1; lsp <= ns; ++lsp) {
        double prhoidrho = prhoi[lsp - 1] * drho_slv[lsp - 1];
        dp_err = dp_err - prhoidrho;
        den_correction = den_correction + prhoidrho * prhoidrho;
    }
    double dpfact = dp_err / (den_correction > 0.000001 ? den_correction : 0.000001);
    p_e = p_e * (1.0 + pede * dpfact);
    for (int lsp = 1; lsp <= ns; ++lsp)
        prhoi[lsp - 1] = prhoi[lsp - 1] * (1.0 + prhoi[lsp - 1] * drho_slv[lsp - 1] * dpfact);
    double cc = p_over_rho / rho * p_e;
    for (int lsp = 1; lsp <= ns; ++lsp) cc = cc + yyroe[lsp - 1] * prhoi[lsp - 1];
    double c = sqrt(cc);
    *c_o = c; *ci_o = 1.0 / c;
    *b3 = etot;
    *b2 = p_e / (rho * cc);
    *b1 = -(*b2) * (etot - 2.0 * qq);
    *rho_o = rho; *h_o = h; *uu_o = uu; *vv_o = vv; *ww_o = ww;
}

/* eigenvectors_x / _y / _z, kernels_gpu.F90:2703-2984.  el(row, field), er(field, row),
 * both nv x nv column-major.  dir picks the normal velocity slot. */
static void eigenvectors(int dir, int ns, double b1, double b2, double b3, double uu, double vv,
                         double ww, double c, double ci, double h, const double *yyroe,
                         const double *prhoi, double *el, double *er)
{
    int nv = ns + 4;
#define EL(r, cc) el[((r) - 1) + (size_t)nv * ((cc) - 1)]
#define ER(r, cc) er[((r) - 1) + (size_t)nv * ((cc) - 1)]
    double cci = ci * ci;
    double un = dir == 1 ? uu : dir == 2 ? vv : ww;
    /* the two transverse velocities in the order the reference uses */
    double ta = dir == 1 ? vv : uu;
    double tb = dir == 3 ? vv : ww;
    for (int q = 0; q < nv * nv; ++q) { el[q] = 0.0; er[q] = 0.0; }
    for (int lsp = 1; lsp <= ns; ++lsp) {
        double b1i = b1 + prhoi[lsp - 1] * cci;
        EL(lsp, 1) = 0.5 * (b1i + un * ci);
        EL(lsp, ns + 2) = -ta;
        EL(lsp, ns + 3) = -tb;
        EL(lsp, ns + 4) = 0.5 * (b1i - un * ci);
        EL(lsp, lsp + 1) = 1.0;
        EL(I_U, 1 + lsp) = b2 * uu * yyroe[lsp - 1];
        EL(I_V, 1 + lsp) = b2 * vv * yyroe[lsp - 1];
        EL(I_W, 1 + lsp) = b2 * ww * yyroe[lsp - 1];
        EL(I_E, 1 + lsp) = -b2 * yyroe[lsp - 1];
        for (int msp = 1; msp <= ns; ++msp) {
            double b1m = b1 + prhoi[msp - 1] * cci;
            EL(msp, 1 + lsp) = EL(msp, 1 + lsp) - b1m * yyroe[lsp - 1];
        }
    }
    EL(I_U, 1) = -0.5 * (b2 * uu + (dir == 1 ? ci : 0.0));
    EL(I_V, 1) = -0.5 * (b2 * vv + (dir == 2 ? ci : 0.0));
    EL(I_W, 1) = -0.5 * (b2 * ww + (dir == 3 ? ci : 0.0));
    EL(I_E, 1) = 0.5 * b2;
    /* transverse fields: unit entries on the transverse momenta */
    int ia = dir == 1 ? I_V : I_U;      /* first transverse momentum */
    int ib = dir == 3 ? I_V : I_W;      /* second transverse momentum */
    EL(I_U, ns + 2) = 0.0; EL(I_V, ns + 2) = 0.0; EL(I_W, ns + 2) = 0.0; EL(I_E, ns + 2) = 0.0;
    EL(I_U, ns + 3) = 0.0; EL(I_V, ns + 3) = 0.0; EL(I_W, ns + 3) = 0.0; EL(I_E, ns + 3) = 0.0;
    EL(ia, ns + 2) = 1.0;
    EL(ib, ns + 3) = 1.0;
    EL(I_U, ns + 4) = -0.5 * (b2 * uu - (dir == 1 ? ci : 0.0));
    EL(I_V, ns + 4) = -0.5 * (b2 * vv - (dir == 2 ? ci : 0.0));
    EL(I_W, ns + 4) = -0.5 * (b2 * ww - (dir == 3 ? ci : 0.0));
    EL(I_E, ns + 4) = 0.5 * b2;
    for (int lsp = 1; lsp <= ns; ++lsp) {
        double b3i = b3 - prhoi[lsp - 1] / b2 * cci;
        ER(1, lsp) = yyroe[lsp - 1];
        ER(1 + lsp, lsp) = 1.0;
        ER(4 + ns, lsp) = yyroe[lsp - 1];
        ER(1 + lsp, I_U) = uu;
        ER(1 + lsp, I_V) = vv;
        ER(1 + lsp, I_W) = ww;
        ER(1 + lsp, I_E) = b3i;
    }
    ER(1, I_U) = uu - (dir == 1 ? c : 0.0);
    ER(1, I_V) = vv - (dir == 2 ? c : 0.0);
    ER(1, I_W) = ww - (dir == 3 ? c : 0.0);
    ER(1, I_E) = h - un * c;
    ER(ns + 2, I_U) = 0.0; ER(ns + 3, I_U) = 0.0;
    ER(ns + 2, I_V) = 0.0; ER(ns + 3, I_V) = 0.0;
    ER(ns + 2, I_W) = 0.0; ER(ns + 3, I_W) = 0.0;
    ER(ns + 2, ia) = 1.0;
    ER(ns + 3, ib) = 1.0;
    ER(ns + 4, I_U) = uu + (dir == 1 ? c : 0.0);
    ER(ns + 4, I_V) = vv + (dir == 2 ? c : 0.0);
    ER(ns + 4, I_W) = ww + (dir == 3 ? c : 0.0);
    ER(ns + 2, I_E) = ta;
    ER(ns + 3, I_E) = tb;
    ER(ns + 4, I_E) = h + un * c;
#undef EL
#undef ER
}

/* wenorec_1d (kernels_gpu.F90:2353-2637) and wenorec_1d_rusanov (:2065-2351); the two
 * differ only in the smoothness-indicator scaling.  vp/vm are nvar x 8 column-major. */
static void wenorec(int ns, int nvar, const double *vp, const double *vm, double *vhat, int iweno,
                    int wenorec_ord, int weno_version, double rho0, double u0, int rusanov)
{
    double betascale[NVMAX];
    double u0_2 = u0 * u0, r2u2 = rho0 * rho0 * u0_2, r2u4 = r2u2 * u0_2;
    if (!rusanov) {
        betascale[0] = 1.0 / r2u2;
        for (int lsp = 1; lsp <= ns + 1; ++lsp) betascale[lsp - 1] = betascale[0];
        betascale[ns + 1] = 1.0 / r2u4;
        betascale[ns + 2] = betascale[ns + 1];
        betascale[ns + 3] = betascale[0];
    } else {
        double r2u4i = 1.0 / r2u4;
        betascale[0] = 1.0 / r2u2;
        for (int lsp = 2; lsp <= ns; ++lsp) betascale[lsp - 1] = betascale[0];
        betascale[I_U - 1] = r2u4i; betascale[I_V - 1] = r2u4i; betascale[I_W - 1] = r2u4i;
        betascale[I_E - 1] = r2u4i / u0_2;
    }
#define VP(m, l) vp[((m) - 1) + (size_t)nvar * ((l) - 1)]
#define VM(m, l) vm[((m) - 1) + (size_t)nvar * ((l) - 1)]
#define SQ(x) ((x) * (x))
    int i = iweno;
    double betap[6], betam[6], dwe[6];
#define BP(l) betap[(l) + 1]
#define BM(l) betam[(l) + 1]
#define DW(l) dwe[(l) + 1]
    if (wenorec_ord == 1) {
        for (int m = 1; m <= nvar; ++m) vhat[m - 1] = VP(m, i) + VM(m, i + 1);
    } else if (wenorec_ord == 2) {
        DW(1) = 2.0 / 3.0; DW(0) = 1.0 / 3.0;
        for (int m = 1; m <= nvar; ++m) {
            BP(0) = SQ(VP(m, i) - VP(m, i - 1));
            BP(1) = SQ(VP(m, i + 1) - VP(m, i));
            BP(1) = betascale[m - 1] * BP(1);
            BP(0) = betascale[m - 1] * BP(0);
            BM(0) = SQ(VM(m, i + 2) - VM(m, i + 1));
            BM(1) = SQ(VM(m, i + 1) - VM(m, i));
            BM(1) = betascale[m - 1] * BM(1);
            BM(0) = betascale[m - 1] * BM(0);
            double sump = 0.0, summ = 0.0;
            for (int l = 0; l <= 1; ++l) {
                BP(l) = DW(l) / SQ(0.000001 + BP(l));
                BM(l) = DW(l) / SQ(0.000001 + BM(l));
                sump = sump + BP(l); summ = summ + BM(l);
            }
            for (int l = 0; l <= 1; ++l) { BP(l) = BP(l) / sump; BM(l) = BM(l) / summ; }
            double vminus = BP(0) * (-VP(m, i - 1) + 3 * VP(m, i)) + BP(1) * (VP(m, i) + VP(m, i + 1));
            double vplus = BM(0) * (-VM(m, i + 2) + 3 * VM(m, i + 1)) + BM(1) * (VM(m, i) + VM(m, i + 1));
            vhat[m - 1] = 0.5 * (vminus + vplus);
        }
    } else if (wenorec_ord == 3) {
        DW(0) = 1.0 / 10.0; DW(1) = 6.0 / 10.0; DW(2) = 3.0 / 10.0;
        double d0 = 13.0 / 12.0, d1 = 1.0 / 4.0;
        double c0 = 1.0 / 3.0, c1 = 5.0 / 6.0, c2 = -1.0 / 6.0, c3 = -7.0 / 6.0, c4 = 11.0 / 6.0;
        for (int m = 1; m <= nvar; ++m) {
            BP(2) = d0 * SQ(VP(m, i) - 2.0 * VP(m, i + 1) + VP(m, i + 2)) + d1 * SQ(3.0 * VP(m, i) - 4.0 * VP(m, i + 1) + VP(m, i + 2));
            BP(1) = d0 * SQ(VP(m, i - 1) - 2.0 * VP(m, i) + VP(m, i + 1)) + d1 * SQ(VP(m, i - 1) - VP(m, i + 1));
            BP(0) = d0 * SQ(VP(m, i) - 2.0 * VP(m, i - 1) + VP(m, i - 2)) + d1 * SQ(3.0 * VP(m, i) - 4.0 * VP(m, i - 1) + VP(m, i - 2));
            BP(2) = betascale[m - 1] * BP(2); BP(1) = betascale[m - 1] * BP(1); BP(0) = betascale[m - 1] * BP(0);
            BM(2) = d0 * SQ(VM(m, i + 1) - 2.0 * VM(m, i) + VM(m, i - 1)) + d1 * SQ(3.0 * VM(m, i + 1) - 4.0 * VM(m, i) + VM(m, i - 1));
            BM(1) = d0 * SQ(VM(m, i + 2) - 2.0 * VM(m, i + 1) + VM(m, i)) + d1 * SQ(VM(m, i + 2) - VM(m, i));
            BM(0) = d0 * SQ(VM(m, i + 1) - 2.0 * VM(m, i + 2) + VM(m, i + 3)) + d1 * SQ(3.0 * VM(m, i + 1) - 4.0 * VM(m, i + 2) + VM(m, i + 3));
            BM(2) = betascale[m - 1] * BM(2); BM(1) = betascale[m - 1] * BM(1); BM(0) = betascale[m - 1] * BM(0);
            double sump = 0.0, summ = 0.0;
            if (weno_version == 0) {
                for (int l = 0; l <= 2; ++l) {
                    BP(l) = DW(l) / SQ(0.000001 + BP(l));
                    BM(l) = DW(l) / SQ(0.000001 + BM(l));
                    sump = sump + BP(l); summ = summ + BM(l);
                }
            } else {
                double eps40 = 1.0e-40;
                double tau5p = fabs(BP(0) - BP(2)) + eps40;
                double tau5m = fabs(BM(0) - BM(2)) + eps40;
                for (int l = 0; l <= 2; ++l) {
                    BP(l) = (BP(l) + eps40) / (BP(l) + tau5p);
                    BM(l) = (BM(l) + eps40) / (BM(l) + tau5m);
                }
                for (int l = 0; l <= 2; ++l) {
                    BP(l) = DW(l) / BP(l);
                    BM(l) = DW(l) / BM(l);
                    sump = sump + BP(l); summ = summ + BM(l);
                }
            }
            for (int l = 0; l <= 2; ++l) { BP(l) = BP(l) / sump; BM(l) = BM(l) / summ; }
            double vminus = BP(2) * (c0 * VP(m, i) + c1 * VP(m, i + 1) + c2 * VP(m, i + 2)) +
                            BP(1) * (c2 * VP(m, i - 1) + c1 * VP(m, i) + c0 * VP(m, i + 1)) +
                            BP(0) * (c0 * VP(m, i - 2) + c3 * VP(m, i - 1) + c4 * VP(m, i));
            double vplus = BM(2) * (c0 * VM(m, i + 1) + c1 * VM(m, i) + c2 * VM(m, i - 1)) +
                           BM(1) * (c2 * VM(m, i + 2) + c1 * VM(m, i + 1) + c0 * VM(m, i)) +
                           BM(0) * (c0 * VM(m, i + 3) + c3 * VM(m, i + 2) + c4 * VM(m, i + 1));
            vhat[m - 1] = vminus + vplus;
        }
    } else if (wenorec_ord == 4) {
        DW(0) = 1.0 / 35.0; DW(1) = 12.0 / 35.0; DW(2) = 18.0 / 35.0; DW(3) = 4.0 / 35.0;
        double d1 = 1.0 / 36.0, d2 = 13.0 / 12.0, d3 = 781.0 / 720.0;
        for (int m = 1; m <= nvar; ++m) {
            BP(3) = d1 * SQ(-11 * VP(m, i) + 18 * VP(m, i + 1) - 9 * VP(m, i + 2) + 2 * VP(m, i + 3)) +
                    d2 * SQ(2 * VP(m, i) - 5 * VP(m, i + 1) + 4 * VP(m, i + 2) - VP(m, i + 3)) +
                    d3 * SQ(-VP(m, i) + 3 * VP(m, i + 1) - 3 * VP(m, i + 2) + VP(m, i + 3));
            BP(2) = d1 * SQ(-2 * VP(m, i - 1) - 3 * VP(m, i) + 6 * VP(m, i + 1) - VP(m, i + 2)) +
                    d2 * SQ(VP(m, i - 1) - 2 * VP(m, i) + VP(m, i + 1)) +
                    d3 * SQ(-VP(m, i - 1) + 3 * VP(m, i) - 3 * VP(m, i + 1) + VP(m, i + 2));
            BP(1) = d1 * SQ(VP(m, i - 2) - 6 * VP(m, i - 1) + 3 * VP(m, i) + 2 * VP(m, i + 1)) +
                    d2 * SQ(VP(m, i - 1) - 2 * VP(m, i) + VP(m, i + 1)) +
                    d3 * SQ(-VP(m, i - 2) + 3 * VP(m, i - 1) - 3 * VP(m, i) + VP(m, i + 1));
            BP(0) = d1 * SQ(-2 * VP(m, i - 3) + 9 * VP(m, i - 2) - 18 * VP(m, i - 1) + 11 * VP(m, i)) +
                    d2 * SQ(-VP(m, i - 3) + 4 * VP(m, i - 2) - 5 * VP(m, i - 1) + 2 * VP(m, i)) +
                    d3 * SQ(-VP(m, i - 3) + 3 * VP(m, i - 2) - 3 * VP(m, i - 1) + VP(m, i));
            for (int l = 0; l <= 3; ++l) BP(l) = betascale[m - 1] * BP(l);
            BM(3) = d1 * SQ(-11 * VM(m, i + 1) + 18 * VM(m, i) - 9 * VM(m, i - 1) + 2 * VM(m, i - 2)) +
                    d2 * SQ(2 * VM(m, i + 1) - 5 * VM(m, i) + 4 * VM(m, i - 1) - VM(m, i - 2)) +
                    d3 * SQ(-VM(m, i + 1) + 3 * VM(m, i) - 3 * VM(m, i - 1) + VM(m, i - 2));
            BM(2) = d1 * SQ(-2 * VM(m, i + 2) - 3 * VM(m, i + 1) + 6 * VM(m, i) - VM(m, i - 1)) +
                    d2 * SQ(VM(m, i + 2) - 2 * VM(m, i + 1) + VM(m, i)) +
                    d3 * SQ(-VM(m, i + 2) + 3 * VM(m, i + 1) - 3 * VM(m, i) + VM(m, i - 1));
            BM(1) = d1 * SQ(VM(m, i + 3) - 6 * VM(m, i + 2) + 3 * VM(m, i + 1) + 2 * VM(m, i)) +
                    d2 * SQ(VM(m, i + 2) - 2 * VM(m, i + 1) + VM(m, i)) +
                    d3 * SQ(-VM(m, i + 3) + 3 * VM(m, i + 2) - 3 * VM(m, i + 1) + VM(m, i));
            BM(0) = d1 * SQ(-2 * VM(m, i + 4) + 9 * VM(m, i + 3) - 18 * VM(m, i + 2) + 11 * VM(m, i + 1)) +
                    d2 * SQ(-VM(m, i + 4) + 4 * VM(m, i + 3) - 5 * VM(m, i + 2) + 2 * VM(m, i + 1)) +
                    d3 * SQ(-VM(m, i + 4) + 3 * VM(m, i + 3) - 3 * VM(m, i + 2) + VM(m, i + 1));
            for (int l = 0; l <= 3; ++l) BM(l) = betascale[m - 1] * BM(l);
            double sump = 0.0, summ = 0.0;
            for (int l = 0; l <= 3; ++l) {
                BP(l) = DW(l) / SQ(0.000001 + BP(l));
                BM(l) = DW(l) / SQ(0.000001 + BM(l));
                sump = sump + BP(l); summ = summ + BM(l);
            }
            for (int l = 0; l <= 3; ++l) { BP(l) = BP(l) / sump; BM(l) = BM(l) / summ; }
            double vminus = BP(3) * (6 * VP(m, i) + 26 * VP(m, i + 1) - 10 * VP(m, i + 2) + 2 * VP(m, i + 3)) +
                            BP(2) * (-2 * VP(m, i - 1) + 14 * VP(m, i) + 14 * VP(m, i + 1) - 2 * VP(m, i + 2)) +
                            BP(1) * (2 * VP(m, i - 2) - 10 * VP(m, i - 1) + 26 * VP(m, i) + 6 * VP(m, i + 1)) +
                            BP(0) * (-6 * VP(m, i - 3) + 26 * VP(m, i - 2) - 46 * VP(m, i - 1) + 50 * VP(m, i));
            double vplus = BM(3) * (6 * VM(m, i + 1) + 26 * VM(m, i) - 10 * VM(m, i - 1) + 2 * VM(m, i - 2)) +
                           BM(2) * (-2 * VM(m, i + 2) + 14 * VM(m, i + 1) + 14 * VM(m, i) - 2 * VM(m, i - 1)) +
                           BM(1) * (2 * VM(m, i + 3) - 10 * VM(m, i + 2) + 26 * VM(m, i + 1) + 6 * VM(m, i)) +
                           BM(0) * (-6 * VM(m, i + 4) + 26 * VM(m, i + 3) - 46 * VM(m, i + 2) + 50 * VM(m, i + 1));
            vhat[m - 1] = (vminus + vplus) / 24.0;
        }
    }
#undef VP
#undef VM
#undef SQ
#undef BP
#undef BM
#undef DW
}

/* The WENO branch of euler_{x,y,z}_hybrid[_rusanov]_kernel for face f of the line
 * (kernels_gpu.F90:338-421 Roe/LLF; :683-727 Rusanov). */
int orc_weno_face(const line_t *L, int f, int weno_scheme, int weno_version, int wenorec_ord,
                  const double *rgas, const double *cv_coeff, const double *cp_coeff,
                  int nsetcv, const double *trange, int indx_cp_l, int indx_cp_r,
                  double tol_iter_nr, double rho0, double u0, int flux_splitting, double *fh)
{
    int ns = L->ns, nv = ns + 4, dir = L->dir;
    int weno_size = 2 * weno_scheme;
    if (weno_size > 8) return -3;
    int jn = dir == 1 ? J_U : dir == 2 ? J_V : J_W;
    int in_mom = dir == 1 ? I_U : dir == 2 ? I_V : I_W;
    double gp[NVMAX * 8], gm[NVMAX * 8], fi[NVMAX];
    if (flux_splitting == 1) {
        double evmax = -1.0;
        for (int l = 1; l <= weno_size; ++l) {
            int ll = f + l - weno_scheme;
            double un = LA(L, ll, jn), c = LA(L, ll, J_C);
            double a = fabs(un - c), b = fabs(un + c);
            double evm = a > b ? a : b;
            if (evm > evmax) evmax = evm;
        }
        for (int l = 1; l <= weno_size; ++l) {
            int ll = f + l - weno_scheme;
            double rho = LA(L, ll, J_R), uu = LA(L, ll, J_U), vv = LA(L, ll, J_V), ww = LA(L, ll, J_W);
            double h = LA(L, ll, J_H), pp = LA(L, ll, J_P);
            double un = LA(L, ll, jn);
            double rhou = rho * un, rhoevm = rho * evmax;
            for (int lsp = 1; lsp <= ns; ++lsp) {
                double yy = LA(L, ll, lsp);
                double evm = rhou * yy;
                double c = 0.5 * (evm + rhoevm * yy);
                gp[(lsp - 1) + nv * (l - 1)] = c;
                gm[(lsp - 1) + nv * (l - 1)] = evm - c;
            }
            double vel[3] = {uu, vv, ww};
            for (int q = 0; q < 3; ++q) {
                double evm = vel[q] * rhou + ((I_U + q) == in_mom ? pp : 0.0);
                double c = 0.5 * (evm + rhoevm * vel[q]);
                gp[(I_U + q - 1) + nv * (l - 1)] = c;
                gm[(I_U + q - 1) + nv * (l - 1)] = evm - c;
            }
            double evm = h * rhou;
            double c = 0.5 * (evm + evmax * (rho * h - pp));
            gp[(I_E - 1) + nv * (l - 1)] = c;
            gm[(I_E - 1) + nv * (l - 1)] = evm - c;
        }
        wenorec(ns, nv, gp, gm, fi, weno_scheme, wenorec_ord, weno_version, rho0, u0, 1);
        for (int m = 0; m < nv; ++m) fh[m] = fi[m];
        return 0;
    }
    double b1, b2, b3, rho, c, ci, h, uu, vv, ww, yyroe[NSMAX], prhoi[NSMAX];
    double el[NVMAX * NVMAX], er[NVMAX * NVMAX], evmax[NVMAX];
    roe_average(L, f, rgas, cv_coeff, cp_coeff, nsetcv, trange, indx_cp_l, indx_cp_r, tol_iter_nr,
                &b1, &b2, &b3, &rho, &c, &ci, &h, &uu, &vv, &ww, yyroe, prhoi);
    eigenvectors(dir, ns, b1, b2, b3, uu, vv, ww, c, ci, h, yyroe, prhoi, el, er);
    for (int m = 0; m < nv; ++m) evmax[m] = -1.0;
    for (int l = 1; l <= weno_size; ++l) {
        int ll = f + l - weno_scheme;
        double un = LA(L, ll, jn), cc = LA(L, ll, J_C);
        double a;
        a = fabs(un - cc); if (a > evmax[0]) evmax[0] = a;
        a = fabs(un); if (a > evmax[1]) evmax[1] = a;
        for (int lsp = 3; lsp <= ns + 3; ++lsp) evmax[lsp - 1] = evmax[1];
        a = fabs(un + cc); if (a > evmax[ns + 3]) evmax[ns + 3] = a;
    }
    for (int l = 1; l <= weno_size; ++l) {
        int ll = f + l - weno_scheme;
        double rl = LA(L, ll, J_R), ul = LA(L, ll, J_U), vl = LA(L, ll, J_V), wl = LA(L, ll, J_W);
        double hl = LA(L, ll, J_H), pl = LA(L, ll, J_P);
        double unl = LA(L, ll, jn);
        double rhou = rl * unl;
        for (int lsp = 1; lsp <= ns; ++lsp) fi[lsp - 1] = rhou * LA(L, ll, lsp);
        fi[I_U - 1] = ul * rhou; fi[I_V - 1] = vl * rhou; fi[I_W - 1] = wl * rhou; fi[I_E - 1] = hl * rhou;
        fi[in_mom - 1] += pl;
        for (int m = 1; m <= nv; ++m) {
            double wc = 0.0, gc = 0.0;
            for (int lsp = 1; lsp <= ns; ++lsp) {
                wc = wc + el[(lsp - 1) + nv * (m - 1)] * rl * LA(L, ll, lsp);
                gc = gc + el[(lsp - 1) + nv * (m - 1)] * fi[lsp - 1];
            }
            wc = wc + el[(I_U - 1) + nv * (m - 1)] * rl * ul;
            gc = gc + el[(I_U - 1) + nv * (m - 1)] * fi[I_U - 1];
            wc = wc + el[(I_V - 1) + nv * (m - 1)] * rl * vl;
            gc = gc + el[(I_V - 1) + nv * (m - 1)] * fi[I_V - 1];
            wc = wc + el[(I_W - 1) + nv * (m - 1)] * rl * wl;
            gc = gc + el[(I_W - 1) + nv * (m - 1)] * fi[I_W - 1];
            wc = wc + el[(I_E - 1) + nv * (m - 1)] * (rl * hl - pl);
            gc = gc + el[(I_E - 1) + nv * (m - 1)] * fi[I_E - 1];
            double cc = 0.5 * (gc + evmax[m - 1] * wc);
            gp[(m - 1) + nv * (l - 1)] = cc;
            gm[(m - 1) + nv * (l - 1)] = gc - cc;
        }
    }
    wenorec(ns, nv, gp, gm, fi, weno_scheme, wenorec_ord, weno_version, rho0, u0, 0);
    for (int m = 1; m <= nv; ++m) {
        double s = 0.0;
        for (int mm = 1; mm <= nv; ++mm) s = s + er[(mm - 1) + nv * (m - 1)] * fi[mm - 1];
        fh[m - 1] = s;
    }
    return 0;
}

static int euler_dir(int dir, int ns, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                     int fstart, int fend, int lmax_base, int nkeep, const double *rgas,
                     const double *w_aux, const double *coeff_deriv1, const double *metric,
                     int zmin, int zmax, int weno_scheme, int weno_version,
                     double sensor_threshold, const double *cv_coeff, const double *cp_coeff,
                     int nsetcv, const double *trange, int indx_cp_l, int indx_cp_r,
                     const int *ep_ord_change, double tol_iter_nr, double rho0, double u0,
                     int flux_splitting, int umin, int umax, double *fl)
{
    (void)nv_aux;
    int n = dir == 1 ? nx : dir == 2 ? ny : nz;
    int t1n = dir == 1 ? ny : nx;           /* transverse extents */
    int t2n = dir == 3 ? ny : nz;
    int status = 0;
    OMP_FOR
    for (int t2 = 1; t2 <= t2n; ++t2) {
    double fh[NVMAX], *fhl = malloc(sizeof(double) * (size_t)(n + 2 * ng + 2) * nv);
    if (!fhl) { status = -4; continue; }
    for (int t1 = 1; t1 <= t1n; ++t1) {
        line_t L = {w_aux, nx, ny, nz, ng, ns, dir, 0, 0, 0};
        if (dir == 1) { L.j = t1; L.k = t2; }
        else if (dir == 2) { L.i = t1; L.k = t2; }
        else { L.i = t1; L.j = t2; }
        for (int f = fstart; f <= fend; ++f) {
            int ishk = 0;
            for (int ii = f - weno_scheme + 1; ii <= f + weno_scheme; ++ii)
                if (LA(&L, ii, J_DUC) > sensor_threshold) ishk = 1;
            int ei = dir == 1 ? f : L.i, ej = dir == 2 ? f : L.j, ek = dir == 3 ? f : L.k;
            int epo = EPO(ei, ej, ek, dir);
            if (ishk == 0) {
                int lmax = lmax_base + epo; if (lmax < 1) lmax = 1;
                int zero = (f == 0 && zmin == 1) || (f == n && zmax == 1);
                keep_flux(&L, f, lmax, nkeep, coeff_deriv1, zero, fh);
            } else {
                int ord = weno_scheme + epo; if (ord < 1) ord = 1;
                int rc = orc_weno_face(&L, f, weno_scheme, weno_version, ord, rgas, cv_coeff, cp_coeff,
                                       nsetcv, trange, indx_cp_l, indx_cp_r, tol_iter_nr, rho0, u0,
                                       flux_splitting, fh);
                if (rc) status = rc;
            }
            for (int m = 0; m < nv; ++m) fhl[(size_t)(f + ng) * nv + m] = fh[m];
        }
        for (int c = umin; c <= umax; ++c) {
            int ci = dir == 1 ? c : L.i, cj = dir == 2 ? c : L.j, ck = dir == 3 ? c : L.k;
            for (int m = 1; m <= nv; ++m)
                F(fl, ci, cj, ck, m) = F(fl, ci, cj, ck, m) +
                    (fhl[(size_t)(c + ng) * nv + m - 1] - fhl[(size_t)(c - 1 + ng) * nv + m - 1]) * metric[c - 1];
        }
    }
    free(fhl);
    }
    return status;
}

int orc_euler_x(int ns, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                int istart_face, int iend_face, int lmax_base, int nkeep,
                const double *rgas, const double *w_aux, const double *coeff_deriv1,
                const double *dcsidx, double *fhat, int force_zero_flux_min,
                int force_zero_flux_max, int weno_scheme, int weno_version,
                double sensor_threshold, int weno_size, const double *cv_coeff,
                const double *cp_coeff, int nsetcv, const double *trange,
                int indx_cp_l, int indx_cp_r, const int *ep_ord_change,
                double tol_iter_nr, double rho0, double u0, int flux_splitting,
                int eul_imin, int eul_imax, double *fl, void *stream)
{
    (void)fhat; (void)weno_size; (void)stream;
    return euler_dir(1, ns, nv, nv_aux, nx, ny, nz, ng, istart_face, iend_face, lmax_base, nkeep, rgas,
                     w_aux, coeff_deriv1, dcsidx, force_zero_flux_min, force_zero_flux_max, weno_scheme,
                     weno_version, sensor_threshold, cv_coeff, cp_coeff, nsetcv, trange, indx_cp_l,
                     indx_cp_r, ep_ord_change, tol_iter_nr, rho0, u0, flux_splitting, eul_imin, eul_imax, fl);
}

int orc_euler_y(int ns, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                int eul_jmin, int eul_jmax, int lmax_base, int nkeep,
                const double *rgas, const double *w_aux, double *fl,
                const double *coeff_deriv1, const double *detady, double *fhat,
                int force_zero_flux_min, int force_zero_flux_max, int weno_scheme,
                int weno_version, double sensor_threshold, int weno_size,
                const double *cv_coeff, const double *cp_coeff, int nsetcv,
                const double *trange, int indx_cp_l, int indx_cp_r,
                const int *ep_ord_change, double tol_iter_nr, double rho0, double u0,
                int flux_splitting, void *stream)
{
    (void)fhat; (void)weno_size; (void)stream;
    return euler_dir(2, ns, nv, nv_aux, nx, ny, nz, ng, eul_jmin - 1, eul_jmax, lmax_base, nkeep, rgas,
                     w_aux, coeff_deriv1, detady, force_zero_flux_min, force_zero_flux_max, weno_scheme,
                     weno_version, sensor_threshold, cv_coeff, cp_coeff, nsetcv, trange, indx_cp_l,
                     indx_cp_r, ep_ord_change, tol_iter_nr, rho0, u0, flux_splitting, eul_jmin, eul_jmax, fl);
}

int orc_euler_z(int ns, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                int eul_kmin, int eul_kmax, int lmax_base, int nkeep,
                const double *rgas, const double *w_aux, double *fl,
                const double *coeff_deriv1, const double *dzitdz, double *fhat,
                int force_zero_flux_min, int force_zero_flux_max, int weno_scheme,
                int weno_version, double sensor_threshold, int weno_size,
                const double *cv_coeff, const double *cp_coeff, int nsetcv,
                const double *trange, int indx_cp_l, int indx_cp_r,
                const int *ep_ord_change, double tol_iter_nr, double rho0, double u0,
                int flux_splitting, void *stream)
{
    (void)fhat; (void)weno_size; (void)stream;
    return euler_dir(3, ns, nv, nv_aux, nx, ny, nz, ng, eul_kmin - 1, eul_kmax, lmax_base, nkeep, rgas,
                     w_aux, coeff_deriv1, dzitdz, force_zero_flux_min, force_zero_flux_max, weno_scheme,
                     weno_version, sensor_threshold, cv_coeff, cp_coeff, nsetcv, trange, indx_cp_l,
                     indx_cp_r, ep_ord_change, tol_iter_nr, rho0, u0, flux_splitting, eul_kmin, eul_kmax, fl);
}

/* ===================================================================================== */
/* a6: viscous / diffusive fluxes                                                         */
/* ===================================================================================== */

/* visflx_x_cuf / _y / _z, kernels_gpu.F90:3658-4031 */
static int visflx_dir(int dir, int ns, int nx, int ny, int nz, int nv, int nv_aux, int ng,
                      const double *coord, const double *w_aux, double *fl,
                      int indx_cp_l, int indx_cp_r, const double *cp_coeff, int nsetcv,
                      const double *trange, const double *rgas, double r_univ)
{
    (void)nv_aux;
    int n = dir == 1 ? nx : dir == 2 ? ny : nz;
    int t1n = dir == 1 ? ny : nx;
    int t2n = dir == 3 ? ny : nz;
    int status = 0;
    OMP_FOR
    for (int t2 = 1; t2 <= t2n; ++t2) {
    double *fhl = malloc(sizeof(double) * (size_t)(n + 1) * nv);
    if (!fhl) { status = -4; continue; }
    for (int t1 = 1; t1 <= t1n; ++t1) {
        int bi, bj, bk;
        if (dir == 1) { bi = 0; bj = t1; bk = t2; }
        else if (dir == 2) { bi = t1; bj = 0; bk = t2; }
        else { bi = t1; bj = t2; bk = 0; }
        for (int f = 0; f <= n; ++f) {
            int i0 = bi, j0 = bj, k0 = bk, i1 = bi, j1 = bj, k1 = bk;
            if (dir == 1) { i0 = f; i1 = f + 1; }
            else if (dir == 2) { j0 = f; j1 = f + 1; }
            else { k0 = f; k1 = f + 1; }
            double uu = W(w_aux, i0, j0, k0, J_U), uup = W(w_aux, i1, j1, k1, J_U);
            double vv = W(w_aux, i0, j0, k0, J_V), vvp = W(w_aux, i1, j1, k1, J_V);
            double ww = W(w_aux, i0, j0, k0, J_W), wwp = W(w_aux, i1, j1, k1, J_W);
            double tt = W(w_aux, i0, j0, k0, J_T), ttp = W(w_aux, i1, j1, k1, J_T);
            double mu = W(w_aux, i0, j0, k0, J_MU), mup = W(w_aux, i1, j1, k1, J_MU);
            double kc = W(w_aux, i0, j0, k0, J_K_COND), kcp = W(w_aux, i1, j1, k1, J_K_COND);
            double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
            double qqp = 0.5 * (uup * uup + vvp * vvp + wwp * wwp);
            double dx = coord[f + 1 + ng - 1] - coord[f + ng - 1];
            double sigx = uup - uu, sigy = vvp - vv, sigz = wwp - ww;
            double sigq_tt = ttp - tt, sigq_qq = qqp - qq;
            double muf = mu + mup;
            muf = 0.5 * muf / dx;
            double kcf = kc + kcp;
            kcf = 0.5 * kcf / dx;
            sigx = sigx * muf; sigy = sigy * muf; sigz = sigz * muf;
            double sigq = sigq_tt * kcf + sigq_qq * muf;
            double rm = rgas[0] * W(w_aux, i0, j0, k0, 1);
            for (int lsp = 2; lsp <= ns; ++lsp) rm = rm + rgas[lsp - 1] * W(w_aux, i0, j0, k0, lsp);
            double mwm = r_univ / rm;
            rm = rgas[0] * W(w_aux, i1, j1, k1, 1);
            for (int lsp = 2; lsp <= ns; ++lsp) rm = rm + rgas[lsp - 1] * W(w_aux, i1, j1, k1, lsp);
            double mwmp = r_univ / rm;
            double mwmif = 0.5 * (mwmp + mwm) / (mwmp * mwm);
            mwmif = mwmif / dx;
            double fhat_tmp = 0.0;
            for (int msp = 1; msp <= ns; ++msp) {
                double rhodiff = 0.5 * (W(w_aux, i0, j0, k0, J_D_START + msp) + W(w_aux, i1, j1, k1, J_D_START + msp));
                double xxp = W(w_aux, i1, j1, k1, msp) * mwmp;
                double xx = W(w_aux, i0, j0, k0, msp) * mwm;
                fhat_tmp = fhat_tmp + (xxp - xx) * rhodiff;
            }
            fhat_tmp = fhat_tmp * mwmif;
            double fhat_tmp_et = 0.0;
            double *fh = fhl + (size_t)f * nv;
            for (int lsp = 1; lsp <= ns; ++lsp) {
                double rhodiff = 0.5 * (W(w_aux, i0, j0, k0, J_D_START + lsp) + W(w_aux, i1, j1, k1, J_D_START + lsp));
                double yyp = W(w_aux, i1, j1, k1, lsp), yy = W(w_aux, i0, j0, k0, lsp);
                double yyf = 0.5 * (yyp + yy);
                double xxp = yyp * mwmp, xx = yy * mwm;
                double siglsp = -rhodiff * (xxp - xx) * mwmif + fhat_tmp * yyf;
                fh[lsp - 1] = siglsp;
                double hh = species_h(ns, tt, indx_cp_l, indx_cp_r, cp_coeff, nsetcv, trange, lsp);
                double hhp = species_h(ns, ttp, indx_cp_l, indx_cp_r, cp_coeff, nsetcv, trange, lsp);
                double hhf = 0.5 * (hh + hhp);
                fhat_tmp_et = fhat_tmp_et + siglsp * hhf;
            }
            fh[I_U - 1] = -sigx;
            fh[I_V - 1] = -sigy;
            fh[I_W - 1] = -sigz;
            fh[I_E - 1] = -sigq + fhat_tmp_et;
        }
        for (int c = 1; c <= n; ++c) {
            int ci = dir == 1 ? c : bi, cj = dir == 2 ? c : bj, ck = dir == 3 ? c : bk;
            double dxhl = 2.0 / (coord[c + 1 + ng - 1] - coord[c - 1 + ng - 1]);
            for (int m = 1; m <= nv; ++m)
                F(fl, ci, cj, ck, m) = F(fl, ci, cj, ck, m) + dxhl * (fhl[(size_t)c * nv + m - 1] - fhl[(size_t)(c - 1) * nv + m - 1]);
        }
    }
    free(fhl);
    }
    return status;
}

int orc_visflx_x(int ns, int nx, int ny, int nz, int nv, int nv_aux, int ng, const double *x,
                 const double *w_aux, double *fl, double *fhat, int indx_cp_l, int indx_cp_r,
                 const double *cp_coeff, int nsetcv, const double *trange, const double *rgas,
                 double r_univ, void *stream)
{
    (void)fhat; (void)stream;
    return visflx_dir(1, ns, nx, ny, nz, nv, nv_aux, ng, x, w_aux, fl, indx_cp_l, indx_cp_r, cp_coeff, nsetcv, trange, rgas, r_univ);
}
int orc_visflx_y(int ns, int nx, int ny, int nz, int nv, int nv_aux, int ng, const double *y,
                 const double *w_aux, double *fl, double *fhat, int indx_cp_l, int indx_cp_r,
                 const double *cp_coeff, int nsetcv, const double *trange, const double *rgas,
                 double r_univ, void *stream)
{
    (void)fhat; (void)stream;
    return visflx_dir(2, ns, nx, ny, nz, nv, nv_aux, ng, y, w_aux, fl, indx_cp_l, indx_cp_r, cp_coeff, nsetcv, trange, rgas, r_univ);
}
int orc_visflx_z(int ns, int nx, int ny, int nz, int nv, int nv_aux, int ng, const double *z,
                 const double *w_aux, double *fl, double *fhat, int indx_cp_l, int indx_cp_r,
                 const double *cp_coeff, int nsetcv, const double *trange, const double *rgas,
                 double r_univ, void *stream)
{
    (void)fhat; (void)stream;
    return visflx_dir(3, ns, nx, ny, nz, nv, nv_aux, ng, z, w_aux, fl, indx_cp_l, indx_cp_r, cp_coeff, nsetcv, trange, rgas, r_univ);
}

/* visflx_reduced_cuf, kernels_gpu.F90:3539-3656 */
int orc_visflx_reduced(int ns, int nx, int ny, int nz, int ng, int nv_aux, int visc_order,
                       const double *w_aux, double *fl, const double *coeff_deriv1,
                       const double *dcsidx, const double *detady, const double *dzitdz,
                       void *stream)
{
    (void)nv_aux; (void)stream;
    int lmax = visc_order / 2;
    OMP_FOR
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        double uu = W(w_aux, i, j, k, J_U), vv = W(w_aux, i, j, k, J_V), ww = W(w_aux, i, j, k, J_W);
        double mu = W(w_aux, i, j, k, J_MU);
        double div3l = W(w_aux, i, j, k, J_DIV);
        double ux = 0, vx = 0, wx = 0, mux = 0, divx3l = 0;
        double uy = 0, vy = 0, wy = 0, muy = 0, divy3l = 0;
        double uz = 0, vz = 0, wz = 0, muz = 0, divz3l = 0;
        for (int l = 1; l <= lmax; ++l) {
            double ccl = CD1(l, lmax);
            ux = ux + ccl * (W(w_aux, i + l, j, k, J_U) - W(w_aux, i - l, j, k, J_U));
            vx = vx + ccl * (W(w_aux, i + l, j, k, J_V) - W(w_aux, i - l, j, k, J_V));
            wx = wx + ccl * (W(w_aux, i + l, j, k, J_W) - W(w_aux, i - l, j, k, J_W));
            mux = mux + ccl * (W(w_aux, i + l, j, k, J_MU) - W(w_aux, i - l, j, k, J_MU));
            divx3l = divx3l + ccl * (W(w_aux, i + l, j, k, J_DIV) - W(w_aux, i - l, j, k, J_DIV));
            uy = uy + ccl * (W(w_aux, i, j + l, k, J_U) - W(w_aux, i, j - l, k, J_U));
            vy = vy + ccl * (W(w_aux, i, j + l, k, J_V) - W(w_aux, i, j - l, k, J_V));
            wy = wy + ccl * (W(w_aux, i, j + l, k, J_W) - W(w_aux, i, j - l, k, J_W));
            muy = muy + ccl * (W(w_aux, i, j + l, k, J_MU) - W(w_aux, i, j - l, k, J_MU));
            divy3l = divy3l + ccl * (W(w_aux, i, j + l, k, J_DIV) - W(w_aux, i, j - l, k, J_DIV));
            uz = uz + ccl * (W(w_aux, i, j, k + l, J_U) - W(w_aux, i, j, k - l, J_U));
            vz = vz + ccl * (W(w_aux, i, j, k + l, J_V) - W(w_aux, i, j, k - l, J_V));
            wz = wz + ccl * (W(w_aux, i, j, k + l, J_W) - W(w_aux, i, j, k - l, J_W));
            muz = muz + ccl * (W(w_aux, i, j, k + l, J_MU) - W(w_aux, i, j, k - l, J_MU));
            divz3l = divz3l + ccl * (W(w_aux, i, j, k + l, J_DIV) - W(w_aux, i, j, k - l, J_DIV));
        }
        double dx = dcsidx[i - 1], dy = detady[j - 1], dz = dzitdz[k - 1];
        ux *= dx; vx *= dx; wx *= dx; mux *= dx; divx3l *= dx;
        uy *= dy; vy *= dy; wy *= dy; muy *= dy; divy3l *= dy;
        uz *= dz; vz *= dz; wz *= dz; muz *= dz; divz3l *= dz;
        double sig11 = ux - 2.0 * div3l, sig12 = vx, sig13 = wx;
        double sig21 = uy, sig22 = vy - 2.0 * div3l, sig23 = wy;
        double sig31 = uz, sig32 = vz, sig33 = wz - 2.0 * div3l;
        double sigx = mux * sig11 + muy * sig12 + muz * sig13 + mu * divx3l;
        double sigy = mux * sig21 + muy * sig22 + muz * sig23 + mu * divy3l;
        double sigz = mux * sig31 + muy * sig32 + muz * sig33 + mu * divz3l;
        double sigah = (sig11 * ux + sig12 * uy + sig13 * uz + sig21 * vx + sig22 * vy + sig23 * vz +
                        sig31 * wx + sig32 * wy + sig33 * wz) * mu;
        double sigq = sigx * uu + sigy * vv + sigz * ww + sigah;
        F(fl, i, j, k, I_U) = F(fl, i, j, k, I_U) - sigx;
        F(fl, i, j, k, I_V) = F(fl, i, j, k, I_V) - sigy;
        F(fl, i, j, k, I_W) = F(fl, i, j, k, I_W) - sigz;
        F(fl, i, j, k, I_E) = F(fl, i, j, k, I_E) - sigq;
    }
    return 0;
}

/* visflx_cuf (kernels_gpu.F90:3357-3537), conservative_viscous = 0: the Laplacian form of
 * the viscous terms, momentum and energy only */
#define CD2(l, lmax) coeff_deriv2[(size_t)(l) + 5 * (size_t)((lmax) - 1)]
int orc_visflx(int ns, int nx, int ny, int nz, int ng, int nv_aux, int visc_order, double *w_aux, double *fl,
               const double *coeff_deriv1, const double *coeff_deriv2, const double *dcsidx, const double *detady,
               const double *dzitdz, const double *dcsidxs, const double *detadys, const double *dzitdzs,
               const double *dcsidx2, const double *detady2, const double *dzitdz2, void *stream)
{
    (void)nv_aux; (void)stream;
    int lmax = visc_order / 2;
    OMP_FOR
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        double uu = W(w_aux, i, j, k, J_U), vv = W(w_aux, i, j, k, J_V), ww = W(w_aux, i, j, k, J_W);
        double tt = W(w_aux, i, j, k, J_T), mu = W(w_aux, i, j, k, J_MU);
        double k_cond = W(w_aux, i, j, k, J_K_COND), div3l = W(w_aux, i, j, k, J_DIV);
        double ux = 0, vx = 0, wx = 0, tx = 0, mux = 0, divx3l = 0;
        double uy = 0, vy = 0, wy = 0, ty = 0, muy = 0, divy3l = 0;
        double uz = 0, vz = 0, wz = 0, tz = 0, muz = 0, divz3l = 0;
        double ulapx = CD2(0, lmax) * uu, ulapy = ulapx, ulapz = ulapx;
        double vlapx = CD2(0, lmax) * vv, vlapy = vlapx, vlapz = vlapx;
        double wlapx = CD2(0, lmax) * ww, wlapy = wlapx, wlapz = wlapx;
        double tlapx = CD2(0, lmax) * tt, tlapy = tlapx, tlapz = tlapx;
        for (int l = 1; l <= lmax; ++l) {
            double clapl = CD2(l, lmax);
            ulapx = ulapx + clapl * (W(w_aux, i + l, j, k, J_U) + W(w_aux, i - l, j, k, J_U));
            ulapy = ulapy + clapl * (W(w_aux, i, j + l, k, J_U) + W(w_aux, i, j - l, k, J_U));
            ulapz = ulapz + clapl * (W(w_aux, i, j, k + l, J_U) + W(w_aux, i, j, k - l, J_U));
            vlapx = vlapx + clapl * (W(w_aux, i + l, j, k, J_V) + W(w_aux, i - l, j, k, J_V));
            vlapy = vlapy + clapl * (W(w_aux, i, j + l, k, J_V) + W(w_aux, i, j - l, k, J_V));
            vlapz = vlapz + clapl * (W(w_aux, i, j, k + l, J_V) + W(w_aux, i, j, k - l, J_V));
            wlapx = wlapx + clapl * (W(w_aux, i + l, j, k, J_W) + W(w_aux, i - l, j, k, J_W));
            wlapy = wlapy + clapl * (W(w_aux, i, j + l, k, J_W) + W(w_aux, i, j - l, k, J_W));
            wlapz = wlapz + clapl * (W(w_aux, i, j, k + l, J_W) + W(w_aux, i, j, k - l, J_W));
            tlapx = tlapx + clapl * (W(w_aux, i + l, j, k, J_T) + W(w_aux, i - l, j, k, J_T));
            tlapy = tlapy + clapl * (W(w_aux, i, j + l, k, J_T) + W(w_aux, i, j - l, k, J_T));
            tlapz = tlapz + clapl * (W(w_aux, i, j, k + l, J_T) + W(w_aux, i, j, k - l, J_T));
        }
        for (int l = 1; l <= lmax; ++l) {
            double ccl = CD1(l, lmax);
            ux = ux + ccl * (W(w_aux, i + l, j, k, J_U) - W(w_aux, i - l, j, k, J_U));
            vx = vx + ccl * (W(w_aux, i + l, j, k, J_V) - W(w_aux, i - l, j, k, J_V));
            wx = wx + ccl * (W(w_aux, i + l, j, k, J_W) - W(w_aux, i - l, j, k, J_W));
            tx = tx + ccl * (W(w_aux, i + l, j, k, J_T) - W(w_aux, i - l, j, k, J_T));
            mux = mux + ccl * (W(w_aux, i + l, j, k, J_MU) - W(w_aux, i - l, j, k, J_MU));
            divx3l = divx3l + ccl * (W(w_aux, i + l, j, k, J_DIV) - W(w_aux, i - l, j, k, J_DIV));
            uy = uy + ccl * (W(w_aux, i, j + l, k, J_U) - W(w_aux, i, j - l, k, J_U));
            vy = vy + ccl * (W(w_aux, i, j + l, k, J_V) - W(w_aux, i, j - l, k, J_V));
            wy = wy + ccl * (W(w_aux, i, j + l, k, J_W) - W(w_aux, i, j - l, k, J_W));
            ty = ty + ccl * (W(w_aux, i, j + l, k, J_T) - W(w_aux, i, j - l, k, J_T));
            muy = muy + ccl * (W(w_aux, i, j + l, k, J_MU) - W(w_aux, i, j - l, k, J_MU));
            divy3l = divy3l + ccl * (W(w_aux, i, j + l, k, J_DIV) - W(w_aux, i, j - l, k, J_DIV));
            uz = uz + ccl * (W(w_aux, i, j, k + l, J_U) - W(w_aux, i, j, k - l, J_U));
            vz = vz + ccl * (W(w_aux, i, j, k + l, J_V) - W(w_aux, i, j, k - l, J_V));
            wz = wz + ccl * (W(w_aux, i, j, k + l, J_W) - W(w_aux, i, j, k - l, J_W));
            tz = tz + ccl * (W(w_aux, i, j, k + l, J_T) - W(w_aux, i, j, k - l, J_T));
            muz = muz + ccl * (W(w_aux, i, j, k + l, J_MU) - W(w_aux, i, j, k - l, J_MU));
            divz3l = divz3l + ccl * (W(w_aux, i, j, k + l, J_DIV) - W(w_aux, i, j, k - l, J_DIV));
        }
        double dx = dcsidx[i - 1], dy = detady[j - 1], dz = dzitdz[k - 1];
        ux *= dx; vx *= dx; wx *= dx; tx *= dx; mux *= dx; divx3l *= dx;
        uy *= dy; vy *= dy; wy *= dy; ty *= dy; muy *= dy; divy3l *= dy;
        uz *= dz; vz *= dz; wz *= dz; tz *= dz; muz *= dz; divz3l *= dz;
        ulapx = ulapx * dcsidxs[i - 1] + ux * dcsidx2[i - 1];
        vlapx = vlapx * dcsidxs[i - 1] + vx * dcsidx2[i - 1];
        wlapx = wlapx * dcsidxs[i - 1] + wx * dcsidx2[i - 1];
        tlapx = tlapx * dcsidxs[i - 1] + tx * dcsidx2[i - 1];
        ulapy = ulapy * detadys[j - 1] + uy * detady2[j - 1];
        vlapy = vlapy * detadys[j - 1] + vy * detady2[j - 1];
        wlapy = wlapy * detadys[j - 1] + wy * detady2[j - 1];
        tlapy = tlapy * detadys[j - 1] + ty * detady2[j - 1];
        ulapz = ulapz * dzitdzs[k - 1] + uz * dzitdz2[k - 1];
        vlapz = vlapz * dzitdzs[k - 1] + vz * dzitdz2[k - 1];
        wlapz = wlapz * dzitdzs[k - 1] + wz * dzitdz2[k - 1];
        tlapz = tlapz * dzitdzs[k - 1] + tz * dzitdz2[k - 1];
        double ulap = ulapx + ulapy + ulapz, vlap = vlapx + vlapy + vlapz;
        double wlap = wlapx + wlapy + wlapz, tlap = tlapx + tlapy + tlapz;
        double sig11 = 2.0 * (ux - div3l), sig12 = uy + vx, sig13 = uz + wx;
        double sig22 = 2.0 * (vy - div3l), sig23 = vz + wy, sig33 = 2.0 * (wz - div3l);
        double sigx = mux * sig11 + muy * sig12 + muz * sig13 + mu * (ulap + divx3l);
        double sigy = mux * sig12 + muy * sig22 + muz * sig23 + mu * (vlap + divy3l);
        double sigz = mux * sig13 + muy * sig23 + muz * sig33 + mu * (wlap + divz3l);
        double sigqt = (mux * tx + muy * ty + muz * tz + mu * tlap) * k_cond / mu;
        double sigah = (sig11 * ux + sig12 * uy + sig13 * uz + sig12 * vx + sig22 * vy + sig23 * vz + sig13 * wx +
                        sig23 * wy + sig33 * wz) * mu;
        double sigq = sigx * uu + sigy * vv + sigz * ww + sigah + sigqt;
        F(fl, i, j, k, I_U) = F(fl, i, j, k, I_U) - sigx;
        F(fl, i, j, k, I_V) = F(fl, i, j, k, I_V) - sigy;
        F(fl, i, j, k, I_W) = F(fl, i, j, k, I_W) - sigz;
        F(fl, i, j, k, I_E) = F(fl, i, j, k, I_E) - sigq;
    }
    return 0;
}

/* ===================================================================================== */
/* a7: explicit chemistry                                                                 */
/* ===================================================================================== */

/* Reaction-rate part shared by compute_chemistry_cuf (kernels_gpu.F90:9551-9649) and
 * rosenbrock_rhs (:9249-9339): for each reaction the rate of progress q_r from the
 * concentrations; calls back into `acc` (species source accumulation) per reaction. */
static double forward_rate(int nreactions, int lr, int ns, double tt, double Rtti, const double *conc,
                           const double *arr_a, const double *arr_b, const double *arr_ea,
                           const double *tb_eff, const double *falloff_coeffs, int ty)
{
    double kf = 0.0, k0, kinf, tb, rpres;
    switch (ty) {
    case 0: case 1:
        kf = RX(arr_a, lr, 1);
        if (RX(arr_b, lr, 1) != 0.0) kf = kf * pow(tt, RX(arr_b, lr, 1));
        if (RX(arr_ea, lr, 1) != 0.0) kf = kf * exp(-RX(arr_ea, lr, 1) * Rtti);
        break;
    case 2:
        k0 = RX(arr_a, lr, 1) * pow(tt, RX(arr_b, lr, 1)) * exp(-RX(arr_ea, lr, 1) * Rtti);
        kinf = RX(arr_a, lr, 2) * pow(tt, RX(arr_b, lr, 2)) * exp(-RX(arr_ea, lr, 2) * Rtti);
        tb = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) tb = tb + RX(tb_eff, lr, lsp) * conc[lsp - 1];
        rpres = k0 * tb / kinf;
        kf = kinf * (rpres / (1.0 + rpres));
        break;
    case 3:
        k0 = RX(arr_a, lr, 1) * pow(tt, RX(arr_b, lr, 1)) * exp(-RX(arr_ea, lr, 1) * Rtti);
        kinf = RX(arr_a, lr, 2) * pow(tt, RX(arr_b, lr, 2)) * exp(-RX(arr_ea, lr, 2) * Rtti);
        tb = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) tb = tb + RX(tb_eff, lr, lsp) * conc[lsp - 1];
        rpres = k0 * tb / kinf;
        if (rpres < 1e-30) {
            kf = k0 * tb;
        } else {
            double fcent = (1.0 - RX(falloff_coeffs, lr, 1)) * exp(-tt / RX(falloff_coeffs, lr, 2)) +
                           RX(falloff_coeffs, lr, 1) * exp(-tt / RX(falloff_coeffs, lr, 3));
            if (RX(falloff_coeffs, lr, 4) != -3.14) fcent = fcent + exp(-RX(falloff_coeffs, lr, 4) / tt);
            fcent = log10(fcent);
            double coff = -0.4 - 0.67 * fcent;
            double noff = 0.75 - 1.27 * fcent;
            double doff = 0.14;
            double lr10 = log10(rpres);
            double fo = (lr10 + coff) / (noff - doff * (lr10 + coff));
            double foff = pow(10.0, pow(1.0 + pow(fo, 2.0), -1.0) * fcent);
            kf = kinf * (rpres / (1.0 + rpres)) * foff;
        }
        break;
    case 4:
        k0 = RX(arr_a, lr, 1) * pow(tt, RX(arr_b, lr, 1)) * exp(-RX(arr_ea, lr, 1) * Rtti);
        kinf = RX(arr_a, lr, 2) * pow(tt, RX(arr_b, lr, 2)) * exp(-RX(arr_ea, lr, 2) * Rtti);
        tb = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) tb = tb + RX(tb_eff, lr, lsp) * conc[lsp - 1];
        rpres = k0 * tb / kinf;
        {
            double xoff = 1.0 / (1.0 + pow(log10(rpres), 2.0));
            double foff = RX(falloff_coeffs, lr, 4) * pow(RX(falloff_coeffs, lr, 1) * exp(-RX(falloff_coeffs, lr, 2) / tt) +
                          exp(-tt / RX(falloff_coeffs, lr, 3)), xoff) * pow(tt, RX(falloff_coeffs, lr, 5));
            kf = kinf * (rpres / (1.0 + rpres)) * foff;
        }
        break;
    }
    return kf;
}

static double progress_rate(int nreactions, int lr, int ns, double tt, double Rtti, int itt, double dtt,
                            int num_t_tab, const double *conc, const double *arr_a, const double *arr_b,
                            const double *arr_ea, const double *tb_eff, const double *falloff_coeffs,
                            const int *reac_ty, const int *isrev, const double *r_coeffs,
                            const double *p_coeffs, const double *kc_tab)
{
    int ty = reac_ty[lr - 1];
    double kf = forward_rate(nreactions, lr, ns, tt, Rtti, conc, arr_a, arr_b, arr_ea, tb_eff, falloff_coeffs, ty);
    double kb;
    if (isrev[lr - 1] == 1) {
        double dkc = KC(itt + 1, lr) - KC(itt, lr);
        double kc = KC(itt, lr) + dkc * dtt;
        kb = kf / kc;
    } else {
        kb = 0.0;
    }
    double q1 = 1.0, q2 = 1.0;
    for (int lsp = 1; lsp <= ns; ++lsp) {
        double rr = RX(r_coeffs, lr, lsp), pp = RX(p_coeffs, lr, lsp);
        if (rr == 1.0) q1 = q1 * conc[lsp - 1];
        else if (rr != 0.0) q1 = q1 * pow(conc[lsp - 1], rr);
        if (pp == 1.0) q2 = q2 * conc[lsp - 1];
        else if (pp != 0.0) q2 = q2 * pow(conc[lsp - 1], pp);
    }
    double tb = 1.0;
    if (ty == 1) {
        tb = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) tb = tb + RX(tb_eff, lr, lsp) * conc[lsp - 1];
    }
    return tb * (kf * q1 - kb * q2);
}

int orc_compute_chemistry(int ns, int nx, int ny, int nz, int nv, int nv_aux, int ng,
                          int nreactions, double *w, double *w_aux, double *fl,
                          const double *mw, const double *arr_a, const double *arr_b,
                          const double *arr_ea, const double *tb_eff,
                          const double *falloff_coeffs, const int *reac_ty,
                          const int *isrev, const double *r_coeffs,
                          const double *p_coeffs, const double *kc_tab, int num_t_tab,
                          double t_min_tab, double dt_tab, double r_univ, int indx_cp_l,
                          int indx_cp_r, const double *cv_coeff, int nsetcv,
                          const double *trange, double tol_iter_nr, void *stream)
{
    (void)nv; (void)nv_aux; (void)stream;
    OMP_FOR
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        double conc[NSMAX], yy[NSMAX];
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double v = W(w, i, j, k, lsp);
            W(w, i, j, k, lsp) = v > 0.0 ? v : 0.0;
        }
        double rho = W(w, i, j, k, 1);
        for (int lsp = 2; lsp <= ns; ++lsp) rho = rho + W(w, i, j, k, lsp);
        double ri = 1.0 / rho;
        for (int lsp = 1; lsp <= ns; ++lsp) {
            W(w_aux, i, j, k, lsp) = W(w, i, j, k, lsp) * ri;
            yy[lsp - 1] = W(w_aux, i, j, k, lsp);
            conc[lsp - 1] = rho * yy[lsp - 1] / mw[lsp - 1];
        }
        double uu = W(w, i, j, k, I_U) * ri, vv = W(w, i, j, k, I_V) * ri, ww = W(w, i, j, k, I_W) * ri;
        double rhoe = W(w, i, j, k, I_E);
        double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
        double ee = rhoe * ri - qq;
        double tt = temperature_from_e(ns, ee, W(w_aux, i, j, k, J_T), cv_coeff, nsetcv, trange,
                                       indx_cp_l, indx_cp_r, tol_iter_nr, yy);
        W(w_aux, i, j, k, J_T) = tt;
        double Rtti = 1.0 / (r_univ * tt);
        int itt = (int)((tt - t_min_tab) / dt_tab) + 1;
        if (itt < 1) itt = 1;
        if (itt > num_t_tab) itt = num_t_tab;
        double ttleft = (itt - 1) * dt_tab + t_min_tab;
        double dtt = (tt - ttleft) / dt_tab;
        for (int lr = 1; lr <= nreactions; ++lr) {
            double qlr = progress_rate(nreactions, lr, ns, tt, Rtti, itt, dtt, num_t_tab, conc, arr_a, arr_b,
                                       arr_ea, tb_eff, falloff_coeffs, reac_ty, isrev, r_coeffs, p_coeffs, kc_tab);
            for (int lsp = 1; lsp <= ns; ++lsp) {
                double wdtkj = (RX(p_coeffs, lr, lsp) - RX(r_coeffs, lr, lsp)) * qlr * mw[lsp - 1];
                F(fl, i, j, k, lsp) = F(fl, i, j, k, lsp) - wdtkj;
            }
        }
    }
    return 0;
}

/* ===================================================================================== */
/* a8: Rosenbrock stiff integrator (kernels_gpu.F90:8928-9501)                            */
/* ===================================================================================== */
typedef struct {
    int ns, nv, nreactions, num_t_tab, nsetcv, indx_cp_l, indx_cp_r;
    const double *mw, *arr_a, *arr_b, *arr_ea, *tb_eff, *falloff_coeffs, *r_coeffs, *p_coeffs, *kc_tab;
    const int *reac_ty, *isrev;
    double t_min_tab, dt_tab, r_univ, tol_iter_nr;
    const double *cv_coeff, *trange;
    /* constant-pressure reactor (pres > 0): state (Y_1..Y_NS, 0, 0, 0, h) */
    const double *cp_coeff, *rgas;
    double pres;
    /* PaSR closure of the LES path (pasr != 0): the cell's mixing time tau_k */
    int pasr;
    double tau_k;
} chem_t;

/* The PaSR factor of rosenbrock_rhs (kernels_gpu.F90:9359-9379) and of
 * rosenbrock_krylov_rhs (:10954-10974): the chemical time scale tau_c is the geometric mean
 * of the slowest and the fastest species time Y_s / |wdot_s / rho| over the species with
 * |wdot_s / rho| > 1e-8, and the species sources are scaled by tau_c / (tau_c + tau_k)
 * (by 1 when tau_k is not above the smallest normal double). */
static void pasr_scale(int ns, double tau_k, double ri, const double *yy, double *rhs)
{
    double tau_cf = 0.0, tau_cs = DBL_MAX;
    for (int lsp = 0; lsp < ns; ++lsp) {
        if (fabs(rhs[lsp] * ri) > 1.0e-8) {
            double tau_lsp = yy[lsp] / fabs(rhs[lsp] * ri);
            if (tau_lsp > tau_cf) tau_cf = tau_lsp;
            if (tau_lsp > 0.0 && tau_lsp < tau_cs) tau_cs = tau_lsp;
        }
    }
    double tau_c = sqrt(tau_cs * tau_cf);
    double gam_pasr = tau_k > DBL_MIN ? tau_c / (tau_c + tau_k) : 1.0;
    for (int lsp = 0; lsp < ns; ++lsp) rhs[lsp] = gam_pasr * rhs[lsp];
}

/* The cell's mixing time of the PaSR closure (kernels_gpu.F90:8970-8980, 9720-9728):
 * Delta = (dxi/dx deta/dy dzeta/dz)^(-1/3), nu_sgs = mu_sgs / rho,
 * eps_sgs = c_eps c_yoshi^(3/2) nu_sgs^3 / Delta^4, tau_k = c_mix sqrt(nu_eff / eps_sgs)
 * with nu_eff = mu / rho (mu already holds mu_sgs). */
static double pasr_tau_k(double rho, double mu_sgs, double mu, double dx, double dy, double dz,
                         double les_c_yoshi, double les_c_mix, double les_c_eps)
{
    double ri = 1.0 / rho;
    double delta = pow(dx * dy * dz, -1.0 / 3.0);
    double nu_sgs = mu_sgs * ri;
    double d2 = delta * delta;
    double eps_sgs = les_c_eps * pow(les_c_yoshi, 1.5) * (nu_sgs * nu_sgs * nu_sgs) / (d2 * d2);
    double nu_eff = mu * ri;
    double r = nu_eff / (eps_sgs + DBL_MIN);
    return les_c_mix * sqrt(r > 0.0 ? r : 0.0);
}

static void ros_rhs_cv(const chem_t *C, double *wv, double *yy, double *tguess, double *rhs);
static void ros_rhs_cp(const chem_t *C, double *wv, double *yy, double *tguess, double *rhs);

/* The chemistry right-hand side of one reactor: the reference's constant-volume cell, or
 * the constant-pressure reactor when C->pres > 0. */
static void ros_rhs(const chem_t *C, double *wv, double *yy, double *tguess, double *rhs)
{
    if (C->pres > 0.0) ros_rhs_cp(C, wv, yy, tguess, rhs);
    else ros_rhs_cv(C, wv, yy, tguess, rhs);
}

/* Constant-pressure adiabatic reactor (no counterpart in the reference, whose reactor is a
 * constant-volume cell).  State wv = (Y_1..Y_NS, 0, 0, 0, h), h the mixture enthalpy per
 * unit mass, constant in time.  Written in rosenbrock_rhs's arithmetic forms
 * (kernels_gpu.F90:9249-9339): species clipped at 0 and renormalised, T from h by the
 * Newton iteration of get_mixture_temperature_from_h_roe_dev (:6810-6900, the e-inversion
 * on the c_p tables), rho = p / (R_mix T), concentrations rho Y_s / W_s, the same rates of
 * progress, and dY_s/dt = sum_r ((p - r) q_r) W_s / rho. */
static void ros_rhs_cp(const chem_t *C, double *wv, double *yy, double *tguess, double *rhs)
{
    int ns = C->ns, nv = C->nv, nreactions = C->nreactions, num_t_tab = C->num_t_tab;
    int nsetcv = C->nsetcv, indx_cp_l = C->indx_cp_l, indx_cp_r = C->indx_cp_r;
    double conc[NSMAX];
    for (int lsp = 1; lsp <= ns; ++lsp) wv[lsp - 1] = wv[lsp - 1] > 0.0 ? wv[lsp - 1] : 0.0;
    double ysum = wv[0];
    for (int lsp = 2; lsp <= ns; ++lsp) ysum = ysum + wv[lsp - 1];
    double yi = 1.0 / ysum;
    double rmix = 0.0;
    for (int lsp = 1; lsp <= ns; ++lsp) {
        yy[lsp - 1] = wv[lsp - 1] * yi;
        rmix = rmix + C->rgas[lsp - 1] * yy[lsp - 1];
    }
    double tt = temperature_from_e(ns, wv[I_E - 1], *tguess, C->cp_coeff, nsetcv, C->trange, indx_cp_l, indx_cp_r,
                                   C->tol_iter_nr, yy);
    *tguess = tt;
    double rho = C->pres / (rmix * tt);
    double ri = 1.0 / rho;
    for (int lsp = 1; lsp <= ns; ++lsp) conc[lsp - 1] = rho * yy[lsp - 1] / C->mw[lsp - 1];
    double Rtti = 1.0 / (C->r_univ * tt);
    int itt = (int)((tt - C->t_min_tab) / C->dt_tab) + 1;
    if (itt < 1) itt = 1;
    if (itt > num_t_tab) itt = num_t_tab;
    double ttleft = (itt - 1) * C->dt_tab + C->t_min_tab;
    double dtt = (tt - ttleft) / C->dt_tab;
    for (int m = 0; m < nv; ++m) rhs[m] = 0.0;
    for (int lr = 1; lr <= nreactions; ++lr) {
        double qlr = progress_rate(nreactions, lr, ns, tt, Rtti, itt, dtt, num_t_tab, conc, C->arr_a, C->arr_b,
                                   C->arr_ea, C->tb_eff, C->falloff_coeffs, C->reac_ty, C->isrev, C->r_coeffs,
                                   C->p_coeffs, C->kc_tab);
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double rr = RX(C->r_coeffs, lr, lsp), pp = RX(C->p_coeffs, lr, lsp);
            if (rr != 0.0 || pp != 0.0) {
                double wdtkj = (pp - rr) * qlr;
                rhs[lsp - 1] = rhs[lsp - 1] + wdtkj * C->mw[lsp - 1];
            }
        }
    }
    for (int lsp = 1; lsp <= ns; ++lsp) rhs[lsp - 1] = rhs[lsp - 1] * ri;
}

/* rosenbrock_rhs (kernels_gpu.F90:9195-9381) on the cell state wv[nv]; clips species in
 * wv, updates yy and the temperature guess *tguess */
static void ros_rhs_cv(const chem_t *C, double *wv, double *yy, double *tguess, double *rhs)
{
    int ns = C->ns, nv = C->nv, nreactions = C->nreactions, num_t_tab = C->num_t_tab;
    int nsetcv = C->nsetcv, indx_cp_l = C->indx_cp_l, indx_cp_r = C->indx_cp_r;
    double conc[NSMAX];
    for (int lsp = 1; lsp <= ns; ++lsp) wv[lsp - 1] = wv[lsp - 1] > 0.0 ? wv[lsp - 1] : 0.0;
    double rho = wv[0];
    for (int lsp = 2; lsp <= ns; ++lsp) rho = rho + wv[lsp - 1];
    double ri = 1.0 / rho;
    for (int lsp = 1; lsp <= ns; ++lsp) {
        yy[lsp - 1] = wv[lsp - 1] * ri;
        conc[lsp - 1] = rho * yy[lsp - 1] / C->mw[lsp - 1];
    }
    double uu = wv[I_U - 1] * ri, vv = wv[I_V - 1] * ri, ww = wv[I_W - 1] * ri;
    double rhoe = wv[I_E - 1];
    double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
    double ee = rhoe * ri - qq;
    double tt = temperature_from_e(ns, ee, *tguess, C->cv_coeff, nsetcv, C->trange, indx_cp_l, indx_cp_r,
                                   C->tol_iter_nr, yy);
    *tguess = tt;
    double Rtti = 1.0 / (C->r_univ * tt);
    int itt = (int)((tt - C->t_min_tab) / C->dt_tab) + 1;
    if (itt < 1) itt = 1;
    if (itt > num_t_tab) itt = num_t_tab;
    double ttleft = (itt - 1) * C->dt_tab + C->t_min_tab;
    double dtt = (tt - ttleft) / C->dt_tab;
    for (int m = 0; m < nv; ++m) rhs[m] = 0.0;
    for (int lr = 1; lr <= nreactions; ++lr) {
        double qlr = progress_rate(nreactions, lr, ns, tt, Rtti, itt, dtt, num_t_tab, conc, C->arr_a, C->arr_b,
                                   C->arr_ea, C->tb_eff, C->falloff_coeffs, C->reac_ty, C->isrev, C->r_coeffs,
                                   C->p_coeffs, C->kc_tab);
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double rr = RX(C->r_coeffs, lr, lsp), pp = RX(C->p_coeffs, lr, lsp);
            if (rr != 0.0 || pp != 0.0) {
                double wdtkj = (pp - rr) * qlr;
                rhs[lsp - 1] = rhs[lsp - 1] + wdtkj * C->mw[lsp - 1];
            }
        }
    }
    if (C->pasr) pasr_scale(ns, C->tau_k, ri, yy, rhs);
}

/* ludcmp (kernels_gpu.F90:9395-9462), Crout with implicit partial pivoting.  A row whose
 * largest magnitude is below tiny (or NaN) ends it early, as in the reference, which leaves
 * INDX unset there; indx starts as the identity so that lubksb stays inside its vector. */
static void ludcmp(int n, double *a, int *indx)
{
    const double tiny = 1.5e-16;
    double vv[NVMAX];
#define A(r, c) a[((r) - 1) + (size_t)n * ((c) - 1)]
    for (int i = 1; i <= n; ++i) indx[i - 1] = i;
    for (int i = 1; i <= n; ++i) {
        double amax = 0.0;
        for (int j = 1; j <= n; ++j) if (fabs(A(i, j)) > amax) amax = fabs(A(i, j));
        if (amax < tiny) return;
        vv[i - 1] = 1.0 / amax;
    }
    int imax = 1;
    for (int j = 1; j <= n; ++j) {
        for (int i = 1; i <= j - 1; ++i) {
            double summ = A(i, j);
            for (int k = 1; k <= i - 1; ++k) summ = summ - A(i, k) * A(k, j);
            A(i, j) = summ;
        }
        double amax = 0.0;
        for (int i = j; i <= n; ++i) {
            double summ = A(i, j);
            for (int k = 1; k <= j - 1; ++k) summ = summ - A(i, k) * A(k, j);
            A(i, j) = summ;
            double dum = vv[i - 1] * fabs(summ);
            if (dum >= amax) { imax = i; amax = dum; }
        }
        if (j != imax) {
            for (int k = 1; k <= n; ++k) { double dum = A(imax, k); A(imax, k) = A(j, k); A(j, k) = dum; }
            vv[imax - 1] = vv[j - 1];
        }
        indx[j - 1] = imax;
        if (fabs(A(j, j)) < tiny) A(j, j) = tiny;
        if (j != n) {
            double dum = 1.0 / A(j, j);
            for (int i = j + 1; i <= n; ++i) A(i, j) = A(i, j) * dum;
        }
    }
}

/* lubksb (kernels_gpu.F90:9464-9501) */
static void lubksb(int n, const double *a, const int *indx, double *b)
{
    int ii = 0;
    for (int i = 1; i <= n; ++i) {
        int ll = indx[i - 1];
        double summ = b[ll - 1];
        b[ll - 1] = b[i - 1];
        if (ii != 0) {
            for (int j = ii; j <= i - 1; ++j) summ = summ - A(i, j) * b[j - 1];
        } else if (summ != 0.0) {
            ii = i;
        }
        b[i - 1] = summ;
    }
    for (int i = n; i >= 1; --i) {
        double summ = b[i - 1];
        for (int j = i + 1; j <= n; ++j) summ = summ - A(i, j) * b[j - 1];
        b[i - 1] = summ / A(i, i);
    }
#undef A
}

/* Solve a x = b with ludcmp + lubksb (kernels_gpu.F90:9395-9501): a is n x n column-major
 * and is overwritten by its LU factors, b by x.  Exported for the brute-force checks. */
int orc_lu_solve(int n, double *a, double *b)
{
    int indx[NVMAX];
    if (n < 1 || n > NVMAX) return -2;
    ludcmp(n, a, indx);
    lubksb(n, a, indx, b);
    return 0;
}

/* Integrate one cell's state wv[nv] from time_start to time_end (rosenbrock_kernel body).
 * Returns accepted steps in *nacc and rejected tries in *nrej. */
int orc_rosenbrock_cell(const chem_t *C, double *wv, double *tguess, double dttry, double time_start,
                        double time_end, int maxsteps, int maxtry, double tol, int *nacc, int *nrej)
{
    int nv = C->nv, ns = C->ns;
    const double safety = 0.9, grow = 1.5, pgrow = -0.25, shrnk = 0.5, pshrnk = -1.0 / 3.0, errcon = 0.1296;
    const double gam = 1.0 / 2.0, a21 = 2.0, a31 = 48.0 / 25.0, a32 = 6.0 / 25.0;
    const double c21 = -8.0, c31 = 372.0 / 25.0, c32 = 12.0 / 5.0;
    const double c41 = -112.0 / 125.0, c42 = -54.0 / 125.0, c43 = -2.0 / 5.0;
    const double b1 = 19.0 / 9.0, b2 = 1.0 / 2.0, b3 = 25.0 / 108.0, b4 = 125.0 / 108.0;
    const double e1 = 17.0 / 54.0, e2 = 7.0 / 36.0, e3 = 0.0, e4 = 125.0 / 108.0;
    double jac[NVMAX * NVMAX], amat[NVMAX * NVMAX];
    double w_sav[NVMAX], rhs_sav[NVMAX], err[NVMAX], yscal[NVMAX], rhs[NVMAX], rhsp[NVMAX], rhsm[NVMAX];
    double g1[NVMAX], g2[NVMAX], g3[NVMAX], g4[NVMAX], yy[NSMAX];
    int indx[NVMAX];
    (void)ns;
    int finish = 0;
    double time = time_start, dt = dttry, itol = 1.0 / tol, dtnext = dt;
    *nacc = 0; *nrej = 0;
    for (int iter = 1; iter <= maxsteps; ++iter) {
        double t0 = time;
        /* rosenbrock_jacobian (:9141-9193) */
        for (int n = 1; n <= nv; ++n) {
            double wtmp = wv[n - 1];
            double h = wtmp * 1.0e-06 + 1.0e-14;
            wv[n - 1] = wtmp + h;
            ros_rhs(C, wv, yy, tguess, rhsp);
            wv[n - 1] = wtmp - h;
            ros_rhs(C, wv, yy, tguess, rhsm);
            wv[n - 1] = wtmp;
            for (int m = 1; m <= nv; ++m) jac[(m - 1) + (size_t)nv * (n - 1)] = (rhsp[m - 1] - rhsm[m - 1]) / (2.0 * h);
        }
        ros_rhs(C, wv, yy, tguess, rhs_sav);
        for (int m = 0; m < nv; ++m) {
            w_sav[m] = wv[m];
            yscal[m] = fabs(w_sav[m]) > 1.0 ? fabs(w_sav[m]) : 1.0;
        }
        double errmax = 0.0;
        for (int jtry = 1; jtry <= maxtry; ++jtry) {
            /* rosenbrock_step (:9046-9139) */
            for (int n = 1; n <= nv; ++n) {
                for (int m = 1; m <= nv; ++m) amat[(m - 1) + (size_t)nv * (n - 1)] = -jac[(m - 1) + (size_t)nv * (n - 1)];
                amat[(n - 1) + (size_t)nv * (n - 1)] = amat[(n - 1) + (size_t)nv * (n - 1)] + 1.0 / (gam * dt);
            }
            ludcmp(nv, amat, indx);
            for (int m = 0; m < nv; ++m) g1[m] = rhs_sav[m];
            lubksb(nv, amat, indx, g1);
            for (int m = 0; m < nv; ++m) wv[m] = w_sav[m] + a21 * g1[m];
            time = t0 + 1.0 * dt;
            ros_rhs(C, wv, yy, tguess, rhs);
            for (int m = 0; m < nv; ++m) g2[m] = rhs[m] + c21 * g1[m] / dt;
            lubksb(nv, amat, indx, g2);
            for (int m = 0; m < nv; ++m) wv[m] = w_sav[m] + a31 * g1[m] + a32 * g2[m];
            time = t0 + 3.0 / 5.0 * dt;
            ros_rhs(C, wv, yy, tguess, rhs);
            for (int m = 0; m < nv; ++m) g3[m] = rhs[m] + (c31 * g1[m] + c32 * g2[m]) / dt;
            lubksb(nv, amat, indx, g3);
            for (int m = 0; m < nv; ++m) g4[m] = rhs[m] + (c41 * g1[m] + c42 * g2[m] + c43 * g3[m]) / dt;
            lubksb(nv, amat, indx, g4);
            for (int m = 0; m < nv; ++m) {
                wv[m] = w_sav[m] + b1 * g1[m] + b2 * g2[m] + b3 * g3[m] + b4 * g4[m];
                err[m] = e1 * g1[m] + e2 * g2[m] + e3 * g3[m] + e4 * g4[m];
            }
            time = t0 + dt;
            errmax = 0.0;
            for (int m = 0; m < nv; ++m) { double e = fabs(err[m] / yscal[m]); if (e > errmax) errmax = e; }
            errmax = errmax * itol;
            if (errmax < 1.0 || finish == 1) {
                if (errmax > errcon) dtnext = safety * dt * pow(errmax, pgrow);
                else dtnext = grow * dt;
                break;
            } else {
                *nrej += 1;
                dtnext = safety * dt * pow(errmax, pshrnk);
                double a = fabs(dtnext), b = shrnk * fabs(dt);
                dt = copysign(a > b ? a : b, dt);
            }
        }
        *nacc += 1;
        double remaining = time_end - time;
        if (finish == 1 || remaining < 1.0e-12 * time_end) break;
        if (dtnext * 1.5 > remaining) { dt = remaining; finish = 1; }
        else { dt = dtnext; finish = 0; }
    }
    return 0;
}

int orc_rosenbrock(int ns, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                   double *w, double *w_aux, const double *fl, const double *mw,
                   int nreactions, const double *arr_a, const double *arr_b,
                   const double *arr_ea, const double *tb_eff,
                   const double *falloff_coeffs, const int *reac_ty, const int *isrev,
                   const double *r_coeffs, const double *p_coeffs,
                   const double *kc_tab, int num_t_tab, double t_min_tab,
                   double dt_tab, double r_univ, double tol_iter_nr,
                   const double *cv_coeff, int nsetcv, const double *trange,
                   int indx_cp_l, int indx_cp_r, double dttry, double time_start,
                   double time_end, int maxsteps, int maxtry, double tol,
                   int simpler_splitting, int enable_pasr, double les_c_yoshi,
                   double les_c_mix, double les_c_eps, const double *dcsidx,
                   const double *detady, const double *dzitdz, void *stream)
{
    (void)fl; (void)simpler_splitting; (void)stream;
    if (enable_pasr != 0 && nv_aux < J_LES1) return -2;
    chem_t C = {ns, nv, nreactions, num_t_tab, nsetcv, indx_cp_l, indx_cp_r, mw, arr_a, arr_b, arr_ea, tb_eff,
                falloff_coeffs, r_coeffs, p_coeffs, kc_tab, reac_ty, isrev, t_min_tab, dt_tab, r_univ,
                tol_iter_nr, cv_coeff, trange};
    OMP_FOR
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        double wv[NVMAX];
        for (int m = 1; m <= nv; ++m) wv[m - 1] = W(w, i, j, k, m);
        double tg = W(w_aux, i, j, k, J_T);
        int na, nr;
        chem_t Cc = C;
        if (enable_pasr > 0) {
            Cc.pasr = 1;
            Cc.tau_k = pasr_tau_k(W(w_aux, i, j, k, J_R), W(w_aux, i, j, k, J_LES1), W(w_aux, i, j, k, J_MU),
                                  dcsidx[i - 1], detady[j - 1], dzitdz[k - 1], les_c_yoshi, les_c_mix, les_c_eps);
        }
        orc_rosenbrock_cell(&Cc, wv, &tg, dttry, time_start, time_end, maxsteps, maxtry, tol, &na, &nr);
        for (int m = 1; m <= nv; ++m) W(w, i, j, k, m) = wv[m - 1];
        /* the kernel leaves w_aux T at the last right-hand-side evaluation (w_aux Y,
         * also scratch there, is rebuilt by the next eval_aux and not written here) */
        W(w_aux, i, j, k, J_T) = tg;
    }
    return 0;
}

/* Batched interface over a flat list of reactors (state w[nv][nbatch] SoA, T guess
 * tg[nbatch]) used by the batched-chemistry tests and the CPU baseline. */
int orc_reactor_batch(int ns, int nbatch, double *w, double *tg, const double *mw, int nreactions,
                      const double *arr_a, const double *arr_b, const double *arr_ea,
                      const double *tb_eff, const double *falloff_coeffs, const int *reac_ty,
                      const int *isrev, const double *r_coeffs, const double *p_coeffs,
                      const double *kc_tab, int num_t_tab, double t_min_tab, double dt_tab,
                      double r_univ, double tol_iter_nr, const double *cv_coeff, int nsetcv,
                      const double *trange, int indx_cp_l, int indx_cp_r, const double *dttry,
                      const double *time_end, int maxsteps, int maxtry, double tol,
                      int *nacc, int *nrej, const double *pressure, const double *rgas,
                      const double *cp_coeff);

int orc_rosenbrock_batch(int ns, int nbatch, double *w, double *tg, const double *mw, int nreactions,
                         const double *arr_a, const double *arr_b, const double *arr_ea,
                         const double *tb_eff, const double *falloff_coeffs, const int *reac_ty,
                         const int *isrev, const double *r_coeffs, const double *p_coeffs,
                         const double *kc_tab, int num_t_tab, double t_min_tab, double dt_tab,
                         double r_univ, double tol_iter_nr, const double *cv_coeff, int nsetcv,
                         const double *trange, int indx_cp_l, int indx_cp_r, const double *dttry,
                         const double *time_end, int maxsteps, int maxtry, double tol,
                         int *nacc, int *nrej)
{
    return orc_reactor_batch(ns, nbatch, w, tg, mw, nreactions, arr_a, arr_b, arr_ea, tb_eff, falloff_coeffs,
                             reac_ty, isrev, r_coeffs, p_coeffs, kc_tab, num_t_tab, t_min_tab, dt_tab, r_univ,
                             tol_iter_nr, cv_coeff, nsetcv, trange, indx_cp_l, indx_cp_r, dttry, time_end, maxsteps,
                             maxtry, tol, nacc, nrej, NULL, NULL, NULL);
}

/* The same over constant-pressure reactors when pressure != NULL (pressure[b] in Pa, state
 * (Y, 0, 0, 0, h), rgas and cp_coeff required), as streams_reactor_batch takes them. */
int orc_reactor_batch(int ns, int nbatch, double *w, double *tg, const double *mw, int nreactions,
                      const double *arr_a, const double *arr_b, const double *arr_ea,
                      const double *tb_eff, const double *falloff_coeffs, const int *reac_ty,
                      const int *isrev, const double *r_coeffs, const double *p_coeffs,
                      const double *kc_tab, int num_t_tab, double t_min_tab, double dt_tab,
                      double r_univ, double tol_iter_nr, const double *cv_coeff, int nsetcv,
                      const double *trange, int indx_cp_l, int indx_cp_r, const double *dttry,
                      const double *time_end, int maxsteps, int maxtry, double tol,
                      int *nacc, int *nrej, const double *pressure, const double *rgas,
                      const double *cp_coeff)
{
    int nv = ns + 4;
    if (pressure != NULL && (rgas == NULL || cp_coeff == NULL)) return -2;
    OMP_FOR
    for (int b = 0; b < nbatch; ++b) {
        chem_t C = {ns, nv, nreactions, num_t_tab, nsetcv, indx_cp_l, indx_cp_r, mw, arr_a, arr_b, arr_ea, tb_eff,
                    falloff_coeffs, r_coeffs, p_coeffs, kc_tab, reac_ty, isrev, t_min_tab, dt_tab, r_univ,
                    tol_iter_nr, cv_coeff, trange, cp_coeff, rgas, pressure ? pressure[b] : 0.0};
        double wv[NVMAX];
        for (int m = 0; m < nv; ++m) wv[m] = w[(size_t)m * nbatch + b];
        int na = 0, nr = 0;
        orc_rosenbrock_cell(&C, wv, &tg[b], dttry[b], 0.0, time_end[b], maxsteps, maxtry, tol, &na, &nr);
        for (int m = 0; m < nv; ++m) w[(size_t)m * nbatch + b] = wv[m];
        if (nacc) nacc[b] = na;
        if (nrej) nrej[b] = nr;
    }
    return 0;
}

/* ===================================================================================== */
/* a9: Rosenbrock-Krylov integrator (kernels_gpu.F90:10360-11205)                         */
/* ===================================================================================== */
#define KRY_M 4        /* M_max, index_define.h:27 */
#define KRY_NSTEP 4    /* nstep, index_define.h:4 */
#define KM(a, r, c) (a)[((r) - 1) + KRY_M * ((c) - 1)]          /* (KRY_M x KRY_M) */

/* NORM2 as the plain root of the sum of squares in index order */
static double kry_norm2(int n, const double *v)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + v[i] * v[i];
    return sqrt(s);
}

/* rosenbrock_krylov_rhs (kernels_gpu.F90:10794-10976): the arithmetic of rosenbrock_rhs
 * (PaSR factor included) on a copy of the state (the species are clipped in the copy only)
 * with the temperature guess passed by value (intent(in) there) */
static void kry_rhs(const chem_t *C, const double *w, double t_iter, double *rhs)
{
    double tmp[NVMAX], yy[NSMAX];
    for (int m = 0; m < C->nv; ++m) tmp[m] = w[m];
    double tg = t_iter;
    ros_rhs(C, tmp, yy, &tg, rhs);
}

/* inv_matrix_krylov (:11087-11119): inverse of the leading d x d block in place, through
 * ludcmp_krylov / lubksb_krylov (the Crout LU of the a8 integrator on that block) */
static void kry_inv(double *a, int d)
{
    if (d == 1) { KM(a, 1, 1) = 1.0 / KM(a, 1, 1); return; }
    double blk[KRY_M * KRY_M], y[KRY_M * KRY_M];
    int indx[KRY_M];
    for (int c = 1; c <= d; ++c)
        for (int r = 1; r <= d; ++r) blk[(r - 1) + d * (c - 1)] = KM(a, r, c);
    ludcmp(d, blk, indx);
    for (int c = 1; c <= d; ++c) {
        for (int r = 1; r <= d; ++r) y[(r - 1) + d * (c - 1)] = r == c ? 1.0 : 0.0;
        lubksb(d, blk, indx, &y[d * (c - 1)]);
    }
    for (int c = 1; c <= d; ++c)
        for (int r = 1; r <= d; ++r) KM(a, r, c) = y[(r - 1) + d * (c - 1)];
}

/* arnoldi (:10664-10792): Krylov basis of the one-sided finite-difference Jacobian at
 * w_sav started from f(w_sav); stops at the first size kk with
 * |dt gam H(kk+1,kk)| |lambda(kk)| < 1e-6 and then returns H, the basis Vb and
 * inv = (I - dt gam H)^-1; when no size up to KRY_M passes, H, Vb and inv stay zero. */
static void kry_arnoldi(const chem_t *C, const double *w_sav, double dt, double t_iter, const double *fls,
                        double *Vb, double *inv, double *rhs_sav)
{
    const int nv = C->nv;
    const double gam = 0.572816062482135, TOL = 1.0e-6, eps = 1.0e-08;
    double HH[(KRY_M + 1) * KRY_M], V[NVMAX * (KRY_M + 1)], Vt[KRY_M * NVMAX], HHinv[KRY_M * KRY_M];
    double rpr[KRY_M], lambda[KRY_M], jac[NVMAX], wl[NVMAX], rhsp[NVMAX];
#define HHX(r, c) HH[((r) - 1) + (KRY_M + 1) * ((c) - 1)]
#define VX(r, c) V[((r) - 1) + (size_t)nv * ((c) - 1)]
#define VTX(r, c) Vt[((r) - 1) + KRY_M * ((c) - 1)]
    for (int q = 0; q < (KRY_M + 1) * KRY_M; ++q) HH[q] = 0.0;
    for (int q = 0; q < nv * (KRY_M + 1); ++q) V[q] = 0.0;
    for (int q = 0; q < KRY_M * nv; ++q) Vt[q] = 0.0;
    for (int q = 0; q < KRY_M * KRY_M; ++q) { HHinv[q] = 0.0; inv[q] = 0.0; }
    for (int q = 0; q < nv * KRY_M; ++q) Vb[q] = 0.0;
    for (int q = 0; q < KRY_M; ++q) lambda[q] = 0.0;
    kry_rhs(C, w_sav, t_iter, rhs_sav);
    if (fls) for (int m = 0; m < nv; ++m) rhs_sav[m] = rhs_sav[m] - fls[m];
    double nrm = kry_norm2(nv, rhs_sav);
    for (int m = 1; m <= nv; ++m) VX(m, 1) = rhs_sav[m - 1] / nrm;
    for (int kk = 1; kk <= KRY_M; ++kk) {
        /* rosenbrock_krylov_jacobian (:10613-10662) */
        for (int m = 1; m <= nv; ++m) wl[m - 1] = w_sav[m - 1] + eps * VX(m, kk);
        kry_rhs(C, wl, t_iter, rhsp);
        if (fls) for (int m = 0; m < nv; ++m) rhsp[m] = rhsp[m] - fls[m];
        for (int m = 0; m < nv; ++m) jac[m] = (rhsp[m] - rhs_sav[m]) / eps;
        for (int jj = 1; jj <= kk; ++jj) {
            double d = 0.0;
            for (int m = 1; m <= nv; ++m) d = d + VX(m, jj) * jac[m - 1];
            HHX(jj, kk) = d;
            for (int m = 1; m <= nv; ++m) jac[m - 1] = jac[m - 1] - HHX(jj, kk) * VX(m, jj);
        }
        HHX(kk + 1, kk) = kry_norm2(nv, jac);
        for (int m = 1; m <= nv; ++m) VX(m, kk + 1) = jac[m - 1] / HHX(kk + 1, kk);
        int lim = kk < KRY_M ? kk : KRY_M;
        for (int ll = 1; ll <= lim; ++ll)
            for (int jj = 1; jj <= nv; ++jj) VTX(ll, jj) = VX(jj, ll);
        for (int ll = 1; ll <= KRY_M; ++ll) {
            rpr[ll - 1] = 0.0;
            for (int jj = 1; jj <= nv; ++jj) rpr[ll - 1] = rpr[ll - 1] + VTX(ll, jj) * rhs_sav[jj - 1];
        }
        for (int n = 1; n <= lim; ++n)
            for (int nn = 1; nn <= lim; ++nn) {
                KM(HHinv, n, nn) = -(dt * gam * HHX(n, nn));
                if (n == nn) KM(HHinv, n, nn) = KM(HHinv, n, nn) + 1.0;
            }
        if (kk < KRY_M)
            for (int n = kk + 1; n <= KRY_M; ++n) KM(HHinv, n, n) = 1.0;
        kry_inv(HHinv, kk);
        for (int ll = 1; ll <= KRY_M; ++ll) {
            lambda[ll - 1] = 0.0;
            for (int jj = 1; jj <= KRY_M; ++jj) lambda[ll - 1] = lambda[ll - 1] + KM(HHinv, ll, jj) * rpr[jj - 1];
        }
        if (fabs(dt * gam * HHX(kk + 1, kk)) * fabs(lambda[kk - 1]) < TOL) {
            for (int c = 1; c <= kk; ++c)
                for (int m = 1; m <= nv; ++m) Vb[(m - 1) + (size_t)nv * (c - 1)] = VX(m, c);
            for (int q = 0; q < KRY_M * KRY_M; ++q) inv[q] = HHinv[q];
            break;
        }
    }
#undef HHX
#undef VX
#undef VTX
}

/* One cell of rosenbrock_krylov_kernel (:10360-10611) from time_start to time_end:
 * 4-stage ROK method, a new Krylov basis for every attempted step, error
 * (yt - ys) / (1e-4 ys + 1e-8) in the 2-norm, dt *= min(5, max(0.2, 0.625 errn^0.1 /
 * errmax^0.175)).  fls is the cell's transport forcing fl_sav (simpler splitting) or NULL. */
int orc_rosenbrock_krylov_cell(const chem_t *C, double *vec, double t_iter, double dttry, double time_start,
                               double time_end, int maxsteps, const double *fls, int *nsteps)
{
    const int nv = C->nv, ns = C->ns;
    const double gam = 0.572816062482135;
    double alpha[5][4] = {{0}}, gamma_c[5][4] = {{0}}, bs[5], bt[5];
    alpha[2][1] = 0.432364435748567;
    alpha[3][1] = -0.514211316876170; alpha[3][2] = 1.382271144617360;
    alpha[4][1] = -0.514211316876170; alpha[4][2] = 1.382271144617360; alpha[4][3] = 0.0;
    bs[1] = 0.194335256262729; bs[2] = 0.483167813989227; bs[3] = 0.0; bs[4] = 0.322496929748044;
    gamma_c[2][1] = -0.602765307997356;
    gamma_c[3][1] = -1.389195789724843; gamma_c[3][2] = 1.072950969011413;
    gamma_c[4][1] = 0.992356412977094; gamma_c[4][2] = -1.390032613873701; gamma_c[4][3] = -0.440875890223325;
    bt[1] = -0.217819895945721; bt[2] = 1.03130847478467; bt[3] = 0.186511421161047; bt[4] = 0.0;
    double Vb[NVMAX * KRY_M], inv[KRY_M * KRY_M], rhs_sav[NVMAX];
    double A[NVMAX * KRY_M], Bm[NVMAX * NVMAX], Ks[NVMAX * (KRY_NSTEP + 1)];
    double ys[NVMAX], yt[NVMAX], F[NVMAX], GK[NVMAX], Cv[NVMAX], errv[NVMAX], yy[NSMAX];
#define VBX(r, c) Vb[((r) - 1) + (size_t)nv * ((c) - 1)]
#define AX(r, c) A[((r) - 1) + (size_t)nv * ((c) - 1)]
#define BX(r, c) Bm[((r) - 1) + (size_t)nv * ((c) - 1)]
#define KSX(r, c) Ks[((r) - 1) + (size_t)nv * ((c) - 1)]
    double time = time_start, dt = dttry, errn = 1.0;
    int it = 0;
    while (time < (time_end - 1.0e-12)) {
        dt = dt < time_end - time ? dt : time_end - time;
        it = it + 1;
        kry_arnoldi(C, vec, dt, t_iter, fls, Vb, inv, rhs_sav);
        /* A = Vb (I - inv), B = A Vb^T */
        for (int kk = 1; kk <= KRY_M; ++kk)
            for (int ll = 1; ll <= nv; ++ll) {
                double a = VBX(ll, 1) * ((kk == 1 ? 1.0 : 0.0) - KM(inv, 1, kk));
                for (int jj = 2; jj <= KRY_M; ++jj) a = a + VBX(ll, jj) * ((jj == kk ? 1.0 : 0.0) - KM(inv, jj, kk));
                AX(ll, kk) = a;
            }
        for (int kk = 1; kk <= nv; ++kk)
            for (int ll = 1; ll <= nv; ++ll) {
                double b = AX(ll, 1) * VBX(kk, 1);
                for (int jj = 2; jj <= KRY_M; ++jj) b = b + AX(ll, jj) * VBX(kk, jj);
                BX(ll, kk) = b;
            }
        for (int ii = 1; ii <= KRY_NSTEP; ++ii) {
            for (int m = 1; m <= nv; ++m) ys[m - 1] = vec[m - 1];
            for (int jj = 1; jj <= ii - 1; ++jj)
                for (int m = 1; m <= nv; ++m) ys[m - 1] = ys[m - 1] + dt * alpha[ii][jj] * KSX(m, jj);
            kry_rhs(C, ys, t_iter, F);
            for (int m = 0; m < nv; ++m) GK[m] = 0.0;
            for (int jj = 1; jj <= ii - 1; ++jj)
                for (int m = 1; m <= nv; ++m) GK[m - 1] = GK[m - 1] + gamma_c[ii][jj] / gam * KSX(m, jj);
            for (int m = 0; m < nv; ++m) F[m] = F[m] + GK[m];
            if (fls) for (int m = 0; m < nv; ++m) F[m] = F[m] - fls[m];
            for (int ll = 1; ll <= nv; ++ll) {
                double c = BX(ll, 1) * F[0];
                for (int jj = 2; jj <= nv; ++jj) c = c + BX(ll, jj) * F[jj - 1];
                Cv[ll - 1] = c;
            }
            for (int m = 1; m <= nv; ++m) KSX(m, ii) = F[m - 1] - Cv[m - 1] - GK[m - 1];
        }
        for (int ll = 1; ll <= nv; ++ll) {
            ys[ll - 1] = vec[ll - 1];
            yt[ll - 1] = vec[ll - 1];
            for (int ii = 1; ii <= KRY_NSTEP; ++ii) {
                ys[ll - 1] = ys[ll - 1] + dt * bs[ii] * KSX(ll, ii);
                yt[ll - 1] = yt[ll - 1] + dt * bt[ii] * KSX(ll, ii);
            }
            errv[ll - 1] = (yt[ll - 1] - ys[ll - 1]) / ((1.0e-4) * ys[ll - 1] + 1.0e-8);
        }
        double errmax = kry_norm2(nv, errv);
        double fac = 0.625 * pow(errn, 0.1) / pow(errmax, 0.175);
        fac = fac > 0.2 ? fac : 0.2;
        fac = fac < 5.0 ? fac : 5.0;
        double dtnext = dt * fac;
        if (errmax < 1.0) {
            time = time + dt;
            for (int m = 0; m < nv; ++m) vec[m] = ys[m];
            double rho = vec[0];
            for (int n = 2; n <= ns; ++n) rho = rho + vec[n - 1];
            double ri = 1.0 / rho;
            for (int n = 0; n < ns; ++n) yy[n] = vec[n] * ri;
            double uu = vec[I_U - 1] * ri, vv = vec[I_V - 1] * ri, ww = vec[I_W - 1] * ri;
            double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
            double ee = vec[I_E - 1] * ri - qq;
            t_iter = temperature_from_e(ns, ee, t_iter, C->cv_coeff, C->nsetcv, C->trange, C->indx_cp_l,
                                        C->indx_cp_r, C->tol_iter_nr, yy);
            dt = dtnext;
            errn = errmax;
        } else {
            dt = dtnext;
            errn = errmax;
        }
        if (it > maxsteps) break;
    }
    if (nsteps) *nsteps = it;
#undef VBX
#undef AX
#undef BX
#undef KSX
    return 0;
}

int orc_rosenbrock_krylov(int ns, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                          double *w, double *w_aux, const double *fl, const double *mw,
                          int nreactions, const double *arr_a, const double *arr_b,
                          const double *arr_ea, const double *tb_eff,
                          const double *falloff_coeffs, const int *reac_ty, const int *isrev,
                          const double *r_coeffs, const double *p_coeffs,
                          const double *kc_tab, int num_t_tab, double t_min_tab,
                          double dt_tab, double r_univ, double tol_iter_nr,
                          const double *cv_coeff, int nsetcv, const double *trange,
                          int indx_cp_l, int indx_cp_r, double dttry, double time_start,
                          double time_end, int simpler_splitting, const double *fl_sav, int maxsteps,
                          int enable_pasr, double les_c_yoshi, double les_c_mix, double les_c_eps,
                          const double *dcsidx, const double *detady, const double *dzitdz, void *stream)
{
    (void)fl; (void)stream;
    if (enable_pasr != 0 && nv_aux < J_LES1) return -2;
    if (simpler_splitting == 1 && fl_sav == NULL) return -1;
    chem_t C = {ns, nv, nreactions, num_t_tab, nsetcv, indx_cp_l, indx_cp_r, mw, arr_a, arr_b, arr_ea, tb_eff,
                falloff_coeffs, r_coeffs, p_coeffs, kc_tab, reac_ty, isrev, t_min_tab, dt_tab, r_univ,
                tol_iter_nr, cv_coeff, trange};
    const size_t fs = (size_t)nx * ny * nz;
    OMP_FOR
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        double zb = W(w_aux, i, j, k, J_Z);
        if (!(zb > 5.0e-4 && zb < 1.0 - 5.0e-4)) continue;
        double vec[NVMAX], fls[NVMAX];
        for (int m = 1; m <= nv; ++m) vec[m - 1] = W(w, i, j, k, m);
        if (simpler_splitting == 1)
            for (int m = 1; m <= nv; ++m)
                fls[m - 1] = fl_sav[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1)) + fs * (m - 1)];
        chem_t Cc = C;
        if (enable_pasr > 0) {
            Cc.pasr = 1;
            Cc.tau_k = pasr_tau_k(W(w_aux, i, j, k, J_R), W(w_aux, i, j, k, J_LES1), W(w_aux, i, j, k, J_MU),
                                  dcsidx[i - 1], detady[j - 1], dzitdz[k - 1], les_c_yoshi, les_c_mix, les_c_eps);
        }
        orc_rosenbrock_krylov_cell(&Cc, vec, W(w_aux, i, j, k, J_T), dttry, time_start, time_end, maxsteps,
                                   simpler_splitting == 1 ? fls : NULL, NULL);
        for (int m = 1; m <= nv; ++m) W(w, i, j, k, m) = vec[m - 1];
    }
    return 0;
}

/* Batched interface of the Krylov integrator over a flat list of reactors (state
 * w[nv][nbatch] SoA, temperature guess tg[nbatch], left as given), for the pinning tests. */
int orc_rosenbrock_krylov_batch(int ns, int nbatch, double *w, const double *tg, const double *mw, int nreactions,
                                const double *arr_a, const double *arr_b, const double *arr_ea,
                                const double *tb_eff, const double *falloff_coeffs, const int *reac_ty,
                                const int *isrev, const double *r_coeffs, const double *p_coeffs,
                                const double *kc_tab, int num_t_tab, double t_min_tab, double dt_tab,
                                double r_univ, double tol_iter_nr, const double *cv_coeff, int nsetcv,
                                const double *trange, int indx_cp_l, int indx_cp_r, const double *dttry,
                                const double *time_end, int maxsteps, int *nsteps)
{
    int nv = ns + 4;
    chem_t C = {ns, nv, nreactions, num_t_tab, nsetcv, indx_cp_l, indx_cp_r, mw, arr_a, arr_b, arr_ea, tb_eff,
                falloff_coeffs, r_coeffs, p_coeffs, kc_tab, reac_ty, isrev, t_min_tab, dt_tab, r_univ,
                tol_iter_nr, cv_coeff, trange};
    OMP_FOR
    for (int b = 0; b < nbatch; ++b) {
        double wv[NVMAX];
        for (int m = 0; m < nv; ++m) wv[m] = w[(size_t)m * nbatch + b];
        int it = 0;
        orc_rosenbrock_krylov_cell(&C, wv, tg[b], dttry[b], 0.0, time_end[b], maxsteps, NULL, &it);
        for (int m = 0; m < nv; ++m) w[(size_t)m * nbatch + b] = wv[m];
        if (nsteps) nsteps[b] = it;
    }
    return 0;
}

/* ===================================================================================== */
/* a10, a12: RK bookkeeping and time step                                                 */
/* ===================================================================================== */
int orc_init_flux(int nx, int ny, int nz, int nv, double *fl, double *fln, double rhodt, void *stream)
{
    (void)stream;
    long n = (long)nx * ny * nz * nv;
    OMP_FOR
    for (long q = 0; q < n; ++q) { fln[q] = -rhodt * fl[q]; fl[q] = 0.0; }
    return 0;
}

int orc_update_flux(int nx, int ny, int nz, int nv, const double *fl, double *fln, double gamdt, void *stream)
{
    (void)stream;
    long n = (long)nx * ny * nz * nv;
    OMP_FOR
    for (long q = 0; q < n; ++q) fln[q] = fln[q] - gamdt * fl[q];
    return 0;
}

/* init_flux_simpler_cuf, kernels_gpu.F90:47-66: fln = -rhodt*(fl - fl_sav); fl = 0 */
int orc_init_flux_simpler(int nx, int ny, int nz, int nv, double *fl, double *fln, const double *fl_sav,
                          double rhodt, void *stream)
{
    (void)stream;
    long n = (long)nx * ny * nz * nv;
    OMP_FOR
    for (long q = 0; q < n; ++q) { fln[q] = -rhodt * (fl[q] - fl_sav[q]); fl[q] = 0.0; }
    return 0;
}

/* sav_flx_cuf, kernels_gpu.F90:68-86: fl_sav = fl */
int orc_sav_flx(int nx, int ny, int nz, int nv, const double *fl, double *fl_sav, void *stream)
{
    (void)stream;
    long n = (long)nx * ny * nz * nv;
    OMP_FOR
    for (long q = 0; q < n; ++q) fl_sav[q] = fl[q];
    return 0;
}

/* update_simpler_flux_cuf, kernels_gpu.F90:3316-3334: fln -= gamdt*(fl - fl_sav) */
int orc_update_simpler_flux(int nx, int ny, int nz, int nv, const double *fl, double *fln, const double *fl_sav,
                            double gamdt, void *stream)
{
    (void)stream;
    long n = (long)nx * ny * nz * nv;
    OMP_FOR
    for (long q = 0; q < n; ++q) fln[q] = fln[q] - gamdt * (fl[q] - fl_sav[q]);
    return 0;
}

int orc_update_field(int nx, int ny, int nz, int ng, int nv, double *w, const double *fln,
                     const int *fluid_mask, void *stream)
{
    (void)stream;
    OMP_FOR
    for (int m = 1; m <= nv; ++m)
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
        if (!fluid_mask || fluid_mask[PIDX(i, j, k)] == 0)
            W(w, i, j, k, m) = W(w, i, j, k, m) + F(fln, i, j, k, m);
    return 0;
}

int orc_compute_dt(int ns, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                   const double *dcsidx, const double *detady, const double *dzitdz,
                   const double *dcsidxs, const double *detadys, const double *dzitdzs,
                   const double *w, const double *w_aux, double *dtmax9,
                   int indx_cp_l, int indx_cp_r, const double *cp_coeff, int nsetcv,
                   const double *trange, const int *fluid_mask, void *stream)
{
    (void)nv; (void)nv_aux; (void)w; (void)stream;
    double mx[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, yy[NSMAX];
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        if (fluid_mask && fluid_mask[PIDX(i, j, k)] != 0) continue;
        double rho = W(w_aux, i, j, k, J_R), ri = 1.0 / rho;
        double uu = W(w_aux, i, j, k, J_U), vv = W(w_aux, i, j, k, J_V), ww = W(w_aux, i, j, k, J_W);
        double tt = W(w_aux, i, j, k, J_T), mu = W(w_aux, i, j, k, J_MU), c = W(w_aux, i, j, k, J_C);
        double kc = W(w_aux, i, j, k, J_K_COND);
        for (int lsp = 1; lsp <= ns; ++lsp) yy[lsp - 1] = W(w_aux, i, j, k, lsp);
        double cploc = mixture_cp(ns, tt, indx_cp_l, indx_cp_r, cp_coeff, nsetcv, trange, yy);
        double nu = ri * mu, kr = kc * ri / cploc;
        double v[9] = {(fabs(uu) + c) * dcsidx[i - 1], (fabs(vv) + c) * detady[j - 1], (fabs(ww) + c) * dzitdz[k - 1],
                       nu * dcsidxs[i - 1], nu * detadys[j - 1], nu * dzitdzs[k - 1],
                       kr * dcsidxs[i - 1], kr * detadys[j - 1], kr * dzitdzs[k - 1]};
        for (int q = 0; q < 9; ++q) if (v[q] > mx[q]) mx[q] = v[q];
    }
    for (int q = 0; q < 9; ++q) dtmax9[q] = mx[q];
    return 0;
}

/* ===================================================================================== */
/* run-loop monitors                                                                       */
/* ===================================================================================== */
/* compute_residual_cuf (kernels_gpu.F90:5991-6012): sum over fluid points of
 * (fln(i,j,k,I_U)/dt)^2 */
int orc_compute_residual(int nx, int ny, int nz, int ng, int nv, const double *fln, double dt,
                         double *residual_rhou, const int *fluid_mask, void *stream)
{
    (void)stream;
    int ns = nv - 4;
    double res = 0.0;
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        if (fluid_mask && fluid_mask[PIDX(i, j, k)] != 0) continue;
        double r = F(fln, i, j, k, I_U) / dt;
        res = res + r * r;
    }
    *residual_rhou = res;
    return 0;
}

/* compute_rho_t_p_minmax_cuf (:6038-6073) */
int orc_compute_rho_t_p_minmax(int ns, int nx, int ny, int nz, int ng, int nv_aux, const double *w_aux,
                               double *rhomin, double *rhomax, double *tmin, double *tmax, double *pmin,
                               double *pmax, const int *fluid_mask, void *stream)
{
    (void)nv_aux; (void)stream;
    double v[6] = {1.7976931348623157e308, -100.0, 1.7976931348623157e308, -100.0, 1.7976931348623157e308, -100.0};
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        if (fluid_mask && fluid_mask[PIDX(i, j, k)] != 0) continue;
        double rho = W(w_aux, i, j, k, J_R), tt = W(w_aux, i, j, k, J_T), pp = W(w_aux, i, j, k, J_P);
        v[0] = fmin(v[0], rho); v[1] = fmax(v[1], rho);
        v[2] = fmin(v[2], tt);  v[3] = fmax(v[3], tt);
        v[4] = fmin(v[4], pp);  v[5] = fmax(v[5], pp);
    }
    *rhomin = v[0]; *rhomax = v[1]; *tmin = v[2]; *tmax = v[3]; *pmin = v[4]; *pmax = v[5];
    return 0;
}

/* find_boom_cuf (:10334-10358): cells whose (fln(i,j,k,2)/dt)^2 is NaN; the reference prints
 * each one's global indices, this returns their count and the first in i-fastest order */
int orc_find_boom(int nx, int ny, int nz, int ng, int nv, const double *fln, double dt, const int *fluid_mask,
                  int ncoords1, int ncoords3, int *nboom, int *first_ijk, void *stream)
{
    (void)nv; (void)stream;
    int cnt = 0;
    first_ijk[0] = first_ijk[1] = first_ijk[2] = 0;
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        if (fluid_mask && fluid_mask[PIDX(i, j, k)] != 0) continue;
        double r = F(fln, i, j, k, 2) / dt;
        double res = r * r;
        if (res != res) {
            if (cnt == 0) { first_ijk[0] = i + ncoords1 * nx; first_ijk[1] = j; first_ijk[2] = k + ncoords3 * nz; }
            ++cnt;
        }
    }
    *nboom = cnt;
    return 0;
}

/* ===================================================================================== */
/* host <-> device field copies (base_gpu.F90:1266-1325): on the CPU both sides are host   */
/* arrays, so these are the reference's transposition loops                               */
/* ===================================================================================== */
int orc_copy_cpu_gpu(int nx, int ny, int nz, int ng, int nv, const double *w_host, double *w_gpu, void *stream)
{
    (void)stream;
    for (int k = 1 - ng; k <= nz + ng; ++k)
    for (int j = 1 - ng; j <= ny + ng; ++j)
    for (int i = 1 - ng; i <= nx + ng; ++i)
    for (int iv = 1; iv <= nv; ++iv)
        W(w_gpu, i, j, k, iv) = w_host[(size_t)(iv - 1) + (size_t)nv * PIDX(i, j, k)];
    return 0;
}

int orc_copy_gpu_cpu(int nx, int ny, int nz, int ng, int nv, const double *w_gpu, double *w_host, void *stream)
{
    (void)stream;
    for (int k = 1 - ng; k <= nz + ng; ++k)
    for (int j = 1 - ng; j <= ny + ng; ++j)
    for (int i = 1 - ng; i <= nx + ng; ++i)
    for (int iv = 1; iv <= nv; ++iv)
        w_host[(size_t)(iv - 1) + (size_t)nv * PIDX(i, j, k)] = W(w_gpu, i, j, k, iv);
    return 0;
}

int orc_field_interior(int nx, int ny, int nz, int ng, int nv, double *w_gpu, double *dense, int to_dense,
                       void *stream)
{
    (void)stream;
    for (int m = 1; m <= nv; ++m)
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        if (to_dense) F(dense, i, j, k, m) = W(w_gpu, i, j, k, m);
        else W(w_gpu, i, j, k, m) = F(dense, i, j, k, m);
    }
    return 0;
}

/* ===================================================================================== */
/* non-reflecting / relaxation / reflective-wall boundary treatment                      */
/* ===================================================================================== */
/* bc_nr_lat_{x,y,z}_kernel (kernels_gpu.F90:4330-4793) for face (dir, start_or_end):
 * one-sided second-order derivative of the conserved variables normal to the face,
 * projected on the characteristic fields of the face cell (eigenvectors_{x,y,z}), the
 * incoming waves removed (nr_type 1), taken from the outer difference to the ghost cell
 * (2) or to the freestream winf (3), or reflected (6), and the result added to fl of the
 * face cell: fl += R (lambda L dw/dn) * metric.  The face cells are excluded from the
 * convective update (bc_preproc, multideal.F90:5671-5718). */
int orc_bc_nr_lat(int dir, int start_or_end, int nr_type, int ns, int nx, int ny, int nz, int ng, int nv,
                  const double *w_aux, const double *w, double *fl, const double *metric, int indx_cp_l,
                  int indx_cp_r, const double *rgas, const double *cv_coeff, const double *cp_coeff, int nsetcv,
                  const double *trange, const double *winf, void *stream)
{
    (void)cp_coeff; (void)stream;
    if (nv != ns + 4 || dir < 1 || dir > 3 || (start_or_end != 1 && start_or_end != 2)) return -2;
    if (nr_type != 1 && nr_type != 2 && nr_type != 3 && nr_type != 6) return -3;
    if ((dir == 1 ? nx : dir == 2 ? ny : nz) < 3 || (nr_type == 3 && !winf)) return -2;
    const double c_one[3] = {-1.5, 2.0, -0.5};
    const int sgn = start_or_end == 1 ? 1 : -1;
    int n1 = dir == 1 ? ny : nx, n2 = dir == 3 ? ny : nz;
    for (int b2 = 1; b2 <= n2; ++b2)
    for (int b1 = 1; b1 <= n1; ++b1) {
        int i, j, k;
        int nb = dir == 1 ? nx : dir == 2 ? ny : nz;
        int bnd = start_or_end == 1 ? 1 : nb;
        if (dir == 1) { i = bnd; j = b1; k = b2; }
        else if (dir == 2) { i = b1; j = bnd; k = b2; }
        else { i = b1; j = b2; k = bnd; }
        int di = dir == 1, dj = dir == 2, dk = dir == 3;
        double dw_dn[NVMAX], dw_dn_outer[NVMAX], dwc_dn[NVMAX], dwc_dn_outer[NVMAX], ev[NVMAX];
        double yyvec[NSMAX], prhoi[NSMAX], el[NVMAX * NVMAX], er[NVMAX * NVMAX];
        for (int m = 1; m <= nv; ++m) {
            dw_dn[m - 1] = 0.0;
            for (int l = 1; l <= 3; ++l)
                dw_dn[m - 1] = dw_dn[m - 1] + sgn * c_one[l - 1] *
                               W(w, i + di * sgn * (l - 1), j + dj * sgn * (l - 1), k + dk * sgn * (l - 1), m);
            double w_target = 0.0;
            if (nr_type == 2) w_target = W(w, i - di * sgn, j - dj * sgn, k - dk * sgn, m);
            else if (nr_type == 3) w_target = winf[m - 1];
            dw_dn_outer[m - 1] = sgn * (W(w, i, j, k, m) - w_target);
        }
        for (int lsp = 1; lsp <= ns; ++lsp) yyvec[lsp - 1] = W(w_aux, i, j, k, lsp);
        double rho = W(w_aux, i, j, k, J_R), uu = W(w_aux, i, j, k, J_U), vv = W(w_aux, i, j, k, J_V);
        double ww = W(w_aux, i, j, k, J_W), h = W(w_aux, i, j, k, J_H), tt = W(w_aux, i, j, k, J_T);
        double pp = W(w_aux, i, j, k, J_P), c = W(w_aux, i, j, k, J_C);
        double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
        double cc = c * c, rmix_tt = pp / rho, gamloc = cc / rmix_tt, gm1loc = gamloc - 1.0, ci = 1.0 / c;
        double p_e = rho * gm1loc, etot = h - rmix_tt;
        double b3 = etot, b2 = p_e / (rho * cc), b1v = -b2 * (etot - 2.0 * qq);
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double eei = species_h(ns, tt, indx_cp_l, indx_cp_r, cv_coeff, nsetcv, trange, lsp);
            prhoi[lsp - 1] = rgas[lsp - 1] * tt + gm1loc * (etot - qq - eei);
        }
        eigenvectors(dir, ns, b1v, b2, b3, uu, vv, ww, c, ci, h, yyvec, prhoi, el, er);
        for (int m = 1; m <= nv; ++m) {
            dwc_dn[m - 1] = 0.0;
            dwc_dn_outer[m - 1] = 0.0;
            for (int mm = 1; mm <= nv; ++mm) {
                dwc_dn[m - 1] = dwc_dn[m - 1] + el[(mm - 1) + (size_t)nv * (m - 1)] * dw_dn[mm - 1];
                dwc_dn_outer[m - 1] = dwc_dn_outer[m - 1] + el[(mm - 1) + (size_t)nv * (m - 1)] * dw_dn_outer[mm - 1];
            }
        }
        double un = dir == 1 ? uu : dir == 2 ? vv : ww;
        ev[0] = un - c;
        for (int lsp = 2; lsp <= ns + 3; ++lsp) ev[lsp - 1] = un;
        ev[nv - 1] = un + c;
        if (nr_type == 1)
            for (int l = 1; l <= nv; ++l) ev[l - 1] = sgn * fmin(sgn * ev[l - 1], 0.0);
        if (nr_type == 2 || nr_type == 3)
            for (int m = 1; m <= nv; ++m)
                if (sgn * ev[m - 1] > 0.0) dwc_dn[m - 1] = dwc_dn_outer[m - 1];
        for (int m = 1; m <= nv; ++m) dwc_dn[m - 1] = ev[m - 1] * dwc_dn[m - 1];
        if (nr_type == 6) {
            for (int lsp = 2; lsp <= ns + 3; ++lsp) dwc_dn[lsp - 1] = 0.0;
            if (start_or_end == 1) dwc_dn[nv - 1] = dwc_dn[0];
            else dwc_dn[0] = dwc_dn[nv - 1];
        }
        double met = metric[(dir == 1 ? i : dir == 2 ? j : k) - 1];
        for (int m = 1; m <= nv; ++m) {
            double df = 0.0;
            for (int mm = 1; mm <= nv; ++mm) df = df + er[(mm - 1) + (size_t)nv * (m - 1)] * dwc_dn[mm - 1];
            F(fl, i, j, k, m) = F(fl, i, j, k, m) + df * met;
        }
    }
    return 0;
}

/* ===================================================================================== */
/* f4: LES auxiliaries (compute_aux_les, multideal_gpu.F90:445-498)                        */
/* ===================================================================================== */
/* eval_velaux_cuf (kernels_gpu.F90:9879-9917): over the padded box, species clipped at 0
 * in w, Y, rho, u, v, w into w_aux and the SGS viscosity slot cleared. */
int orc_eval_velaux(int ns, int nx, int ny, int nz, int ng, int nv, int nv_aux, double *w, double *w_aux,
                    void *stream)
{
    (void)stream;
    if (nv != ns + 4 || nv_aux < J_LES1) return -2;
    OMP_FOR
    for (int k = 1 - ng; k <= nz + ng; ++k)
    for (int j = 1 - ng; j <= ny + ng; ++j)
    for (int i = 1 - ng; i <= nx + ng; ++i) {
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double v = W(w, i, j, k, lsp);
            W(w, i, j, k, lsp) = v > 0.0 ? v : 0.0;
        }
        double rho = W(w, i, j, k, 1);
        for (int lsp = 2; lsp <= ns; ++lsp) rho = rho + W(w, i, j, k, lsp);
        double ri = 1.0 / rho;
        for (int lsp = 1; lsp <= ns; ++lsp) W(w_aux, i, j, k, lsp) = W(w, i, j, k, lsp) * ri;
        W(w_aux, i, j, k, J_R) = rho;
        W(w_aux, i, j, k, J_U) = W(w, i, j, k, I_U) * ri;
        W(w_aux, i, j, k, J_V) = W(w, i, j, k, I_V) * ri;
        W(w_aux, i, j, k, J_W) = W(w, i, j, k, I_W) * ri;
        W(w_aux, i, j, k, J_LES1) = 0.0;
    }
    return 0;
}

/* les_wale_mut_cuf (kernels_gpu.F90:9921-10047): on the interior, the velocity gradient by
 * central differences of order ep_order, the Ducros sensor and div/3 (as eval_aux2 does),
 * and the WALE eddy viscosity
 *   mu_sgs = rho C_w^2 Delta^2 |S^d|^3 / (|S|^5 + |S^d|^(5/2) + tiny),  Delta^2 = (dxi/dx deta/dy dzeta/dz)^(-2/3),
 * S the strain rate, S^d the traceless symmetric part of the squared velocity gradient. */
int orc_les_wale_mut(int ns, int nx, int ny, int nz, int ng, int nv_aux, int ep_order, double *w_aux,
                     const double *coeff_deriv1, const double *dcsidx, const double *detady, const double *dzitdz,
                     double les_c_wale, double eps_sensor, int sensor_type, void *stream)
{
    (void)stream;
    int lmax = ep_order / 2;
    if (nv_aux < J_LES1) return -2;
    if (lmax < 1 || lmax > 4 || lmax > ng) return -3;
    OMP_FOR
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        double delta2 = dcsidx[i - 1] * detady[j - 1] * dzitdz[k - 1];
        delta2 = pow(delta2, -2.0 / 3.0);
        double dudx = 0, dudy = 0, dudz = 0, dvdx = 0, dvdy = 0, dvdz = 0, dwdx = 0, dwdy = 0, dwdz = 0;
        for (int l = 1; l <= lmax; ++l) {
            double ccl = CD1(l, lmax);
            dudx = dudx + ccl * (W(w_aux, i + l, j, k, J_U) - W(w_aux, i - l, j, k, J_U));
            dudy = dudy + ccl * (W(w_aux, i, j + l, k, J_U) - W(w_aux, i, j - l, k, J_U));
            dudz = dudz + ccl * (W(w_aux, i, j, k + l, J_U) - W(w_aux, i, j, k - l, J_U));
            dvdx = dvdx + ccl * (W(w_aux, i + l, j, k, J_V) - W(w_aux, i - l, j, k, J_V));
            dvdy = dvdy + ccl * (W(w_aux, i, j + l, k, J_V) - W(w_aux, i, j - l, k, J_V));
            dvdz = dvdz + ccl * (W(w_aux, i, j, k + l, J_V) - W(w_aux, i, j, k - l, J_V));
            dwdx = dwdx + ccl * (W(w_aux, i + l, j, k, J_W) - W(w_aux, i - l, j, k, J_W));
            dwdy = dwdy + ccl * (W(w_aux, i, j + l, k, J_W) - W(w_aux, i, j - l, k, J_W));
            dwdz = dwdz + ccl * (W(w_aux, i, j, k + l, J_W) - W(w_aux, i, j, k - l, J_W));
        }
        dudx *= dcsidx[i - 1]; dudy *= detady[j - 1]; dudz *= dzitdz[k - 1];
        dvdx *= dcsidx[i - 1]; dvdy *= detady[j - 1]; dvdz *= dzitdz[k - 1];
        dwdx *= dcsidx[i - 1]; dwdy *= detady[j - 1]; dwdz *= dzitdz[k - 1];
        double div = dudx + dvdy + dwdz;
        double div3l = div / 3.0;
        double omegax = dwdy - dvdz, omegay = dudz - dwdx, omegaz = dvdx - dudy;
        double omod2 = omegax * omegax + omegay * omegay + omegaz * omegaz;
        double s = -div / sqrt(omod2 + div * div + eps_sensor);
        if (sensor_type != 0 && s < 0.0) s = 0.0;
        W(w_aux, i, j, k, J_DUC) = s * s;
        double S_xx = dudx, S_xy = 0.5 * (dudy + dvdx), S_xz = 0.5 * (dudz + dwdx);
        double S_yy = dvdy, S_yz = 0.5 * (dvdz + dwdy), S_zz = dwdz;
        double g_xx = dudx * dudx + dudy * dvdx + dudz * dwdx;
        double g_xy = dudx * dudy + dudy * dvdy + dudz * dwdy;
        double g_xz = dudx * dudz + dudy * dvdz + dudz * dwdz;
        double g_yx = dvdx * dudx + dvdy * dvdx + dvdz * dwdx;
        double g_yy = dvdx * dudy + dvdy * dvdy + dvdz * dwdy;
        double g_yz = dvdx * dudz + dvdy * dvdz + dvdz * dwdz;
        double g_zx = dwdx * dudx + dwdy * dvdx + dwdz * dwdx;
        double g_zy = dwdx * dudy + dwdy * dvdy + dwdz * dwdy;
        double g_zz = dwdx * dudz + dwdy * dvdz + dwdz * dwdz;
        double tce = (g_xx + g_yy + g_zz) / 3.0;
        double Sd_xx = g_xx - tce, Sd_xy = 0.5 * (g_xy + g_yx), Sd_xz = 0.5 * (g_xz + g_zx);
        double Sd_yy = g_yy - tce, Sd_yz = 0.5 * (g_yz + g_zy), Sd_zz = g_zz - tce;
        double normS = S_xx * S_xx + S_yy * S_yy + S_zz * S_zz + 2.0 * (S_xy * S_xy + S_xz * S_xz + S_yz * S_yz);
        normS = sqrt(normS);
        double normSd = Sd_xx * Sd_xx + Sd_yy * Sd_yy + Sd_zz * Sd_zz +
                        2.0 * (Sd_xy * Sd_xy + Sd_xz * Sd_xz + Sd_yz * Sd_yz);
        normSd = sqrt(normSd);
        double rho = W(w_aux, i, j, k, J_R);
        double nS2 = normS * normS, nSd2 = normSd * normSd;
        double mu_sgs = rho * les_c_wale * les_c_wale * delta2 * (nSd2 * normSd) /
                        (nS2 * nS2 * normS + sqrt(nSd2 * nSd2 * normSd) + DBL_MIN);
        W(w_aux, i, j, k, J_LES1) = mu_sgs;
        W(w_aux, i, j, k, J_DIV) = div3l;
    }
    return 0;
}

/* eval_aux_les_cuf (kernels_gpu.F90:10049-10161): over the padded box, from the rho, u,
 * v, w and Y that eval_velaux left in w_aux: T by Newton, h, p, c and the mixture
 * transport of eval_aux with the SGS terms added, mu + mu_sgs, k + mu_sgs c_p / Pr_sgs,
 * rho D_s + mu_sgs / Sc_sgs.  The mixture fraction is not touched. */
int orc_eval_aux_les(int ns, int nx, int ny, int nz, int ng, int nv, int nv_aux, double *w, double *w_aux,
                     double p0, double t_min_tab, double dt_tab, double r_univ, const double *rgas,
                     const double *cv_coeff, const double *cp_coeff, int nsetcv, const double *trange,
                     int indx_cp_l, int indx_cp_r, double tol_iter_nr, const double *mw, const double *mwinv,
                     const double *visc_species, const double *lambda_species, const double *diffbin_species,
                     int num_t_tab, double les_pr, double les_sc, void *stream)
{
    (void)stream;
    if (nv != ns + 4 || nv_aux < J_LES1) return -2;
    OMP_FOR
    for (int k = 1 - ng; k <= nz + ng; ++k)
    for (int j = 1 - ng; j <= ny + ng; ++j)
    for (int i = 1 - ng; i <= nx + ng; ++i) {
        double yy[NSMAX];
        for (int lsp = 1; lsp <= ns; ++lsp) yy[lsp - 1] = W(w_aux, i, j, k, lsp);
        double rho = W(w_aux, i, j, k, J_R);
        double ri = 1.0 / rho;
        double uu = W(w_aux, i, j, k, J_U), vv = W(w_aux, i, j, k, J_V), ww = W(w_aux, i, j, k, J_W);
        double mu_sgs = W(w_aux, i, j, k, J_LES1);
        double qq = 0.5 * (uu * uu + vv * vv + ww * ww);
        double rhoe = W(w, i, j, k, I_E);
        double ee = rhoe * ri - qq;
        double tt = temperature_from_e(ns, ee, W(w_aux, i, j, k, J_T), cv_coeff, nsetcv, trange,
                                       indx_cp_l, indx_cp_r, tol_iter_nr, yy);
        double rmixtloc = rgas[0] * yy[0];
        for (int lsp = 2; lsp <= ns; ++lsp) rmixtloc = rmixtloc + rgas[lsp - 1] * yy[lsp - 1];
        double cploc = mixture_cp(ns, tt, indx_cp_l, indx_cp_r, cp_coeff, nsetcv, trange, yy);
        double gamloc = cploc / (cploc - rmixtloc);
        double pp = rho * tt * rmixtloc;
        double c = sqrt(gamloc * pp * ri);
        W(w_aux, i, j, k, J_H) = (rhoe + pp) * ri;
        W(w_aux, i, j, k, J_T) = tt;
        W(w_aux, i, j, k, J_P) = pp;
        W(w_aux, i, j, k, J_C) = c;
        int itt = (int)((tt - t_min_tab) / dt_tab) + 1;
        if (itt < 1) itt = 1;
        if (itt > num_t_tab) itt = num_t_tab;
        double tloc = t_min_tab + (itt - 1) * dt_tab;
        double dtt = (tt - tloc) / dt_tab;
        double mwmixt = r_univ / rmixtloc;
        double k_cond = 0.0, k_cond2 = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double dlam = TAB2(lambda_species, itt + 1, lsp) - TAB2(lambda_species, itt, lsp);
            double klsp = TAB2(lambda_species, itt, lsp) + dlam * dtt;
            double xlsp = yy[lsp - 1] * mwmixt * mwinv[lsp - 1];
            k_cond = k_cond + xlsp * klsp;
            k_cond2 = k_cond2 + xlsp / klsp;
        }
        k_cond = 0.5 * (k_cond + 1.0 / k_cond2);
        double mu = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double dmu = TAB2(visc_species, itt + 1, lsp) - TAB2(visc_species, itt, lsp);
            double mulsp = TAB2(visc_species, itt, lsp) + dmu * dtt;
            double mu_den = 0.0;
            for (int msp = 1; msp <= ns; ++msp) {
                dmu = TAB2(visc_species, itt + 1, msp) - TAB2(visc_species, itt, msp);
                double mumsp = TAB2(visc_species, itt, msp) + dmu * dtt;
                double t1 = 1.0 + sqrt(mumsp / mulsp) * pow(mw[msp - 1] * mwinv[lsp - 1], 0.25);
                double phi_lm = 1.0 / sqrt(1.0 + mw[lsp - 1] * mwinv[msp - 1]) * (t1 * t1);
                mu_den = mu_den + yy[msp - 1] * phi_lm * mwinv[msp - 1];
            }
            mu = mu + mulsp * yy[lsp - 1] * mwinv[lsp - 1] / mu_den;
        }
        mu = sqrt(8.0) * mu;
        W(w_aux, i, j, k, J_MU) = mu + mu_sgs;
        W(w_aux, i, j, k, J_K_COND) = k_cond + mu_sgs * cploc / les_pr;
        for (int lsp = 1; lsp <= ns; ++lsp) {
            double diff_den = 0.0;
            for (int msp = 1; msp <= ns; ++msp) {
                if (msp != lsp) {
                    double ddiff = TAB3(diffbin_species, itt + 1, lsp, msp) - TAB3(diffbin_species, itt, lsp, msp);
                    double diff_ij = TAB3(diffbin_species, itt, lsp, msp) + ddiff * dtt;
                    double xmsp = yy[msp - 1] * mwmixt * mwinv[msp - 1];
                    diff_den = diff_den + xmsp / diff_ij;
                }
            }
            double rd = diff_den > 1.0e-15 ? (1.0 - yy[lsp - 1]) / diff_den * p0 / pp * rho : 0.0;
            W(w_aux, i, j, k, J_D_START + lsp) = rd + mu_sgs / les_sc;
        }
    }
    return 0;
}

/* eval_chem_aux_cuf (kernels_gpu.F90:9676-9877), the output-time chemistry diagnostics on
 * the interior: the net production rates wdot_s = sum_r (p - r) q_r W_s at the state of
 * w_aux (rho, T, Y), scaled by the PaSR factor when enable_pasr (which is then stored in
 * J_GAM_PASR), and the heat release rate -sum_s h298_s wdot_s / W_s into J_HRR (the same
 * slot: with PaSR the factor is overwritten by the heat release rate, as there). */
int orc_eval_chem_aux(int ns, int nx, int ny, int nz, int nv_aux, int ng, int nreactions, double *w_aux,
                      const double *mw, const double *arr_a, const double *arr_b, const double *arr_ea,
                      const double *tb_eff, const double *falloff_coeffs, const int *reac_ty, const int *isrev,
                      const double *r_coeffs, const double *p_coeffs, const double *kc_tab, int num_t_tab,
                      double t_min_tab, double dt_tab, double r_univ, const double *dcsidx, const double *detady,
                      const double *dzitdz, const double *h298, int enable_pasr, double les_c_yoshi,
                      double les_c_mix, double les_c_eps, void *stream)
{
    (void)stream;
    if (nv_aux < J_HRR) return -2;
    OMP_FOR
    for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
        double conc[NSMAX], yy[NSMAX], wdot[NSMAX];
        double rho = W(w_aux, i, j, k, J_R);
        double ri = 1.0 / rho;
        double tt = W(w_aux, i, j, k, J_T);
        double Rtti = 1.0 / (r_univ * tt);
        for (int lsp = 1; lsp <= ns; ++lsp) {
            wdot[lsp - 1] = 0.0;
            yy[lsp - 1] = W(w_aux, i, j, k, lsp);
            conc[lsp - 1] = rho * yy[lsp - 1] / mw[lsp - 1];
        }
        double tau_k = 0.0;
        if (enable_pasr > 0)
            tau_k = pasr_tau_k(rho, W(w_aux, i, j, k, J_LES1), W(w_aux, i, j, k, J_MU), dcsidx[i - 1],
                               detady[j - 1], dzitdz[k - 1], les_c_yoshi, les_c_mix, les_c_eps);
        int itt = (int)((tt - t_min_tab) / dt_tab) + 1;
        if (itt < 1) itt = 1;
        if (itt > num_t_tab) itt = num_t_tab;
        double ttleft = (itt - 1) * dt_tab + t_min_tab;
        double dtt = (tt - ttleft) / dt_tab;
        for (int lr = 1; lr <= nreactions; ++lr) {
            double qlr = progress_rate(nreactions, lr, ns, tt, Rtti, itt, dtt, num_t_tab, conc, arr_a, arr_b, arr_ea,
                                       tb_eff, falloff_coeffs, reac_ty, isrev, r_coeffs, p_coeffs, kc_tab);
            for (int lsp = 1; lsp <= ns; ++lsp)
                wdot[lsp - 1] = wdot[lsp - 1] + (RX(p_coeffs, lr, lsp) - RX(r_coeffs, lr, lsp)) * qlr * mw[lsp - 1];
        }
        if (enable_pasr > 0) {
            double tau_cf = 0.0, tau_cs = DBL_MAX;
            for (int lsp = 0; lsp < ns; ++lsp) {
                if (fabs(wdot[lsp] * ri) > 1.0e-8) {
                    double tau_lsp = yy[lsp] / fabs(wdot[lsp] * ri);
                    if (tau_lsp > tau_cf) tau_cf = tau_lsp;
                    if (tau_lsp > 0.0 && tau_lsp < tau_cs) tau_cs = tau_lsp;
                }
            }
            double tau_c = sqrt(tau_cs * tau_cf);
            double gam_pasr = tau_k > DBL_MIN ? tau_c / (tau_c + tau_k) : 1.0;
            for (int lsp = 0; lsp < ns; ++lsp) wdot[lsp] = gam_pasr * wdot[lsp];
            W(w_aux, i, j, k, J_GAM_PASR) = gam_pasr;
        }
        double hrr = 0.0;
        for (int lsp = 1; lsp <= ns; ++lsp) {
            W(w_aux, i, j, k, J_WDOT_START + lsp) = wdot[lsp - 1];
            hrr = hrr - h298[lsp - 1] * wdot[lsp - 1] / mw[lsp - 1];
        }
        W(w_aux, i, j, k, J_HRR) = hrr;
    }
    return 0;
}
