/*
 * streams_hip.h — C-ABI of libstreams_hip.so, the gfx950 (MI355X) implementation of the
 * STREAmS-2 multideal reacting RHS and stiff chemistry.
 *
 * Every entry point replaces one routine of the reference's CUDA-Fortran kernel module
 * `streams_kernels_gpu` (src/multideal/kernels_gpu.F90) or of the halo layer
 * (src/base_gpu.F90), and keeps its argument order.  Conventions:
 *   - arrays are device pointers with the reference's Fortran layout (column-major,
 *     first index fastest, lower bound 1-ng on padded fields: w(1-ng:nx+ng,1-ng:ny+ng,
 *     1-ng:nz+ng,1:nv)); tables keep their Fortran shapes (e.g. cp_coeff(indx_cp_l:
 *     indx_cp_r+2,N_S,nsetcv));
 *   - N_S is a compile-time macro in the reference (src/multideal/index_define.h:3);
 *     here it is the explicit first argument `n_s`;
 *   - arrays that are assumed-shape in the reference carry their extent explicitly
 *     (nv_aux);
 *   - `stream` is a hipStream_t (NULL = default stream); work is enqueued, not waited
 *     for, except where an output is returned to the host;
 *   - the return value is 0 on success, a positive hipError_t on a HIP failure, or a
 *     negative STREAMS_E_* code for an invalid argument.
 * Indices of conserved / auxiliary variables follow index_define.h:
 *   w:     1..N_S = rho*Y_s, N_S+1..N_S+4 = rho*u, rho*v, rho*w, rho*E
 *   w_aux: 1..N_S = Y_s, N_S+1.. = u, v, w, h, T, p, c, mu, Ducros, div/3, rho, Z, k,
 *          N_S+13+s = rho*D_s.
 */
#ifndef STREAMS_HIP_H
#define STREAMS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define STREAMS_E_NSPECIES   (-1)  /* N_S outside the compiled set            */
#define STREAMS_E_DIMS       (-2)  /* inconsistent sizes or index ranges      */
#define STREAMS_E_OPTION     (-3)  /* unsupported numerical option            */
#define STREAMS_E_NOMEM      (-4)  /* device allocation for scratch failed    */

/* Version / capability query: returns the number of compiled N_S values and writes up to
 * `cap` of them into `list`. */
int streams_species_supported(int *list, int cap);

/* ---- auxiliary variables (a1, a2) ------------------------------------------------- */

/* eval_aux_cuf, kernels_gpu.F90:6154-6308.  Clips rho*Y_s >= 0 in w (side effect), writes
 * Y, u, v, w, h, T (Newton on e from the previous w_aux T), p, c, mu (Wilke), k (Mathur),
 * rho*D_s (Bird), Z over the box [istart:iend]x[jstart:jend]x[kstart:kend]. */
int streams_eval_aux(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                     int istart, int iend, int jstart, int jend, int kstart, int kend,
                     double *w, double *w_aux, double p0, double t_min_tab, double dt_tab,
                     double r_univ, const double *rgas, const double *cv_coeff,
                     const double *cp_coeff, int nsetcv, const double *trange,
                     int indx_cp_l, int indx_cp_r, double tol_iter_nr, void *stream,
                     const double *mw, const double *mwinv, const double *visc_species,
                     const double *lambda_species, const double *diffbin_species,
                     int num_t_tab, int n_eoi, const double *aw_eoi, const int *nainsp,
                     const double *beta0, const double *coeff_eoi);

/* eval_aux2_cuf, kernels_gpu.F90:6310-6382: Ducros sensor and div/3 on the interior.
 * sensor_type < 0 (no counterpart in the reference) computes div/3 only and leaves the
 * Ducros column untouched, for substeps whose sensor threshold is at least 1. */
int streams_eval_aux2(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order,
                      double *w_aux, const double *coeff_deriv1, const double *dcsidx,
                      const double *detady, const double *dzitdz, double eps_sensor,
                      int sensor_type, void *stream);

/* ---- LES auxiliaries (compute_aux_les, multideal_gpu.F90:445-498) ------------------ */

/* eval_velaux_cuf, kernels_gpu.F90:9879-9917: over the padded box, species clipped at 0
 * in w, Y, rho, u, v, w into w_aux and the eddy-viscosity slot J_LES1 = 14 + 2 N_S cleared.
 * nv_aux >= 2 n_s + 14. */
int streams_eval_velaux(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                        double *w, double *w_aux, void *stream);

/* les_wale_mut_cuf, kernels_gpu.F90:9921-10047: on the interior, the velocity gradient by
 * central differences of order ep_order (lmax = ep_order / 2 <= min(4, ng)), the Ducros
 * sensor and div/3 as eval_aux2 computes them, and the WALE eddy viscosity
 * rho C_w^2 Delta^2 |S^d|^3 / (|S|^5 + |S^d|^(5/2)) into J_LES1. */
int streams_les_wale_mut(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int ep_order,
                         double *w_aux, const double *coeff_deriv1, const double *dcsidx,
                         const double *detady, const double *dzitdz, double les_c_wale,
                         double eps_sensor, int sensor_type, void *stream);

/* eval_aux_les_cuf, kernels_gpu.F90:10049-10161: over the padded box, from the rho, u, v,
 * w and Y in w_aux: T, h, p, c and the mixture transport of eval_aux with the eddy
 * viscosity added (mu + mu_sgs, k + mu_sgs c_p / les_pr, rho D_s + mu_sgs / les_sc); the
 * mixture fraction J_Z is left as it is. */
int streams_eval_aux_les(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                         double *w, double *w_aux, double p0, double t_min_tab, double dt_tab,
                         double r_univ, const double *rgas, const double *cv_coeff,
                         const double *cp_coeff, int nsetcv, const double *trange,
                         int indx_cp_l, int indx_cp_r, double tol_iter_nr, const double *mw,
                         const double *mwinv, const double *visc_species,
                         const double *lambda_species, const double *diffbin_species,
                         int num_t_tab, double les_pr, double les_sc, void *stream);

/* eval_chem_aux_cuf, kernels_gpu.F90:9676-9877 (compute_chem_aux, multideal_gpu.F90:500-534,
 * at output times): on the interior, the production rates sum_r (p - r) q_r W_s at the
 * rho, T, Y of w_aux into J_WDOT_START + s (scaled by the PaSR factor when enable_pasr)
 * and the heat release rate -sum_s h298_s wdot_s / W_s into J_HRR = 15 + 3 N_S (the slot
 * the reference also gives the PaSR factor, which the heat release rate overwrites).
 * nv_aux >= 3 n_s + 15. */
int streams_eval_chem_aux(int n_s, int nx, int ny, int nz, int nv_aux, int ng, int nreactions,
                          double *w_aux, const double *mw, const double *arr_a,
                          const double *arr_b, const double *arr_ea, const double *tb_eff,
                          const double *falloff_coeffs, const int *reac_ty, const int *isrev,
                          const double *r_coeffs, const double *p_coeffs, const double *kc_tab,
                          int num_t_tab, double t_min_tab, double dt_tab, double r_univ,
                          const double *dcsidx, const double *detady, const double *dzitdz,
                          const double *h298, int enable_pasr, double les_c_yoshi,
                          double les_c_mix, double les_c_eps, void *stream);

/* ---- ghost cells and halo exchange (a3, a11) --------------------------------------- */

/* bcextr_var_cuf, kernels_gpu.F90:5147-5207: zero-order extrapolation of one variable
 * (w_var points at its first element) on all six faces. */
int streams_bcextr_var(int nx, int ny, int nz, int ng, double *w_var, void *stream);

/* base_gpu%bcswap / bcswap_var, base_gpu.F90:1132-1264, for one rank: a face whose
 * neighbour is the rank itself (periodic axis) receives the loopback copy of the
 * opposite ng-deep interior slab; a face with no neighbour is left alone.  All slabs are
 * read before any ghost is written, as pack -> copy -> unpack does.  periodic[3] gives
 * the x, y, z flags.  nv = number of consecutive variables starting at w. */
int streams_bcswap(int nx, int ny, int nz, int ng, int nv, double *w,
                   const int *periodic, void *stream);

/* bcextr_cuf, kernels_gpu.F90:5209-5288 (face ilat = 1..6). */
int streams_bcextr(int ilat, int nx, int ny, int nz, int ng, int nv, double *w, void *stream);

/* bcwall_ad_cuf, kernels_gpu.F90:5749-5852: reflecting adiabatic wall on face ilat. */
int streams_bcwall_ad(int n_s, int ilat, int nx, int ny, int nz, int ng, double *w,
                      void *stream);

/* bc_nr_lat_{x,y,z}_kernel, kernels_gpu.F90:4330-4793 (launched by bc_nr,
 * multideal_gpu.F90:1116-1163) on face (dir = 1..3, start_or_end = 1 min / 2 max):
 * characteristic boundary treatment of the face cells, nr_type 1 non-reflecting, 2
 * relaxation to the ghost cell, 3 relaxation to the freestream state winf(1:nv), 6
 * reflective wall: fl += R (lambda L dw/dn) * metric(face cell).  The face cells must be
 * excluded from the convective update (eul_imin = 2 etc., bc_preproc, multideal.F90:5671-
 * 5718).  metric is dcsidx / detady / dzitdz; winf may be NULL unless nr_type = 3. */
int streams_bc_nr_lat(int dir, int start_or_end, int nr_type, int n_s, int nx, int ny, int nz,
                      int ng, int nv, const double *w_aux, const double *w, double *fl,
                      const double *metric, int indx_cp_l, int indx_cp_r, const double *rgas,
                      const double *cv_coeff, const double *cp_coeff, int nsetcv,
                      const double *trange, const double *winf, void *stream);

/* ---- convective flux (a4, a5) ------------------------------------------------------ */

/* euler_x (multideal_gpu.F90:948-1002): euler_x_fluxes_hybrid_kernel
 * (kernels_gpu.F90:88-425; flux_splitting=1: the Rusanov variant :427-730) on faces
 * istart_face..iend_face, then euler_x_update_cuf (:732-750) fl += (f_i - f_{i-1}) dcsidx
 * on cells eul_imin..eul_imax.  fhat is scratch: its content afterwards is unspecified.
 * With sensor_threshold < 1 the shocked-face worklist is stream-ordered scratch of 4 bytes
 * per face (STREAMS_E_NOMEM if it cannot be allocated); the same holds for euler_y/z and
 * rhs_sweep. */
int streams_euler_x(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                    int istart_face, int iend_face, int lmax_base, int nkeep,
                    const double *rgas, const double *w_aux, const double *coeff_deriv1,
                    const double *dcsidx, double *fhat, int force_zero_flux_min,
                    int force_zero_flux_max, int weno_scheme, int weno_version,
                    double sensor_threshold, int weno_size, const double *cv_coeff,
                    const double *cp_coeff, int nsetcv, const double *trange,
                    int indx_cp_l, int indx_cp_r, const int *ep_ord_change,
                    double tol_iter_nr, double rho0, double u0, int flux_splitting,
                    int eul_imin, int eul_imax, double *fl, void *stream);

/* euler_y (multideal_gpu.F90:1004-1045): euler_y_hybrid_kernel (kernels_gpu.F90:1427-1759)
 * on faces eul_jmin-1..eul_jmax and euler_y_update_cuf (:752-771). */
int streams_euler_y(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                    int eul_jmin, int eul_jmax, int lmax_base, int nkeep,
                    const double *rgas, const double *w_aux, double *fl,
                    const double *coeff_deriv1, const double *detady, double *fhat,
                    int force_zero_flux_min, int force_zero_flux_max, int weno_scheme,
                    int weno_version, double sensor_threshold, int weno_size,
                    const double *cv_coeff, const double *cp_coeff, int nsetcv,
                    const double *trange, int indx_cp_l, int indx_cp_r,
                    const int *ep_ord_change, double tol_iter_nr, double rho0, double u0,
                    int flux_splitting, void *stream);

/* euler_z (multideal_gpu.F90:1047-1090): euler_z_hybrid_kernel (kernels_gpu.F90:795-1123)
 * on faces eul_kmin-1..eul_kmax and euler_z_update_cuf (:773-793). */
int streams_euler_z(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                    int eul_kmin, int eul_kmax, int lmax_base, int nkeep,
                    const double *rgas, const double *w_aux, double *fl,
                    const double *coeff_deriv1, const double *dzitdz, double *fhat,
                    int force_zero_flux_min, int force_zero_flux_max, int weno_scheme,
                    int weno_version, double sensor_threshold, int weno_size,
                    const double *cv_coeff, const double *cp_coeff, int nsetcv,
                    const double *trange, int indx_cp_l, int indx_cp_r,
                    const int *ep_ord_change, double tol_iter_nr, double rho0, double u0,
                    int flux_splitting, void *stream);

/* ---- viscous and diffusive fluxes (a6) --------------------------------------------- */

/* visflx_x_cuf / _y / _z, kernels_gpu.F90:3658-4031 (conservative_viscous=1, mode 1):
 * staggered-face stress, conduction, Hirschfelder-Curtiss diffusion with correction
 * velocity and enthalpy diffusion, then the face-difference update of fl.  `coord` is
 * x (1-ng:nx+ng+1), y or z (1-ng:n+ng).  fhat is scratch. */
int streams_visflx_x(int n_s, int nx, int ny, int nz, int nv, int nv_aux, int ng,
                     const double *x, const double *w_aux, double *fl, double *fhat,
                     int indx_cp_l, int indx_cp_r, const double *cp_coeff, int nsetcv,
                     const double *trange, const double *rgas, double r_univ, void *stream);
int streams_visflx_y(int n_s, int nx, int ny, int nz, int nv, int nv_aux, int ng,
                     const double *y, const double *w_aux, double *fl, double *fhat,
                     int indx_cp_l, int indx_cp_r, const double *cp_coeff, int nsetcv,
                     const double *trange, const double *rgas, double r_univ, void *stream);
int streams_visflx_z(int n_s, int nx, int ny, int nz, int nv, int nv_aux, int ng,
                     const double *z, const double *w_aux, double *fl, double *fhat,
                     int indx_cp_l, int indx_cp_r, const double *cp_coeff, int nsetcv,
                     const double *trange, const double *rgas, double r_univ, void *stream);

/* visflx_reduced_cuf, kernels_gpu.F90:3539-3656 (mode 2): non-Laplacian stress terms and
 * viscous work. */
int streams_visflx_reduced(int n_s, int nx, int ny, int nz, int ng, int nv_aux,
                           int visc_order, const double *w_aux, double *fl,
                           const double *coeff_deriv1, const double *dcsidx,
                           const double *detady, const double *dzitdz, void *stream);

/* visflx_cuf, kernels_gpu.F90:3357-3537 (conservative_viscous = 0, mode 0): the Laplacian
 * form of the viscous terms from cell-centred first and second derivatives (coeff_deriv2(0:4,4),
 * second-derivative metrics dcsidxs/dcsidx2 etc.), with conduction as (grad mu . grad T +
 * mu lap T) k/mu; momentum and energy only (no species diffusion in this form). */
int streams_visflx(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order,
                   double *w_aux, double *fl, const double *coeff_deriv1,
                   const double *coeff_deriv2, const double *dcsidx, const double *detady,
                   const double *dzitdz, const double *dcsidxs, const double *detadys,
                   const double *dzitdzs, const double *dcsidx2, const double *detady2,
                   const double *dzitdz2, void *stream);

/* ---- chemistry (a7, a8) ------------------------------------------------------------ */

/* compute_chemistry_cuf, kernels_gpu.F90:9503-9674: explicit finite-rate source,
 * fl(1:N_S) -= omega_s W_s; clips w, rewrites w_aux Y and T. */
int streams_compute_chemistry(int n_s, int nx, int ny, int nz, int nv, int nv_aux, int ng,
                              int nreactions, double *w, double *w_aux, double *fl,
                              const double *mw, const double *arr_a, const double *arr_b,
                              const double *arr_ea, const double *tb_eff,
                              const double *falloff_coeffs, const int *reac_ty,
                              const int *isrev, const double *r_coeffs,
                              const double *p_coeffs, const double *kc_tab, int num_t_tab,
                              double t_min_tab, double dt_tab, double r_univ, int indx_cp_l,
                              int indx_cp_r, const double *cv_coeff, int nsetcv,
                              const double *trange, double tol_iter_nr, void *stream);

/* rosenbrock_kernel, kernels_gpu.F90:8928-9501 (launched by rosenbrock,
 * multideal_gpu.F90:694-767): per cell adaptive 4-stage Rosenbrock from time_start to
 * time_end with a central-difference Jacobian and partially pivoted LU, all FP64.
 * enable_pasr = 1 scales the species sources of every right-hand side by the PaSR factor
 * tau_c / (tau_c + tau_k), tau_k from w_aux's rho, mu and eddy viscosity (J_LES1, so
 * nv_aux >= 2 n_s + 14) and the metrics dcsidx, detady, dzitdz (kernels_gpu.F90:8970-8980,
 * 9359-9379). */
int streams_rosenbrock(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng,
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
                       const double *detady, const double *dzitdz, void *stream);

/* rosenbrock_krylov_kernel, kernels_gpu.F90:10360-11205 (launched by rosenbrock_krylov,
 * multideal_gpu.F90:769-838): per cell 4-stage Rosenbrock-Krylov (ROK) from time_start to
 * time_end, a Krylov space of at most M_max = 4 vectors rebuilt by Arnoldi on one-sided
 * finite-difference Jacobian-vector products for every attempted step, error norm
 * ||(yt - ys) / (1e-4 ys + 1e-8)||_2, only for cells with 5e-4 < Z < 1 - 5e-4.
 * simpler_splitting = 1 subtracts the transport forcing fl_sav(1:nx,1:ny,1:nz,1:nv) from
 * every right-hand side.  w_aux is read only (T guess, Z).  enable_pasr = 1 applies the
 * PaSR factor to every right-hand side as in streams_rosenbrock (:10475-10486,
 * 10954-10974). */
int streams_rosenbrock_krylov(int n_s, int nv, int nv_aux, int nx, int ny, int nz, int ng,
                              double *w, double *w_aux, const double *fl, const double *mw,
                              int nreactions, const double *arr_a, const double *arr_b,
                              const double *arr_ea, const double *tb_eff,
                              const double *falloff_coeffs, const int *reac_ty,
                              const int *isrev, const double *r_coeffs, const double *p_coeffs,
                              const double *kc_tab, int num_t_tab, double t_min_tab,
                              double dt_tab, double r_univ, double tol_iter_nr,
                              const double *cv_coeff, int nsetcv, const double *trange,
                              int indx_cp_l, int indx_cp_r, double dttry, double time_start,
                              double time_end, int simpler_splitting, const double *fl_sav,
                              int maxsteps, int enable_pasr, double les_c_yoshi, double les_c_mix,
                              double les_c_eps, const double *dcsidx, const double *detady,
                              const double *dzitdz, void *stream);

/* ---- Runge-Kutta bookkeeping and time step (a10, a12) ------------------------------ */

/* init_flux_cuf (kernels_gpu.F90:28-45): fln = -rhodt*fl; fl = 0. */
int streams_init_flux(int nx, int ny, int nz, int nv, double *fl, double *fln,
                      double rhodt, void *stream);
/* update_flux_cuf (:3297-3314): fln -= gamdt*fl. */
int streams_update_flux(int nx, int ny, int nz, int nv, const double *fl, double *fln,
                        double gamdt, void *stream);
/* Simpler operator splitting (operator_splitting = 1, rk_sync(simpler_splitting=1),
 * multideal_gpu.F90:536-621 and advance_solution :1875-1893): the RK stages accumulate
 * fl - fl_sav, where fl_sav holds the transport right-hand side saved by sav_flx
 * (:623-692) and fed to rosenbrock_krylov as its forcing.
 * init_flux_simpler_cuf (kernels_gpu.F90:47-66): fln = -rhodt*(fl - fl_sav); fl = 0. */
int streams_init_flux_simpler(int nx, int ny, int nz, int nv, double *fl, double *fln,
                              const double *fl_sav, double rhodt, void *stream);
/* sav_flx_cuf (:68-86): fl_sav = fl. */
int streams_sav_flx(int nx, int ny, int nz, int nv, const double *fl, double *fl_sav, void *stream);
/* update_simpler_flux_cuf (:3316-3334): fln -= gamdt*(fl - fl_sav). */
int streams_update_simpler_flux(int nx, int ny, int nz, int nv, const double *fl, double *fln,
                                const double *fl_sav, double gamdt, void *stream);
/* update_field_cuf (:3336-3355): w += fln where fluid_mask == 0 (fluid_mask may be NULL
 * for an all-fluid domain). */
int streams_update_field(int nx, int ny, int nz, int ng, int nv, double *w,
                         const double *fln, const int *fluid_mask, void *stream);

/* compute_dt_cuf (kernels_gpu.F90:6075-6152): the nine inverse-time-step maxima
 * (convective x,y,z; viscous x,y,z; conductive x,y,z) written to host `dtmax9`; the call
 * waits for the stream.  fluid_mask may be NULL. */
int streams_compute_dt(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                       const double *dcsidx, const double *detady, const double *dzitdz,
                       const double *dcsidxs, const double *detadys, const double *dzitdzs,
                       const double *w, const double *w_aux, double *dtmax9,
                       int indx_cp_l, int indx_cp_r, const double *cp_coeff, int nsetcv,
                       const double *trange, const int *fluid_mask, void *stream);

/* No counterpart in the reference: streams_compute_dt for callers that cannot allocate
 * device memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results; the partial maxima live in a fixed device array
 * of the library (147,528 bytes per GPU, 9 doubles for each of at most 2048 blocks plus the
 * 9 results, held until the process ends) instead of stream-ordered scratch.  Calls that
 * overlap on one GPU share that array and race.  Any box size is accepted (the grid is
 * capped at 2048 blocks); nv_aux below 2 n_s + 13 returns STREAMS_E_DIMS. */
int streams_compute_dt_noalloc(int n_s, int nx, int ny, int nz, int ng, int nv, int nv_aux,
                               const double *dcsidx, const double *detady, const double *dzitdz,
                               const double *dcsidxs, const double *detadys, const double *dzitdzs,
                               const double *w, const double *w_aux, double *dtmax9,
                               int indx_cp_l, int indx_cp_r, const double *cp_coeff, int nsetcv,
                               const double *trange, const int *fluid_mask, void *stream);


/* ---- host <-> device field copies and checkpoints (base_gpu.F90:1266-1325) ---------- */

/* copy_cpu_gpu (base_gpu.F90:1266-1286): the host field w_host(nv, 1-ng:nx+ng, 1-ng:ny+ng,
 * 1-ng:nz+ng) (variable fastest, host memory) into the device field w_gpu(1-ng:..,..,..,
 * 1:nv); the transposition runs on the device.  nv <= 32.  Waits for the stream. */
int streams_copy_cpu_gpu(int nx, int ny, int nz, int ng, int nv, const double *w_host,
                         double *w_gpu, void *stream);

/* copy_gpu_cpu (:1303-1325): the reverse, into host memory.  Waits for the stream. */
int streams_copy_gpu_cpu(int nx, int ny, int nz, int ng, int nv, const double *w_gpu,
                         double *w_host, void *stream);

/* The interior w_gpu(1:nx,1:ny,1:nz,1:nv) of a padded device field to a dense device array
 * dense(nx,ny,nz,nv) (to_dense = 1; the layout of rst.bin, field.F90:427-477) or back
 * (to_dense = 0).  No counterpart as one routine: the reference writes rst.bin from the host
 * copy through an MPI-IO subarray view. */
int streams_field_interior(int nx, int ny, int nz, int ng, int nv, double *w_gpu, double *dense,
                           int to_dense, void *stream);

/* ---- run-loop monitors (the time loop of run, multideal_gpu.F90:1659-1851) ----------- */
/* Each writes its results to host memory and waits for the stream.  The reductions combine
 * partial results in a fixed order, so a result does not depend on the launch. */

/* compute_residual_cuf (kernels_gpu.F90:5991-6012): residual_rhou (host) = sum over fluid
 * points of (fln(i,j,k,I_U)/dt)^2.  fluid_mask may be NULL. */
int streams_compute_residual(int nx, int ny, int nz, int ng, int nv, const double *fln, double dt,
                             double *residual_rhou, const int *fluid_mask, void *stream);

/* compute_rho_t_p_minmax_cuf (:6038-6073): extrema of w_aux(J_R), w_aux(J_T), w_aux(J_P)
 * over fluid points, written to the six host scalars (minima start at huge(), maxima at
 * -100, as there). */
int streams_compute_rho_t_p_minmax(int n_s, int nx, int ny, int nz, int ng, int nv_aux,
                                   const double *w_aux, double *rhomin, double *rhomax,
                                   double *tmin, double *tmax, double *pmin, double *pmax,
                                   const int *fluid_mask, void *stream);

/* find_boom_cuf (:10334-10358): the fluid cells where (fln(i,j,k,2)/dt)^2 is NaN.  The
 * reference prints each one's global indices (i + ncoords1*nx, j, k + ncoords3*nz); here
 * *nboom (host) receives their number and first_ijk[3] (host) those indices of the first
 * one in i-fastest order (zeros when there is none). */
int streams_find_boom(int nx, int ny, int nz, int ng, int nv, const double *fln, double dt,
                      const int *fluid_mask, int ncoords1, int ncoords3, int *nboom,
                      int *first_ijk, void *stream);

/* ---- fused substep stages (MI355X-first orchestration of rk_sync) ------------------- */
/* These have no single counterpart in the reference: each folds consecutive routines of
 * one RK substep (multideal_gpu.F90:562-618) into one pass so that face fluxes and the
 * RK bookkeeping stay on chip.  Results agree with the routine-by-routine sequence to
 * rounding (the order of the fl accumulation differs). */

/* Convective (euler_{x,y,z}_fluxes_hybrid + update) and, with visc=1, viscous
 * (visflx_{x,y,z}_cuf) fluxes of direction dir (1..3) on faces 0..n with one update of
 * fl for cells 1..n; init=1 also performs init_flux_cuf (fln = -rhodt*fl_old); init=2
 * sets fl to this direction's part without reading it or touching fln, for a substep whose
 * fln streams_rhs_finish_next already set.  metric is dcsidx/detady/dzitdz, coord is x/y/z
 * of that direction. */
int streams_rhs_sweep(int dir, int n_s, int nv_aux, int nx, int ny, int nz, int ng, int lmax_base,
                      int nkeep, const double *rgas, const double *w_aux, const double *coeff_deriv1,
                      const double *metric, const double *coord, double *fhat,
                      int force_zero_flux_min, int force_zero_flux_max, int weno_scheme,
                      int weno_version, double sensor_threshold, const double *cv_coeff,
                      const double *cp_coeff, int nsetcv, const double *trange, int indx_cp_l,
                      int indx_cp_r, const int *ep_ord_change, double tol_iter_nr, double rho0,
                      double u0, int flux_splitting, int visc, double r_univ, int init,
                      double rhodt, double *fl, double *fln, void *stream);

/* Kinetics records for the fused chemistry: packs the reference's kinetics tables (HOST
 * arrays in their Fortran layouts, as read by multideal.F90:2763-2926: arr_a/b/ea(nr,2),
 * tb_eff/r_coeffs/p_coeffs(nr,N_S), falloff_coeffs(nr,5), reac_ty/isRev(nr)) and the molar
 * masses into nreactions records of 80 doubles each, the layout documented in
 * csrc/kinpack.h.  Returns the number of doubles of the pack (pass out = NULL to query),
 * STREAMS_E_DIMS if cap is too small, STREAMS_E_OPTION if a reaction has more than 4
 * reactant or product species, STREAMS_E_NSPECIES for N_S > 16.  The caller copies the pack to the device once, as the reference uploads its
 * tables once (multideal_gpu.F90:1363-1424). */
int streams_kinetics_pack(int n_s, int nreactions, const double *arr_a, const double *arr_b,
                          const double *arr_ea, const double *tb_eff, const double *falloff_coeffs,
                          const int *reac_ty, const int *isrev, const double *r_coeffs,
                          const double *p_coeffs, const double *mw, double *out, int cap);

/* visflx_reduced_cuf (reduced=1), compute_chemistry_cuf (chem=1, with rho, Y, T taken from
 * w_aux; kinetics from the device copy of streams_kinetics_pack's records), update_flux_cuf
 * and update_field_cuf in one pass over the interior. */
int streams_rhs_finish(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order,
                       double *w, const double *w_aux, double *fl, double *fln,
                       const double *coeff_deriv1, const double *dcsidx, const double *detady,
                       const double *dzitdz, int reduced, int chem, int nreactions,
                       const double *mw, const double *kin_pack, const double *kc_tab,
                       int num_t_tab, double t_min_tab, double dt_tab, double r_univ,
                       double gamdt, const int *fluid_mask, void *stream);

/* streams_rhs_finish followed by the next substep's init_flux_cuf (:28-45), for the
 * substeps of one rk_sync but its last: fln receives -rhodt_next * fl (fl the substep's
 * final right-hand side) in place of the update increment, and fl is not stored.  The next
 * substep's first streams_rhs_sweep then runs with init=2.  w is updated as by
 * streams_rhs_finish, and the pair gives bit-identical results to streams_rhs_finish and a
 * sweep with init=1. */
int streams_rhs_finish_next(int n_s, int nx, int ny, int nz, int ng, int nv_aux, int visc_order,
                            double *w, const double *w_aux, double *fl, double *fln,
                            const double *coeff_deriv1, const double *dcsidx, const double *detady,
                            const double *dzitdz, int reduced, int chem, int nreactions,
                            const double *mw, const double *kin_pack, const double *kc_tab,
                            int num_t_tab, double t_min_tab, double dt_tab, double r_univ,
                            double gamdt, double rhodt_next, const int *fluid_mask, void *stream);

/* ---- batched chemistry ---------------------------------------------------------------- */
/* nbatch independent reactors, each integrated like one cell of rosenbrock_kernel
 * (kernels_gpu.F90:8928-9193, launched by rosenbrock, multideal_gpu.F90:694-767) from t = 0
 * to time_end[b] with first trial step dttry[b].  State is SoA on the device,
 * w[m*nbatch + b] for m = 0..N_S+3:
 *   pressure == NULL: constant volume, w = the conserved variables of w(i,j,k,1:nv);
 *   pressure != NULL: constant pressure pressure[b] (adiabatic), w = (Y_1..Y_N_S, 0, 0, 0,
 *     h) with h the mixture enthalpy; T from h by Newton on cp_coeff, rho = p/(R T)
 *     (rgas and cp_coeff are then required).  The reference's reactor is constant volume.
 * tguess[b] is the Newton start in and the last temperature out; nacc/nrej (may be NULL)
 * receive the accepted / rejected step counts.  t_done/dt_next (both NULL, or both given)
 * make the call resumable with a step budget: reactor b starts at t_done[b] with the step
 * dt_next[b] (dttry[b] if not positive), takes at most maxsteps steps and returns the time
 * reached (time_end[b] when done) and the step it would take next.  Kinetics come from the device copy of
 * streams_kinetics_pack.  work is one unsigned of device memory for the shared work
 * counter (zeroed by the call on the stream); wg_per_cu workgroups of 256 lanes per CU form
 * the persistent grid (4 when wg_per_cu <= 0; never more than the batch needs).
 * mixed = 0 is the plain FP64 integrator in the reference's arithmetic forms (central-
 * difference Jacobian over all N_S + 4 variables, FP64 LU, rosenbrock_rhs's forms of the
 * rate expressions and of the temperature inversion): the yardstick of the other forms.
 * mixed = 1 solves on the species block (the momentum / energy rows of a reactor's RHS
 * are zero), with a one-sided species Jacobian factored in FP32 and every stage solve
 * refined once against the FP64 residual (same stages, error norm and step control);
 * mixed = 2 is that species-block integrator with an FP64 LU.  No
 * counterpart as one routine in the reference: its rosenbrock_kernel takes the cells of a
 * field, this takes a flat batch. */
int streams_reactor_batch(int n_s, int nbatch, double *w, double *tguess, const double *dttry,
                          const double *time_end, const double *pressure, const double *mw,
                          const double *rgas, const double *cp_coeff, int nreactions,
                          const double *kin_pack, const double *kc_tab, int num_t_tab,
                          double t_min_tab, double dt_tab, double r_univ, double tol_iter_nr,
                          const double *cv_coeff, int nsetcv, const double *trange, int indx_cp_l,
                          int indx_cp_r, int maxsteps, int maxtry, double tol, int *nacc, int *nrej,
                          double *t_done, double *dt_next, unsigned *work, int wg_per_cu, int mixed,
                          void *stream);

#ifdef __cplusplus
}
#endif
#endif /* STREAMS_HIP_H */
