! dropin_check.F90 -- calls the gfx950 library through the Fortran drop-in module exactly as
! the reference's host code calls its kernel module: device arrays allocated with hipMalloc
! and passed as c_ptr handles, scalars by value.
!
! Without arguments: checks the RK bookkeeping (init_flux_cuf, update_flux_cuf,
! update_field_cuf) and the periodic halo exchange of the face layers (bcswap) against the
! same arithmetic on the host, and prints "dropin_check: ok".
!
! With two arguments (case file, output file): reads a state and the tables from the case
! file, runs eval_aux_cuf, eval_aux2_cuf, euler_x, visflx_x_cuf, compute_chemistry_cuf and
! rosenbrock_krylov through the module with the reference's argument lists, and writes
! w_aux, fhat, fl and w to the output file, for a comparison with the same sequence made
! through the C-ABI from another host language.  Case file (stream access, little endian):
! 16 int32 (nx ny nz ng n_s nv nv_aux nsetcv indx_cp_l indx_cp_r num_t_tab nreactions lmax
! nkeep weno_scheme maxsteps), 16 float64 (p0 t_min_tab dt_tab R_univ tol_iter_nr rho0 u0
! sensor_threshold eps_sensor dt_chem, then unused), then NBLK arrays, each an int32 kind
! (1 float64, 2 int32), an int64 count and the values, in the order of the block indices
! below.
#ifndef N_S
#define N_S 9
#endif
program dropin_check
  use, intrinsic :: iso_c_binding
  use streams_kernels_hip
  implicit none
  interface
    integer(c_int) function hipMalloc(p, n) bind(C, name="hipMalloc")
      import :: c_ptr, c_size_t, c_int
      type(c_ptr) :: p
      integer(c_size_t), value :: n
    end function hipMalloc
    integer(c_int) function hipMemcpy(dst, src, n, kind) bind(C, name="hipMemcpy")
      import :: c_ptr, c_size_t, c_int
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: n
      integer(c_int), value :: kind
    end function hipMemcpy
    integer(c_int) function hipFree(p) bind(C, name="hipFree")
      import :: c_ptr, c_int
      type(c_ptr), value :: p
    end function hipFree
    integer(c_int) function hipDeviceSynchronize() bind(C, name="hipDeviceSynchronize")
      import :: c_int
    end function hipDeviceSynchronize
  end interface
  integer, parameter :: nx = 12, ny = 10, nz = 8, ng = 3, nv = N_S + 4
  integer(c_int), parameter :: h2d = 1, d2h = 2
  real(c_double), parameter :: rhodt = -0.7d0, gamdt = 0.45d0
  real(c_double), allocatable, target :: w(:,:,:,:), fl(:,:,:,:), fln(:,:,:,:)
  real(c_double), allocatable, target :: w_d(:,:,:,:), fl_d(:,:,:,:), fln_d(:,:,:,:)
  real(c_double), allocatable, target :: fl0(:,:,:,:)
  real(c_double), allocatable :: fln_ref(:,:,:,:), w_ref(:,:,:,:)
  integer(c_int), target :: periodic(3), list(64)
  type(c_ptr) :: dw, dfl, dfln
  integer(c_size_t) :: bw, bf
  integer :: i, j, k, m, nsp, ii, jj, kk, nout
  real(c_double) :: err, expect

  nsp = streams_species_supported(c_loc(list), 64_c_int)
  if (.not. any(list(1:nsp) == N_S)) then
    write(*, '(a,i0,a)') 'dropin_check: N_S = ', N_S, ' not compiled in the library'
    error stop 1
  end if
  if (command_argument_count() >= 2) then
    call run_case()
    stop
  end if

  allocate(w(1-ng:nx+ng, 1-ng:ny+ng, 1-ng:nz+ng, nv), fl(nx, ny, nz, nv), fln(nx, ny, nz, nv))
  allocate(w_d, mold=w)
  allocate(fl_d, mold=fl)
  allocate(fln_d, mold=fln)
  call random_number(w)
  call random_number(fl)
  call random_number(fln)
  fl0 = fl
  bw = int(size(w), c_size_t) * 8_c_size_t
  bf = int(size(fl), c_size_t) * 8_c_size_t
  call ok(hipMalloc(dw, bw), 'hipMalloc')
  call ok(hipMalloc(dfl, bf), 'hipMalloc')
  call ok(hipMalloc(dfln, bf), 'hipMalloc')
  call ok(hipMemcpy(dw, c_loc(w), bw, h2d), 'hipMemcpy')
  call ok(hipMemcpy(dfl, c_loc(fl), bf, h2d), 'hipMemcpy')
  call ok(hipMemcpy(dfln, c_loc(fln), bf, h2d), 'hipMemcpy')

  ! init_flux: fln = -rhodt*fl, fl = 0; then fl = fl0 again on the host side of the check
  call init_flux_cuf(nx, ny, nz, nv, dfl, dfln, rhodt, c_null_ptr)
  call ok(hipMemcpy(dfl, c_loc(fl0), bf, h2d), 'hipMemcpy')
  call update_flux_cuf(nx, ny, nz, nv, dfl, dfln, gamdt, c_null_ptr)
  call update_field_cuf(nx, ny, nz, ng, nv, dw, dfln, c_null_ptr, c_null_ptr)
  periodic = 1
  call bcswap(nx, ny, nz, ng, nv, dw, c_loc(periodic), c_null_ptr)
  call ok(hipDeviceSynchronize(), 'hipDeviceSynchronize')
  call ok(hipMemcpy(c_loc(fln_d), dfln, bf, d2h), 'hipMemcpy')
  call ok(hipMemcpy(c_loc(w_d), dw, bw, d2h), 'hipMemcpy')

  fln_ref = -rhodt * fl0
  fln_ref = fln_ref - gamdt * fl0
  w_ref = w
  w_ref(1:nx, 1:ny, 1:nz, :) = w(1:nx, 1:ny, 1:nz, :) + fln_ref
  err = maxval(abs(fln_d - fln_ref))
  if (err > 1d-15) then
    write(*, '(a,es12.4)') 'dropin_check: update_flux mismatch ', err
    error stop 1
  end if
  do m = 1, nv
    do k = 1 - ng, nz + ng
      kk = modulo(k - 1, nz) + 1
      do j = 1 - ng, ny + ng
        jj = modulo(j - 1, ny) + 1
        do i = 1 - ng, nx + ng
          ii = modulo(i - 1, nx) + 1
          ! bcswap fills the six face layers (base_gpu.F90:155-207); edge and corner ghosts
          ! keep their values
          nout = merge(1, 0, i < 1 .or. i > nx) + merge(1, 0, j < 1 .or. j > ny) + merge(1, 0, k < 1 .or. k > nz)
          if (nout <= 1) then
            expect = w_ref(ii, jj, kk, m)
          else
            expect = w(i, j, k, m)
          end if
          if (abs(w_d(i, j, k, m) - expect) > 1d-15) then
            write(*, '(a,4i4)') 'dropin_check: field/halo mismatch at ', i, j, k, m
            error stop 1
          end if
        end do
      end do
    end do
  end do
  call ok(hipFree(dw), 'hipFree')
  call ok(hipFree(dfl), 'hipFree')
  call ok(hipFree(dfln), 'hipFree')
  write(*, '(a,i0,a)') 'dropin_check: ok (N_S = ', N_S, ')'
contains
  subroutine run_case()
    integer, parameter :: NBLK = 32
    integer, parameter :: B_W = 1, B_WAUX = 2, B_RGAS = 3, B_CV = 4, B_CP = 5, B_TRANGE = 6, B_MW = 7, &
        B_MWINV = 8, B_VISC = 9, B_LAMBDA = 10, B_DIFF = 11, B_EOI = 12, B_NAINSP = 13, B_BETA0 = 14, &
        B_CD1 = 15, B_DCSIDX = 16, B_DETADY = 17, B_DZITDZ = 18, B_X = 19, B_EPO = 20, B_ARRA = 21, &
        B_ARRB = 22, B_ARREA = 23, B_TBEFF = 24, B_FALL = 25, B_RTY = 26, B_REV = 27, B_RCO = 28, &
        B_PCO = 29, B_KC = 30, B_FL = 31, B_FHAT = 32
    character(len=4096) :: path_in, path_out
    integer(c_int32_t), target :: hi(16)
    real(c_double), target :: hd(16)
    type(c_ptr) :: p(NBLK)
    integer(c_int64_t) :: cnt(NBLK)
    integer(c_int32_t) :: kind(NBLK)
    integer :: u, v, b
    integer(c_int) :: cnx, cny, cnz, cng, cnv, cnv_aux, nsetcv, lo, hi_, num_t_tab, nr, lmax, nkeep, wsch
    integer(c_int) :: maxsteps
    real(c_double) :: p0, t_min_tab, dt_tab, r_univ, tol, rho0, u0, thr, eps_sensor, dt_chem

    call get_command_argument(1, path_in)
    call get_command_argument(2, path_out)
    open(newunit=u, file=trim(path_in), access='stream', form='unformatted', status='old')
    read(u) hi, hd
    cnx = hi(1); cny = hi(2); cnz = hi(3); cng = hi(4); cnv = hi(6); cnv_aux = hi(7)
    nsetcv = hi(8); lo = hi(9); hi_ = hi(10); num_t_tab = hi(11); nr = hi(12); lmax = hi(13)
    nkeep = hi(14); wsch = hi(15); maxsteps = hi(16)
    if (hi(5) /= N_S) then
      write(*, '(a,i0,a,i0)') 'dropin_check: case has N_S = ', hi(5), ', module built for ', N_S
      error stop 1
    end if
    p0 = hd(1); t_min_tab = hd(2); dt_tab = hd(3); r_univ = hd(4); tol = hd(5); rho0 = hd(6)
    u0 = hd(7); thr = hd(8); eps_sensor = hd(9); dt_chem = hd(10)
    do b = 1, NBLK
      read(u) kind(b), cnt(b)
      call upload(u, kind(b), cnt(b), p(b))
    end do
    close(u)

    open(newunit=v, file=trim(path_out), access='stream', form='unformatted', status='replace')
    call eval_aux_cuf(cnx, cny, cnz, cng, cnv, cnv_aux, 1 - cng, cnx + cng, 1 - cng, cny + cng, 1 - cng, &
        cnz + cng, p(B_W), p(B_WAUX), p0, t_min_tab, dt_tab, r_univ, p(B_RGAS), p(B_CV), p(B_CP), nsetcv, &
        p(B_TRANGE), lo, hi_, tol, c_null_ptr, p(B_MW), p(B_MWINV), p(B_VISC), p(B_LAMBDA), p(B_DIFF), &
        num_t_tab, 1_c_int, p(B_EOI), p(B_NAINSP), p(B_BETA0), p(B_EOI))
    call eval_aux2_cuf(cnx, cny, cnz, cng, cnv_aux, 2 * lmax, p(B_WAUX), p(B_CD1), p(B_DCSIDX), p(B_DETADY), &
        p(B_DZITDZ), eps_sensor, 1_c_int, c_null_ptr)
    call download(v, p(B_WAUX), cnt(B_WAUX))
    call euler_x(cnv, cnv_aux, cnx, cny, cnz, cng, 0_c_int, cnx, lmax, nkeep, p(B_RGAS), p(B_WAUX), p(B_CD1), &
        p(B_DCSIDX), p(B_FHAT), 0_c_int, 0_c_int, wsch, 1_c_int, thr, 2 * wsch, p(B_CV), p(B_CP), nsetcv, &
        p(B_TRANGE), lo, hi_, p(B_EPO), tol, rho0, u0, 0_c_int, 1_c_int, cnx, p(B_FL), c_null_ptr)
    call visflx_x_cuf(cnx, cny, cnz, cnv, cnv_aux, cng, p(B_X), p(B_WAUX), p(B_FL), p(B_FHAT), lo, hi_, p(B_CP), &
        nsetcv, p(B_TRANGE), p(B_RGAS), r_univ, c_null_ptr)
    call download(v, p(B_FHAT), cnt(B_FHAT))
    call compute_chemistry_cuf(cnx, cny, cnz, cnv, cnv_aux, cng, nr, p(B_W), p(B_WAUX), p(B_FL), p(B_MW), &
        p(B_ARRA), p(B_ARRB), p(B_ARREA), p(B_TBEFF), p(B_FALL), p(B_RTY), p(B_REV), p(B_RCO), p(B_PCO), &
        p(B_KC), num_t_tab, t_min_tab, dt_tab, r_univ, lo, hi_, p(B_CV), nsetcv, p(B_TRANGE), tol, c_null_ptr)
    call download(v, p(B_FL), cnt(B_FL))
    call rosenbrock_krylov(cnv, cnv_aux, cnx, cny, cnz, cng, p(B_W), p(B_WAUX), p(B_FL), p(B_MW), nr, &
        p(B_ARRA), p(B_ARRB), p(B_ARREA), p(B_TBEFF), p(B_FALL), p(B_RTY), p(B_REV), p(B_RCO), p(B_PCO), &
        p(B_KC), num_t_tab, t_min_tab, dt_tab, r_univ, tol, p(B_CV), nsetcv, p(B_TRANGE), lo, hi_, dt_chem, &
        0.0_c_double, dt_chem, 0_c_int, c_null_ptr, maxsteps, 0_c_int, 0.0_c_double, 0.0_c_double, &
        0.0_c_double, p(B_DCSIDX), p(B_DETADY), p(B_DZITDZ), c_null_ptr)
    call download(v, p(B_W), cnt(B_W))
    close(v)
    do b = 1, NBLK
      call ok(hipFree(p(b)), 'hipFree')
    end do
    write(*, '(a,i0,a)') 'dropin_check: case done (N_S = ', N_S, ')'
  end subroutine run_case

  subroutine upload(u, kind, n, d)
    integer, intent(in) :: u
    integer(c_int32_t), intent(in) :: kind
    integer(c_int64_t), intent(in) :: n
    type(c_ptr), intent(out) :: d
    real(c_double), allocatable, target :: a8(:)
    integer(c_int32_t), allocatable, target :: a4(:)
    integer(c_size_t) :: nb
    if (kind == 1) then
      allocate(a8(max(n, 1_c_int64_t)))
      if (n > 0) read(u) a8(1:n)
      nb = int(max(n, 1_c_int64_t), c_size_t) * 8_c_size_t
      call ok(hipMalloc(d, nb), 'hipMalloc')
      call ok(hipMemcpy(d, c_loc(a8), nb, 1_c_int), 'hipMemcpy')
    else
      allocate(a4(max(n, 1_c_int64_t)))
      if (n > 0) read(u) a4(1:n)
      nb = int(max(n, 1_c_int64_t), c_size_t) * 4_c_size_t
      call ok(hipMalloc(d, nb), 'hipMalloc')
      call ok(hipMemcpy(d, c_loc(a4), nb, 1_c_int), 'hipMemcpy')
    end if
  end subroutine upload

  subroutine download(v, d, n)
    integer, intent(in) :: v
    type(c_ptr), intent(in) :: d
    integer(c_int64_t), intent(in) :: n
    real(c_double), allocatable, target :: a8(:)
    allocate(a8(n))
    call ok(hipDeviceSynchronize(), 'hipDeviceSynchronize')
    call ok(hipMemcpy(c_loc(a8), d, int(n, c_size_t) * 8_c_size_t, 2_c_int), 'hipMemcpy')
    write(v) a8
  end subroutine download

  subroutine ok(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      write(*, '(a,a,i0)') what, ' failed: ', rc
      error stop 1
    end if
  end subroutine ok
end program dropin_check
