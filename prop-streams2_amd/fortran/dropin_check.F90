! dropin_check.F90 -- calls the gfx950 library through the Fortran drop-in module exactly as
! the reference's host code calls its kernel module: device arrays allocated with hipMalloc
! and passed as c_ptr handles, scalars by value.  Checks the RK bookkeeping
! (init_flux_cuf, update_flux_cuf, update_field_cuf) and the periodic halo exchange of the
! face layers (bcswap) against the same arithmetic on the host, and prints "dropin_check: ok".
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
  subroutine ok(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      write(*, '(a,a,i0)') what, ' failed: ', rc
      error stop 1
    end if
  end subroutine ok
end program dropin_check
