! Fortran driver over psb_hip_cbind_mod: the psb_spasb(mold=HLG) + psb_spmm +
! psb_gedot + psb_krylov('CG') call path of the reference GPU plug-in, issued
! from Fortran through the C ABI.  SPD Poisson (6,-1) on an L^3 grid,
! x = ones: checks the exact product A*1, then solves A x = A*1 with fused
! CG + DIAG.  Prints PASS or FAIL.
program hip_pde3d_driver
  use iso_c_binding
  use psb_hip_cbind_mod
  implicit none
  integer, parameter :: L = 20, hksz = 64
  integer(c_int) :: n, info, ix, iy, iz, row, k, noffs, iter
  integer(c_int64_t) :: nz, nzout, isz
  integer(c_int), allocatable :: ia(:), ja(:), irn(:), idisp(:), hkoffs(:)
  real(c_double), allocatable :: val(:)
  real(c_double), allocatable, target :: xh(:), yh(:)
  type(c_ptr) :: a, x, y, d, b
  real(c_double), target :: dres(1)
  real(c_double) :: err
  logical :: ok

  n = L*L*L
  allocate(ia(7*n), ja(7*n), val(7*n))
  k = 0
  do iz = 1, L
    do iy = 1, L
      do ix = 1, L
        row = ((iz-1)*L + (iy-1))*L + ix
        call put(row, 6.0d0)
        if (ix > 1) call put(row-1, -1.0d0)
        if (ix < L) call put(row+1, -1.0d0)
        if (iy > 1) call put(row-L, -1.0d0)
        if (iy < L) call put(row+L, -1.0d0)
        if (iz > 1) call put(row-L*L, -1.0d0)
        if (iz < L) call put(row+L*L, -1.0d0)
      end do
    end do
  end do
  nz = k
  info = psb_hip_init(0_c_int)
  ok = (info == psb_hip_success_)
  info = psb_hip_fix_coo_d(n, n, nz, ia, ja, val, 1_c_int, nzout)
  ok = ok .and. (info == 0) .and. (nzout == nz)
  noffs = (n + hksz - 1)/hksz
  allocate(irn(n), idisp(n+1), hkoffs(noffs+1))
  info = psb_hip_hll_offsets(n, nz, ia, hksz, irn, idisp, hkoffs, isz)
  ok = ok .and. (info == 0)
  info = psb_hip_hll_alloc(a, hksz, n, int(nz, c_int), int(isz, c_int), psb_hip_type_double_, 1_c_int)
  ok = ok .and. (info == 0)
  info = psb_hip_hll_copy_coo(n, n, int(nz, c_int), hksz, noffs, int(isz, c_int), irn, hkoffs, &
       & idisp, ja, val, a)
  ok = ok .and. (info == 0)

  allocate(xh(n), yh(n))
  xh = 1.0d0
  info = psb_hip_mv_alloc(x, 1_c_int, n, psb_hip_type_double_)
  info = psb_hip_mv_alloc(y, 1_c_int, n, psb_hip_type_double_)
  info = psb_hip_mv_write(x, c_loc(xh))
  ! y = A*1: exact row sums 6 - #neighbours
  info = psb_hip_hll_spmv(a, 1.0d0, x, 0.0d0, y)
  ok = ok .and. (info == 0)
  info = psb_hip_mv_read(y, c_loc(yh))
  ok = ok .and. all(yh == 7.0d0 - real(irn, c_double))
  ! dot(y, 1) = sum of row sums
  info = psb_hip_mv_dot(dres, n, y, x)
  ok = ok .and. (dres(1) == sum(7.0d0 - real(irn, c_double)))

  ! solve A x = y with fused CG + DIAG from x = 0
  info = psb_hip_mv_alloc(d, 1_c_int, n, psb_hip_type_double_)
  info = psb_hip_hll_get_diag(a, d)
  info = psb_hip_mv_diag_inv(n, d)
  b = y
  xh = 0.0d0
  info = psb_hip_mv_write(x, c_loc(xh))
  info = psb_hip_krylov_fused(0_c_int, psb_hip_fmt_hll_, a, d, b, x, 1.0d-10, 1000_c_int, &
       & 2_c_int, 8_c_int, iter, err, c_null_ptr)
  ok = ok .and. (info == 0) .and. (err <= 1.0d-10)
  info = psb_hip_mv_read(x, c_loc(xh))
  ok = ok .and. (maxval(abs(xh - 1.0d0)) < 1.0d-7)
  write(*, '(a,i0,a,es10.3)') 'fortran driver: cg iterations ', iter, ' err ', err
  call psb_hip_mv_free(x)
  call psb_hip_mv_free(y)
  call psb_hip_mv_free(d)
  call psb_hip_hll_free(a)
  info = psb_hip_exit()
  if (ok) then
    write(*, '(a)') 'PASS'
  else
    write(*, '(a)') 'FAIL'
    stop 1
  end if

contains

  subroutine put(col, v)
    integer(c_int), intent(in) :: col
    real(c_double), intent(in) :: v
    k = k + 1
    ia(k) = row
    ja(k) = col
    val(k) = v
  end subroutine put

end program hip_pde3d_driver
