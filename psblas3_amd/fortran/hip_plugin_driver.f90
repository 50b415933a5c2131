! Fortran driver over psb_hip_plugin_mod: the dual-memory plug-in types used
! as a PSBLAS program uses psb_d_cuda_hlg_sparse_mat / psb_d_vect_cuda.
! SPD Poisson (6,-1) on an L^3 grid:
!  - a matrix built on the host is "host"; its host product (csmv) gives the
!    exact row sums of A*1;
!  - a device product (vect_mv) syncs it ("sync"), runs on the device and
!    leaves y "dev"; y read back equals the host product;
!  - a matrix marked "dev" is read back by the next host product;
!  - CG (krylov/psb_dcg.F90 over the vector type) on A x = A*1 converges in
!    the fused solver's iteration count +-2 with x close to 1.
! Prints PASS or FAIL.
program hip_plugin_driver
  use iso_c_binding
  use psb_hip_cbind_mod
  use psb_hip_plugin_mod
  implicit none
  integer, parameter :: L = 24
  integer(c_int) :: n, info, ix, iy, iz, row, k, iter, iter_fused
  integer(c_int), allocatable :: ia(:), ja(:)
  real(c_double), allocatable :: val(:), yh(:), yd(:), ones(:), w(:)
  real(c_double), allocatable, target :: xs(:)
  type(psb_d_hip_hlg_sparse_mat) :: a
  type(psb_d_hip_vect) :: x, y, d, b, xc
  real(c_double) :: err, err_fused
  logical :: ok

  n = L*L*L
  allocate(ia(7*n), ja(7*n), val(7*n))
  k = 0
  ! entries in the generator's order (diagonal first, neighbours after); the
  ! COO fix sorts them
  do iz = 1, L
    do iy = 1, L
      do ix = 1, L
        row = ((iz-1)*L + (iy-1))*L + ix
        call put(row, 6.0d0)
        if (iz < L) call put(row+L*L, -1.0d0)
        if (ix > 1) call put(row-1, -1.0d0)
        if (iy < L) call put(row+L, -1.0d0)
        if (ix < L) call put(row+1, -1.0d0)
        if (iz > 1) call put(row-L*L, -1.0d0)
        if (iy > 1) call put(row-L, -1.0d0)
      end do
    end do
  end do
  info = psb_hip_init(0_c_int)
  ok = info == 0

  call a%cp_from_coo(n, n, k, ia, ja, val, 64_c_int, info)
  ok = ok .and. info == 0 .and. a%is_host() .and. trim(a%get_fmt()) == 'HLG'
  ok = ok .and. a%nz == k
  allocate(ones(n), yh(n))
  ones = 1.0d0
  yh = 0.0d0
  ! host product on the host copy: exact row sums 7 - (entries in the row)
  call a%csmv(1.0d0, ones, 0.0d0, yh, info)
  ok = ok .and. info == 0 .and. a%is_host()
  ok = ok .and. all(yh == 7.0d0 - real(a%irn(1:n), c_double))
  if (.not. ok) write(*, '(a)') 'host build / host product failed'

  ! device product with device vectors: the matrix is synced first
  call x%bld(ones, info)
  ok = ok .and. x%is_host()
  call y%all(n, info)
  call a%vect_mv(1.0d0, x, 0.0d0, y, info)
  ok = ok .and. info == 0 .and. a%is_sync() .and. x%is_sync() .and. y%is_dev()
  yd = y%get_vect()
  ok = ok .and. y%is_sync() .and. all(yd == yh)
  if (.not. ok) write(*, '(a)') 'device product failed'

  ! a non-trivial x: host and device products within the per-row bound
  allocate(xs(n), w(n))
  do row = 1, n
    xs(row) = sin(real(row, c_double))
  end do
  call xc%bld(xs, info)
  call a%vect_mv(1.0d0, xc, 0.0d0, y, info)
  yd = y%get_vect()
  w = 0.0d0
  call a%set_dev()           ! the device copy marked newer: read back first
  call a%csmv(1.0d0, xs, 0.0d0, w, info)
  ok = ok .and. info == 0 .and. a%is_sync()
  ok = ok .and. maxval(abs(yd - w)) <= 4.0d0 * 8.0d0 * epsilon(1.0d0) * 12.0d0
  if (.not. ok) write(*, '(a)') 'host / device products differ'

  ! CG + Jacobi over the plug-in types against the fused solver
  call d%all(n, info)
  call a%get_diag(d, info)
  info = psb_hip_mv_diag_inv(n, d%dev)
  call b%all(n, info)
  call a%vect_mv(1.0d0, x, 0.0d0, b, info)
  call xc%all(n, info)
  call psb_hip_dcg(a, d, b, xc, 1.0d-9, 1000_c_int, iter, err, info)
  ok = ok .and. info == 0 .and. err <= 1.0d-9
  yd = xc%get_vect()
  ok = ok .and. maxval(abs(yd - 1.0d0)) < 1.0d-6
  call y%all(n, info)
  info = psb_hip_krylov_fused(0_c_int, psb_hip_fmt_hll_, a%dev, d%dev, b%dev, y%dev, 1.0d-9, &
       & 1000_c_int, 2_c_int, 8_c_int, iter_fused, err_fused, c_null_ptr)
  ok = ok .and. info == 0 .and. abs(iter - iter_fused) <= 2
  write(*, '(a,i0,a,i0,a,es10.3)') 'plugin driver: cg iterations ', iter, ' (fused ', &
       & iter_fused, ') err ', err

  call x%free()
  call y%free()
  call d%free()
  call b%free()
  call xc%free()
  call a%free()
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

end program hip_plugin_driver
