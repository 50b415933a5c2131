! psb_hip_plugin_mod: the dual-memory plug-in types over libpsb_hip, in Fortran.
!
! psb_d_hip_vect mirrors psb_d_vect_cuda (cuda/psb_d_cuda_vect_mod.F90:48-103):
! a host copy v(:) and a device multivector, with the host / sync / dev state
! of cuda/psb_d_cuda_hlg_mat_mod.F90:40-42; device operations bring the
! device copy up to date first and leave the result "dev", get_vect brings
! the host copy up to date.
!
! psb_d_hip_hlg_sparse_mat mirrors psb_d_cuda_hlg_sparse_mat
! (cuda/psb_d_cuda_hlg_mat_mod.F90:44-84): the CPU HLL arrays (hkoffs, irn,
! ja, val, idiag of psb_d_hll_sparse_mat) as the host copy, the device HLL
! behind a c_ptr, the same state and overrides: get_fmt ('HLG'),
! cp_from_coo, to_gpu / from_gpu / sync, vect_mv (device product with device
! vectors, syncing a "host" matrix first), csmv (host arrays: the CPU
! format's product on the host copy, syncing a "dev" matrix first, as
! cuda/impl/psb_d_cuda_hlg_vect_mv.F90:77-108 falls back to the host), get_diag,
! free.  The reference's types extend the CPU formats of PSBLAS's own modules;
! these stand alone and carry those formats' arrays themselves.
!
! psb_hip_dcg is the CG iteration of krylov/psb_dcg.F90:241-294 (istop 2,
! Jacobi preconditioner) over these types.
module psb_hip_plugin_mod
  use iso_c_binding
  use psb_hip_cbind_mod
  implicit none
  private

  integer, parameter, public :: psb_hip_host_ = -1, psb_hip_sync_ = 0, psb_hip_dev_ = 1

  type, public :: psb_d_hip_vect
    real(c_double), allocatable :: v(:)
    type(c_ptr) :: dev = c_null_ptr
    integer :: state = psb_hip_host_
  contains
    procedure :: all => hv_all
    procedure :: bld => hv_bld
    procedure :: sync => hv_sync
    procedure :: is_host => hv_is_host
    procedure :: is_dev => hv_is_dev
    procedure :: is_sync => hv_is_sync
    procedure :: set_host => hv_set_host
    procedure :: set_dev => hv_set_dev
    procedure :: set_sync => hv_set_sync
    procedure :: get_nrows => hv_get_nrows
    procedure :: get_vect => hv_get_vect
    procedure :: dot => hv_dot
    procedure :: nrm2 => hv_nrm2
    procedure :: axpby => hv_axpby
    procedure :: mlt => hv_mlt
    procedure :: free => hv_free
  end type psb_d_hip_vect

  type, public :: psb_d_hip_hlg_sparse_mat
    integer(c_int) :: nr = 0, nc = 0, nz = 0, hksz = 64, isz = 0
    integer(c_int), allocatable :: hkoffs(:), irn(:), ja(:), idiag(:)
    real(c_double), allocatable :: val(:)
    type(c_ptr) :: dev = c_null_ptr
    integer :: devstate = psb_hip_host_
  contains
    procedure :: get_fmt => hm_get_fmt
    procedure :: cp_from_coo => hm_cp_from_coo
    procedure :: to_gpu => hm_to_gpu
    procedure :: from_gpu => hm_from_gpu
    procedure :: sync => hm_sync
    procedure :: is_host => hm_is_host
    procedure :: is_dev => hm_is_dev
    procedure :: is_sync => hm_is_sync
    procedure :: set_host => hm_set_host
    procedure :: set_dev => hm_set_dev
    procedure :: set_sync => hm_set_sync
    procedure :: vect_mv => hm_vect_mv
    procedure :: csmv => hm_csmv
    procedure :: get_diag => hm_get_diag
    procedure :: free => hm_free
  end type psb_d_hip_hlg_sparse_mat

  public :: psb_hip_dcg

contains

  ! ------------------------------------------------------------------ vectors
  ! d_cuda_all: n entries, zero on both sides
  subroutine hv_all(x, n, info)
    class(psb_d_hip_vect), intent(inout), target :: x
    integer(c_int), intent(in) :: n
    integer(c_int), intent(out) :: info
    call x%free()
    allocate(x%v(n))
    x%v = 0.0d0
    info = psb_hip_mv_alloc(x%dev, 1_c_int, n, psb_hip_type_double_)
    if (info == 0) info = psb_hip_mv_write(x%dev, c_loc(x%v))
    x%state = psb_hip_sync_
  end subroutine hv_all

  ! d_cuda_bld_x: host values, "host" until the next device operation
  subroutine hv_bld(x, vals, info)
    class(psb_d_hip_vect), intent(inout) :: x
    real(c_double), intent(in) :: vals(:)
    integer(c_int), intent(out) :: info
    call x%free()
    allocate(x%v(size(vals)))
    x%v = vals
    info = psb_hip_mv_alloc(x%dev, 1_c_int, int(size(vals), c_int), psb_hip_type_double_)
    x%state = psb_hip_host_
  end subroutine hv_bld

  ! d_cuda_sync: host -> device write, dev -> host read, then "sync"
  subroutine hv_sync(x, info)
    class(psb_d_hip_vect), intent(inout), target :: x
    integer(c_int), intent(out) :: info
    info = 0
    if (x%state == psb_hip_host_) then
      info = psb_hip_mv_write(x%dev, c_loc(x%v))
    else if (x%state == psb_hip_dev_) then
      info = psb_hip_mv_read(x%dev, c_loc(x%v))
    end if
    x%state = psb_hip_sync_
  end subroutine hv_sync

  logical function hv_is_host(x)
    class(psb_d_hip_vect), intent(in) :: x
    hv_is_host = x%state == psb_hip_host_
  end function hv_is_host
  logical function hv_is_dev(x)
    class(psb_d_hip_vect), intent(in) :: x
    hv_is_dev = x%state == psb_hip_dev_
  end function hv_is_dev
  logical function hv_is_sync(x)
    class(psb_d_hip_vect), intent(in) :: x
    hv_is_sync = x%state == psb_hip_sync_
  end function hv_is_sync
  subroutine hv_set_host(x)
    class(psb_d_hip_vect), intent(inout) :: x
    x%state = psb_hip_host_
  end subroutine hv_set_host
  subroutine hv_set_dev(x)
    class(psb_d_hip_vect), intent(inout) :: x
    x%state = psb_hip_dev_
  end subroutine hv_set_dev
  subroutine hv_set_sync(x)
    class(psb_d_hip_vect), intent(inout) :: x
    x%state = psb_hip_sync_
  end subroutine hv_set_sync

  integer(c_int) function hv_get_nrows(x)
    class(psb_d_hip_vect), intent(in) :: x
    hv_get_nrows = 0
    if (allocated(x%v)) hv_get_nrows = int(size(x%v), c_int)
  end function hv_get_nrows

  ! the host values, read back first when the device copy is the newer one
  function hv_get_vect(x) result(res)
    class(psb_d_hip_vect), intent(inout) :: x
    real(c_double), allocatable :: res(:)
    integer(c_int) :: info
    if (x%state == psb_hip_dev_) call x%sync(info)
    res = x%v
  end function hv_get_vect

  ! bring the device copy up to date for a device operation
  subroutine to_dev(x, info)
    class(psb_d_hip_vect), intent(inout) :: x
    integer(c_int), intent(out) :: info
    info = 0
    if (x%state == psb_hip_host_) call x%sync(info)
  end subroutine to_dev

  ! d_cuda_dot_v: x . y on the device
  function hv_dot(x, y, info) result(res)
    class(psb_d_hip_vect), intent(inout) :: x
    type(psb_d_hip_vect), intent(inout) :: y
    integer(c_int), intent(out) :: info
    real(c_double) :: res, r(1)
    call to_dev(x, info)
    if (info == 0) call to_dev(y, info)
    if (info == 0) info = psb_hip_mv_dot(r, x%get_nrows(), x%dev, y%dev)
    res = r(1)
  end function hv_dot

  ! d_cuda_nrm2
  function hv_nrm2(x, info) result(res)
    class(psb_d_hip_vect), intent(inout) :: x
    integer(c_int), intent(out) :: info
    real(c_double) :: res, r(1)
    call to_dev(x, info)
    if (info == 0) info = psb_hip_mv_nrm2(r, x%get_nrows(), x%dev)
    res = r(1)
  end function hv_nrm2

  ! d_cuda_axpby_v: y = alpha x + beta y, y left "dev"
  subroutine hv_axpby(y, alpha, x, beta, info)
    class(psb_d_hip_vect), intent(inout) :: y
    real(c_double), intent(in) :: alpha, beta
    type(psb_d_hip_vect), intent(inout) :: x
    integer(c_int), intent(out) :: info
    call to_dev(x, info)
    if (info == 0) call to_dev(y, info)
    if (info == 0) info = psb_hip_mv_axpby(y%get_nrows(), alpha, x%dev, beta, y%dev)
    call y%set_dev()
  end subroutine hv_axpby

  ! d_cuda_mlt_v_2: z = alpha x .* y + beta z, z left "dev"
  subroutine hv_mlt(z, alpha, x, y, beta, info)
    class(psb_d_hip_vect), intent(inout) :: z
    real(c_double), intent(in) :: alpha, beta
    type(psb_d_hip_vect), intent(inout) :: x, y
    integer(c_int), intent(out) :: info
    call to_dev(x, info)
    if (info == 0) call to_dev(y, info)
    if (info == 0) call to_dev(z, info)
    if (info == 0) info = psb_hip_mv_axybz(z%get_nrows(), alpha, x%dev, y%dev, beta, z%dev)
    call z%set_dev()
  end subroutine hv_mlt

  subroutine hv_free(x)
    class(psb_d_hip_vect), intent(inout) :: x
    if (c_associated(x%dev)) call psb_hip_mv_free(x%dev)
    x%dev = c_null_ptr
    if (allocated(x%v)) deallocate(x%v)
    x%state = psb_hip_host_
  end subroutine hv_free

  ! ------------------------------------------------------------------ HLG matrix
  ! psb_d_cuda_hlg_get_fmt (cuda/psb_d_cuda_hlg_mat_mod.F90:268-272)
  function hm_get_fmt(a) result(res)
    class(psb_d_hip_hlg_sparse_mat), intent(in) :: a
    character(len=5) :: res
    res = 'HLG'
  end function hm_get_fmt

  ! cp_from_coo on the host: COO fix (sort by row, then column, duplicates
  ! summed), then the CPU HLL arrays with hack size hksz; the matrix is
  ! "host" until to_gpu / sync writes the device copy
  subroutine hm_cp_from_coo(a, nr, nc, nz, ia, ja, val, hksz, info)
    class(psb_d_hip_hlg_sparse_mat), intent(inout), target :: a
    integer(c_int), intent(in) :: nr, nc, nz, hksz
    integer(c_int), intent(in) :: ia(:), ja(:)
    real(c_double), intent(in) :: val(:)
    integer(c_int), intent(out) :: info
    integer(c_int), allocatable :: cia(:), cja(:), idisp(:)
    real(c_double), allocatable, target :: cval(:)
    integer(c_int64_t) :: nzout, isz
    call a%free()
    cia = ia(1:nz)
    cja = ja(1:nz)
    cval = val(1:nz)
    info = psb_hip_fix_coo_d(nr, nc, int(nz, c_int64_t), cia, cja, cval, 1_c_int, nzout)
    if (info /= 0) return
    a%nr = nr
    a%nc = nc
    a%nz = int(nzout, c_int)
    a%hksz = hksz
    allocate(a%irn(max(nr, 1)), idisp(nr + 1), a%hkoffs((nr + hksz - 1) / hksz + 1))
    info = psb_hip_hll_offsets(nr, nzout, cia, hksz, a%irn, idisp, a%hkoffs, isz)
    if (info /= 0) return
    a%isz = int(isz, c_int)
    allocate(a%ja(max(a%isz, 1)), a%val(max(a%isz, 1)), a%idiag(max(nr, 1)))
    info = psb_hip_host_hll_fill(nr, hksz, a%irn, a%hkoffs, idisp, cja, c_loc(cval), c_null_ptr, &
         & psb_hip_type_double_, a%ja, c_loc(a%val), a%idiag)
    a%devstate = psb_hip_host_
  end subroutine hm_cp_from_coo

  ! psb_d_cuda_hlg_to_gpu: allocate the device HLL when needed and write the host copy
  subroutine hm_to_gpu(a, info)
    class(psb_d_hip_hlg_sparse_mat), intent(inout), target :: a
    integer(c_int), intent(out) :: info
    info = 0
    if (.not. c_associated(a%dev)) &
         & info = psb_hip_hll_alloc(a%dev, a%hksz, a%nr, a%nz, a%isz, psb_hip_type_double_, 1_c_int)
    if (info == 0) info = psb_hip_hll_write(a%dev, c_loc(a%val), a%ja, a%hkoffs, a%irn, a%idiag)
    if (info == 0) a%devstate = psb_hip_sync_
  end subroutine hm_to_gpu

  ! psb_d_cuda_hlg_from_gpu: read the device arrays into the host copy
  subroutine hm_from_gpu(a, info)
    class(psb_d_hip_hlg_sparse_mat), intent(inout), target :: a
    integer(c_int), intent(out) :: info
    info = psb_hip_hll_read(a%dev, c_loc(a%val), a%ja, a%hkoffs, a%irn, a%idiag)
    if (info == 0) a%devstate = psb_hip_sync_
  end subroutine hm_from_gpu

  ! d_cuda_hlg_sync: host -> to_gpu, dev -> from_gpu
  subroutine hm_sync(a, info)
    class(psb_d_hip_hlg_sparse_mat), intent(inout) :: a
    integer(c_int), intent(out) :: info
    info = 0
    if (a%devstate == psb_hip_host_) then
      call a%to_gpu(info)
    else if (a%devstate == psb_hip_dev_) then
      call a%from_gpu(info)
    end if
  end subroutine hm_sync

  logical function hm_is_host(a)
    class(psb_d_hip_hlg_sparse_mat), intent(in) :: a
    hm_is_host = a%devstate == psb_hip_host_
  end function hm_is_host
  logical function hm_is_dev(a)
    class(psb_d_hip_hlg_sparse_mat), intent(in) :: a
    hm_is_dev = a%devstate == psb_hip_dev_
  end function hm_is_dev
  logical function hm_is_sync(a)
    class(psb_d_hip_hlg_sparse_mat), intent(in) :: a
    hm_is_sync = a%devstate == psb_hip_sync_
  end function hm_is_sync
  subroutine hm_set_host(a)
    class(psb_d_hip_hlg_sparse_mat), intent(inout) :: a
    a%devstate = psb_hip_host_
  end subroutine hm_set_host
  subroutine hm_set_dev(a)
    class(psb_d_hip_hlg_sparse_mat), intent(inout) :: a
    a%devstate = psb_hip_dev_
  end subroutine hm_set_dev
  subroutine hm_set_sync(a)
    class(psb_d_hip_hlg_sparse_mat), intent(inout) :: a
    a%devstate = psb_hip_sync_
  end subroutine hm_set_sync

  ! psb_d_cuda_hlg_vect_mv with device vectors: y = alpha A x + beta y on the
  ! device (a "host" matrix is synced first), y left "dev"
  subroutine hm_vect_mv(a, alpha, x, beta, y, info)
    class(psb_d_hip_hlg_sparse_mat), intent(inout) :: a
    real(c_double), intent(in) :: alpha, beta
    type(psb_d_hip_vect), intent(inout) :: x, y
    integer(c_int), intent(out) :: info
    info = 0
    if (a%is_host()) call a%sync(info)
    if (info == 0) call to_dev(x, info)
    if (info == 0 .and. beta /= 0.0d0) call to_dev(y, info)
    if (info == 0) info = psb_hip_hll_spmv(a%dev, alpha, x%dev, beta, y%dev)
    call y%set_dev()
  end subroutine hm_vect_mv

  ! csmv on host arrays: the CPU HLL product on the host copy (a "dev"
  ! matrix is read back first), as the reference does for vectors that are
  ! not device vectors (cuda/impl/psb_d_cuda_hlg_vect_mv.F90:77-108)
  subroutine hm_csmv(a, alpha, x, beta, y, info)
    class(psb_d_hip_hlg_sparse_mat), intent(inout), target :: a
    real(c_double), intent(in) :: alpha, beta
    real(c_double), intent(in), target, contiguous :: x(:)
    real(c_double), intent(inout), target, contiguous :: y(:)
    integer(c_int), intent(out) :: info
    info = 0
    if (a%is_dev()) call a%sync(info)
    if (info == 0) info = psb_hip_host_hll_spmv(a%nr, a%hksz, a%hkoffs, a%irn, a%ja, &
         & c_loc(a%val), c_null_ptr, psb_hip_type_double_, alpha, c_loc(x), beta, c_loc(y))
  end subroutine hm_csmv

  ! get_diag on the device (psb_d_diag_precbld's first step), d left "dev"
  subroutine hm_get_diag(a, d, info)
    class(psb_d_hip_hlg_sparse_mat), intent(inout) :: a
    type(psb_d_hip_vect), intent(inout) :: d
    integer(c_int), intent(out) :: info
    info = 0
    if (a%is_host()) call a%sync(info)
    if (info == 0) info = psb_hip_hll_get_diag(a%dev, d%dev)
    call d%set_dev()
  end subroutine hm_get_diag

  subroutine hm_free(a)
    class(psb_d_hip_hlg_sparse_mat), intent(inout) :: a
    if (c_associated(a%dev)) call psb_hip_hll_free(a%dev)
    a%dev = c_null_ptr
    if (allocated(a%hkoffs)) deallocate(a%hkoffs)
    if (allocated(a%irn)) deallocate(a%irn)
    if (allocated(a%ja)) deallocate(a%ja)
    if (allocated(a%val)) deallocate(a%val)
    if (allocated(a%idiag)) deallocate(a%idiag)
    a%devstate = psb_hip_host_
  end subroutine hm_free

  ! ------------------------------------------------------------------ CG
  ! krylov/psb_dcg.F90:241-294 with the Jacobi preconditioner d (inverse
  ! diagonal) and istop 2: ||r|| / ||b|| <= eps.  x holds the initial guess.
  subroutine psb_hip_dcg(a, d, b, x, eps, itmax, iter, err, info)
    type(psb_d_hip_hlg_sparse_mat), intent(inout) :: a
    type(psb_d_hip_vect), intent(inout) :: d, b, x
    real(c_double), intent(in) :: eps
    integer(c_int), intent(in) :: itmax
    integer(c_int), intent(out) :: iter, info
    real(c_double), intent(out) :: err
    type(psb_d_hip_vect) :: r, z, p, q
    real(c_double) :: rho, rho_old, sigma, alpha, beta, bn, rn
    integer(c_int) :: n
    n = b%get_nrows()
    call r%all(n, info)
    call z%all(n, info)
    call p%all(n, info)
    call q%all(n, info)
    ! r = b - A x
    call r%axpby(1.0d0, b, 0.0d0, info)
    call a%vect_mv(-1.0d0, x, 1.0d0, r, info)
    bn = b%nrm2(info)
    rn = r%nrm2(info)
    err = rn / bn
    iter = 0
    rho = 0.0d0
    do while (err > eps .and. iter < itmax .and. info == 0)
      iter = iter + 1
      call z%mlt(1.0d0, d, r, 0.0d0, info)
      rho_old = rho
      rho = r%dot(z, info)
      if (iter == 1) then
        call p%axpby(1.0d0, z, 0.0d0, info)
      else
        if (rho_old == 0.0d0) exit
        beta = rho / rho_old
        call p%axpby(1.0d0, z, beta, info)
      end if
      call a%vect_mv(1.0d0, p, 0.0d0, q, info)
      sigma = p%dot(q, info)
      if (sigma == 0.0d0) exit
      alpha = rho / sigma
      call x%axpby(alpha, p, 1.0d0, info)
      call r%axpby(-alpha, q, 1.0d0, info)
      rn = r%nrm2(info)
      err = rn / bn
    end do
    call r%free()
    call z%free()
    call p%free()
    call q%free()
  end subroutine psb_hip_dcg

end module psb_hip_plugin_mod
