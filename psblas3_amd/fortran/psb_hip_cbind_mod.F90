! psb_hip_cbind_mod: iso_c_binding interfaces to libpsb_hip (include/psb_hip.h).
!
! This is the module a PSBLAS maintainer adds in place of the bind(c) sets of
! cuda/hlldev_mod.F90, cuda/elldev_mod.F90, cuda/psb_d_vectordev_mod.F90 and
! cuda/psb_cuda_env_mod.F90: every interface below binds one entry point of the
! C ABI with the same argument meaning as the CUDA bridge function it replaces.
module psb_hip_cbind_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: psb_hip_success_     = 0
  integer(c_int), parameter :: psb_hip_unsupported_ = 1
  integer(c_int), parameter :: psb_hip_invalid_     = 2
  integer(c_int), parameter :: psb_hip_outofmem_    = 3
  integer(c_int), parameter :: psb_hip_type_int_    = 0
  integer(c_int), parameter :: psb_hip_type_float_  = 1
  integer(c_int), parameter :: psb_hip_type_double_ = 2
  integer(c_int), parameter :: psb_hip_fmt_hll_     = 0
  integer(c_int), parameter :: psb_hip_fmt_ell_     = 1
  integer(c_int), parameter :: psb_hip_fmt_csr_     = 2

  interface
    ! ---------------------------------------------------------------- environment
    function psb_hip_init(dev) result(res) bind(c, name='psb_hip_init')
      import :: c_int
      integer(c_int), value :: dev
      integer(c_int) :: res
    end function psb_hip_init
    function psb_hip_exit() result(res) bind(c, name='psb_hip_exit')
      import :: c_int
      integer(c_int) :: res
    end function psb_hip_exit
    function psb_hip_device_sync() result(res) bind(c, name='psb_hip_device_sync')
      import :: c_int
      integer(c_int) :: res
    end function psb_hip_device_sync
    function psb_hip_get_multiprocessors() result(res) bind(c, name='psb_hip_get_multiprocessors')
      import :: c_int
      integer(c_int) :: res
    end function psb_hip_get_multiprocessors
    function psb_hip_memory_peak_bandwidth() result(res) &
         & bind(c, name='psb_hip_memory_peak_bandwidth')
      import :: c_double
      real(c_double) :: res
    end function psb_hip_memory_peak_bandwidth

    ! ---------------------------------------------------------------- multivectors
    function psb_hip_mv_alloc(mv, count, size, elem_type) result(res) &
         & bind(c, name='psb_hip_mv_alloc')
      import :: c_int, c_ptr
      type(c_ptr) :: mv
      integer(c_int), value :: count, size, elem_type
      integer(c_int) :: res
    end function psb_hip_mv_alloc
    subroutine psb_hip_mv_free(mv) bind(c, name='psb_hip_mv_free')
      import :: c_ptr
      type(c_ptr), value :: mv
    end subroutine psb_hip_mv_free
    function psb_hip_mv_write(mv, host) result(res) bind(c, name='psb_hip_mv_write')
      import :: c_int, c_ptr
      type(c_ptr), value :: mv, host
      integer(c_int) :: res
    end function psb_hip_mv_write
    function psb_hip_mv_read(mv, host) result(res) bind(c, name='psb_hip_mv_read')
      import :: c_int, c_ptr
      type(c_ptr), value :: mv, host
      integer(c_int) :: res
    end function psb_hip_mv_read
    function psb_hip_mv_dot(res_, n, x, y) result(res) bind(c, name='psb_hip_mv_dot')
      import :: c_int, c_ptr, c_double
      real(c_double) :: res_(*)
      integer(c_int), value :: n
      type(c_ptr), value :: x, y
      integer(c_int) :: res
    end function psb_hip_mv_dot
    function psb_hip_mv_nrm2(res_, n, x) result(res) bind(c, name='psb_hip_mv_nrm2')
      import :: c_int, c_ptr, c_double
      real(c_double) :: res_(*)
      integer(c_int), value :: n
      type(c_ptr), value :: x
      integer(c_int) :: res
    end function psb_hip_mv_nrm2
    function psb_hip_mv_axpby(n, alpha, x, beta, y) result(res) bind(c, name='psb_hip_mv_axpby')
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: n
      real(c_double), value :: alpha, beta
      type(c_ptr), value :: x, y
      integer(c_int) :: res
    end function psb_hip_mv_axpby
    function psb_hip_mv_axybz(n, alpha, x, y, beta, z) result(res) bind(c, name='psb_hip_mv_axybz')
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: n
      real(c_double), value :: alpha, beta
      type(c_ptr), value :: x, y, z
      integer(c_int) :: res
    end function psb_hip_mv_axybz
    function psb_hip_mv_diag_inv(n, d) result(res) bind(c, name='psb_hip_mv_diag_inv')
      import :: c_int, c_ptr
      integer(c_int), value :: n
      type(c_ptr), value :: d
      integer(c_int) :: res
    end function psb_hip_mv_diag_inv

    ! ---------------------------------------------------------------- host assembly
    function psb_hip_fix_coo_d(nr, nc, nz, ia, ja, val, dupl, nzout) result(res) &
         & bind(c, name='psb_hip_fix_coo_d')
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: nr, nc, dupl
      integer(c_int64_t), value :: nz
      integer(c_int) :: ia(*), ja(*)
      real(c_double) :: val(*)
      integer(c_int64_t) :: nzout
      integer(c_int) :: res
    end function psb_hip_fix_coo_d
    function psb_hip_hll_offsets(nr, nz, ia, hksz, irn, idisp, hkoffs, isz) result(res) &
         & bind(c, name='psb_hip_hll_offsets')
      import :: c_int, c_int64_t
      integer(c_int), value :: nr, hksz
      integer(c_int64_t), value :: nz
      integer(c_int) :: ia(*), irn(*), idisp(*), hkoffs(*)
      integer(c_int64_t) :: isz
      integer(c_int) :: res
    end function psb_hip_hll_offsets

    ! host copy of the HLL format (psi_d_convert_hll_from_coo) and its product
    function psb_hip_host_hll_fill(nr, hksz, irn, hkoffs, idisp, cja, cval, perm, elem_type, &
         & ja, val, idiag) result(res) bind(c, name='psb_hip_host_hll_fill')
      import :: c_int, c_ptr
      integer(c_int), value :: nr, hksz, elem_type
      integer(c_int) :: irn(*), hkoffs(*), idisp(*), cja(*), ja(*), idiag(*)
      type(c_ptr), value :: cval, perm, val
      integer(c_int) :: res
    end function psb_hip_host_hll_fill
    function psb_hip_host_hll_spmv(nr, hksz, hkoffs, irn, ja, val, perm, elem_type, alpha, x, &
         & beta, y) result(res) bind(c, name='psb_hip_host_hll_spmv')
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: nr, hksz, elem_type
      integer(c_int) :: hkoffs(*), irn(*), ja(*)
      type(c_ptr), value :: val, perm, x, y
      real(c_double), value :: alpha, beta
      integer(c_int) :: res
    end function psb_hip_host_hll_spmv

    ! ---------------------------------------------------------------- HLL (HLG)
    function psb_hip_hll_write(dmat, val, ja, hkoffs, irn, idiag) result(res) &
         & bind(c, name='psb_hip_hll_write')
      import :: c_int, c_ptr
      type(c_ptr), value :: dmat, val
      integer(c_int) :: ja(*), hkoffs(*), irn(*), idiag(*)
      integer(c_int) :: res
    end function psb_hip_hll_write
    function psb_hip_hll_read(dmat, val, ja, hkoffs, irn, idiag) result(res) &
         & bind(c, name='psb_hip_hll_read')
      import :: c_int, c_ptr
      type(c_ptr), value :: dmat, val
      integer(c_int) :: ja(*), hkoffs(*), irn(*), idiag(*)
      integer(c_int) :: res
    end function psb_hip_hll_read
    function psb_hip_hll_alloc(dmat, hksz, rows, nzeros, allocsize, elem_type, first_index) &
         & result(res) bind(c, name='psb_hip_hll_alloc')
      import :: c_int, c_ptr
      type(c_ptr) :: dmat
      integer(c_int), value :: hksz, rows, nzeros, allocsize, elem_type, first_index
      integer(c_int) :: res
    end function psb_hip_hll_alloc
    subroutine psb_hip_hll_free(dmat) bind(c, name='psb_hip_hll_free')
      import :: c_ptr
      type(c_ptr), value :: dmat
    end subroutine psb_hip_hll_free
    function psb_hip_hll_copy_coo(nr, nc, nza, hksz, noffs, isz, irn, hoffs, idisp, ja, val, &
         & dmat) result(res) bind(c, name='psb_hip_hll_copy_coo')
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: nr, nc, nza, hksz, noffs, isz
      integer(c_int) :: irn(*), hoffs(*), idisp(*), ja(*)
      real(c_double) :: val(*)
      type(c_ptr), value :: dmat
      integer(c_int) :: res
    end function psb_hip_hll_copy_coo
    function psb_hip_hll_spmv(dmat, alpha, x, beta, y) result(res) bind(c, name='psb_hip_hll_spmv')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: dmat, x, y
      real(c_double), value :: alpha, beta
      integer(c_int) :: res
    end function psb_hip_hll_spmv
    function psb_hip_hll_get_diag(dmat, d) result(res) bind(c, name='psb_hip_hll_get_diag')
      import :: c_int, c_ptr
      type(c_ptr), value :: dmat, d
      integer(c_int) :: res
    end function psb_hip_hll_get_diag

    ! ---------------------------------------------------------------- ELL / CSR
    function psb_hip_ell_spmv(dmat, alpha, x, beta, y) result(res) bind(c, name='psb_hip_ell_spmv')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: dmat, x, y
      real(c_double), value :: alpha, beta
      integer(c_int) :: res
    end function psb_hip_ell_spmv
    function psb_hip_csr_spmv(dmat, alpha, x, beta, y) result(res) bind(c, name='psb_hip_csr_spmv')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: dmat, x, y
      real(c_double), value :: alpha, beta
      integer(c_int) :: res
    end function psb_hip_csr_spmv
    ! scals / scal of the three formats, on the device copy
    function psb_hip_hll_scals(dmat, s) result(res) bind(c, name='psb_hip_hll_scals')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: dmat
      real(c_double), value :: s
      integer(c_int) :: res
    end function psb_hip_hll_scals
    function psb_hip_hll_scal(dmat, d, left) result(res) bind(c, name='psb_hip_hll_scal')
      import :: c_int, c_ptr
      type(c_ptr), value :: dmat, d
      integer(c_int), value :: left
      integer(c_int) :: res
    end function psb_hip_hll_scal
    function psb_hip_ell_scals(dmat, s) result(res) bind(c, name='psb_hip_ell_scals')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: dmat
      real(c_double), value :: s
      integer(c_int) :: res
    end function psb_hip_ell_scals
    function psb_hip_ell_scal(dmat, d, left) result(res) bind(c, name='psb_hip_ell_scal')
      import :: c_int, c_ptr
      type(c_ptr), value :: dmat, d
      integer(c_int), value :: left
      integer(c_int) :: res
    end function psb_hip_ell_scal
    function psb_hip_csr_scals(dmat, s) result(res) bind(c, name='psb_hip_csr_scals')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: dmat
      real(c_double), value :: s
      integer(c_int) :: res
    end function psb_hip_csr_scals
    function psb_hip_csr_scal(dmat, d, left) result(res) bind(c, name='psb_hip_csr_scal')
      import :: c_int, c_ptr
      type(c_ptr), value :: dmat, d
      integer(c_int), value :: left
      integer(c_int) :: res
    end function psb_hip_csr_scal
    ! spsvCSRGDevice: y = alpha * op(T)^-1 x + beta * y, lower / unit from
    ! a%is_lower() / a%is_unit(), trans 0 'N', 1 'T', 2 'C'
    function psb_hip_csr_spsv(dmat, lower, unit, trans, alpha, x, beta, y) result(res) &
         & bind(c, name='psb_hip_csr_spsv')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: dmat, x, y
      integer(c_int), value :: lower, unit, trans
      real(c_double), value :: alpha, beta
      integer(c_int) :: res
    end function psb_hip_csr_spsv

    ! ---------------------------------------------------------------- solvers
    function psb_hip_krylov_fused(method, fmt, dmat, diag, b, x, eps, itmax, istop, &
         & check_every, iter, err, hist) result(res) bind(c, name='psb_hip_krylov_fused')
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: method, fmt, itmax, istop, check_every
      type(c_ptr), value :: dmat, diag, b, x, hist
      real(c_double), value :: eps
      integer(c_int) :: iter
      real(c_double) :: err
      integer(c_int) :: res
    end function psb_hip_krylov_fused

    function psb_hip_krylov_mixed(method, dmat, diag, b, x, eps, outer_itmax, inner_eps, &
         & inner_itmax, check_every, variant, outer_its, inner_its, err) result(res) &
         & bind(c, name='psb_hip_krylov_mixed')
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: method, outer_itmax, inner_itmax, check_every, variant
      type(c_ptr), value :: dmat, diag, b, x
      real(c_double), value :: eps, inner_eps
      integer(c_int) :: outer_its, inner_its
      real(c_double) :: err
      integer(c_int) :: res
    end function psb_hip_krylov_mixed

    ! ---------------------------------------------------------------- batched solves
    function psb_hip_batch_create(batch, method, max_systems, arena_bytes, check_every, &
         & delta, inner_eps, use_graph) result(res) bind(c, name='psb_hip_batch_create')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr) :: batch
      integer(c_int), value :: method, max_systems, check_every, use_graph
      integer(c_int64_t), value :: arena_bytes
      real(c_double), value :: delta, inner_eps
      integer(c_int) :: res
    end function psb_hip_batch_create

    function psb_hip_batch_free(batch) result(res) bind(c, name='psb_hip_batch_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int) :: res
    end function psb_hip_batch_free

    function psb_hip_batch_register(batch, dmat, diag, mat_id) result(res) &
         & bind(c, name='psb_hip_batch_register')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch, dmat, diag
      integer(c_int) :: mat_id
      integer(c_int) :: res
    end function psb_hip_batch_register

    function psb_hip_batch_footprint(batch, mat_id) result(res) &
         & bind(c, name='psb_hip_batch_footprint')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: batch
      integer(c_int), value :: mat_id
      integer(c_int64_t) :: res
    end function psb_hip_batch_footprint

    function psb_hip_batch_can_admit(batch, mat_id) result(res) &
         & bind(c, name='psb_hip_batch_can_admit')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: mat_id
      integer(c_int) :: res
    end function psb_hip_batch_can_admit

    function psb_hip_batch_submit(batch, mat_id, b, x0, eps, itmax, sys_id) result(res) &
         & bind(c, name='psb_hip_batch_submit')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch, b, x0
      integer(c_int), value :: mat_id, itmax
      real(c_double), value :: eps
      integer(c_int) :: sys_id
      integer(c_int) :: res
    end function psb_hip_batch_submit

    function psb_hip_batch_step(batch) result(res) bind(c, name='psb_hip_batch_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int) :: res
    end function psb_hip_batch_step

    function psb_hip_batch_poll(batch, wait_last, max_out, ids, iters, outer, relres, status, &
         & nout, nlive) result(res) bind(c, name='psb_hip_batch_poll')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), value :: wait_last, max_out
      integer(c_int) :: ids(*), iters(*), outer(*), status(*)
      real(c_double) :: relres(*)
      integer(c_int) :: nout, nlive
      integer(c_int) :: res
    end function psb_hip_batch_poll

    function psb_hip_batch_result(batch, sys_id, x) result(res) &
         & bind(c, name='psb_hip_batch_result')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch, x
      integer(c_int), value :: sys_id
      integer(c_int) :: res
    end function psb_hip_batch_result

    ! ---------------------------------------------------------------- BJAC / ILU(0)
    function psb_hip_ilu0_fact(n, irp, ja, val, lirp, lja, lval, uirp, uja, uval, d, &
         & bad_row) result(res) bind(c, name='psb_hip_ilu0_fact')
      import :: c_int, c_double
      integer(c_int), value :: n
      integer(c_int) :: irp(*), ja(*), lirp(*), lja(*), uirp(*), uja(*), bad_row
      real(c_double) :: val(*), lval(*), uval(*), d(*)
      integer(c_int) :: res
    end function psb_hip_ilu0_fact

    function psb_hip_ilu_alloc(ilu, n, lirp, lja, lval, uirp, uja, uval, d) result(res) &
         & bind(c, name='psb_hip_ilu_alloc')
      import :: c_int, c_ptr, c_double
      type(c_ptr) :: ilu
      integer(c_int), value :: n
      integer(c_int) :: lirp(*), lja(*), uirp(*), uja(*)
      real(c_double) :: lval(*), uval(*), d(*)
      integer(c_int) :: res
    end function psb_hip_ilu_alloc

    function psb_hip_ilu_free(ilu) result(res) bind(c, name='psb_hip_ilu_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: ilu
      integer(c_int) :: res
    end function psb_hip_ilu_free

    function psb_hip_ilu_apply(ilu, alpha, x, beta, y) result(res) &
         & bind(c, name='psb_hip_ilu_apply')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ilu, x, y
      real(c_double), value :: alpha, beta
      integer(c_int) :: res
    end function psb_hip_ilu_apply

    function psb_hip_hll_tune(dmat, x, y, reps, best, us) result(res) &
         & bind(c, name='psb_hip_hll_tune')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: dmat, x, y
      integer(c_int), value :: reps
      integer(c_int) :: best
      real(c_float) :: us(4)
      integer(c_int) :: res
    end function psb_hip_hll_tune
  end interface

end module psb_hip_cbind_mod
