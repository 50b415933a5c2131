﻿!mod$ v1 sum:8419c3678168e143
!need$ 0bde2ac47243ead2 i iso_c_binding
module psb_hip_cbind_mod
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
integer(4),parameter::psb_hip_success_=0_4
integer(4),parameter::psb_hip_unsupported_=1_4
integer(4),parameter::psb_hip_invalid_=2_4
integer(4),parameter::psb_hip_outofmem_=3_4
integer(4),parameter::psb_hip_type_int_=0_4
integer(4),parameter::psb_hip_type_float_=1_4
integer(4),parameter::psb_hip_type_double_=2_4
integer(4),parameter::psb_hip_fmt_hll_=0_4
integer(4),parameter::psb_hip_fmt_ell_=1_4
integer(4),parameter::psb_hip_fmt_csr_=2_4
interface
function psb_hip_init(dev) bind(c,name="psb_hip_init") result(res)
integer(4),value::dev
integer(4)::res
end
end interface
interface
function psb_hip_exit() bind(c,name="psb_hip_exit") result(res)
integer(4)::res
end
end interface
interface
function psb_hip_device_sync() bind(c,name="psb_hip_device_sync") result(res)
integer(4)::res
end
end interface
interface
function psb_hip_get_multiprocessors() bind(c,name="psb_hip_get_multiprocessors") result(res)
integer(4)::res
end
end interface
interface
function psb_hip_memory_peak_bandwidth() bind(c,name="psb_hip_memory_peak_bandwidth") result(res)
real(8)::res
end
end interface
interface
function psb_hip_mv_alloc(mv,count,size,elem_type) bind(c,name="psb_hip_mv_alloc") result(res)
import::c_ptr
type(c_ptr)::mv
integer(4),value::count
integer(4),value::size
integer(4),value::elem_type
integer(4)::res
end
end interface
interface
subroutine psb_hip_mv_free(mv) bind(c,name="psb_hip_mv_free")
import::c_ptr
type(c_ptr),value::mv
end
end interface
interface
function psb_hip_mv_write(mv,host) bind(c,name="psb_hip_mv_write") result(res)
import::c_ptr
type(c_ptr),value::mv
type(c_ptr),value::host
integer(4)::res
end
end interface
interface
function psb_hip_mv_read(mv,host) bind(c,name="psb_hip_mv_read") result(res)
import::c_ptr
type(c_ptr),value::mv
type(c_ptr),value::host
integer(4)::res
end
end interface
interface
function psb_hip_mv_dot(res_,n,x,y) bind(c,name="psb_hip_mv_dot") result(res)
import::c_ptr
real(8)::res_(1_8:*)
integer(4),value::n
type(c_ptr),value::x
type(c_ptr),value::y
integer(4)::res
end
end interface
interface
function psb_hip_mv_nrm2(res_,n,x) bind(c,name="psb_hip_mv_nrm2") result(res)
import::c_ptr
real(8)::res_(1_8:*)
integer(4),value::n
type(c_ptr),value::x
integer(4)::res
end
end interface
interface
function psb_hip_mv_axpby(n,alpha,x,beta,y) bind(c,name="psb_hip_mv_axpby") result(res)
import::c_ptr
integer(4),value::n
real(8),value::alpha
type(c_ptr),value::x
real(8),value::beta
type(c_ptr),value::y
integer(4)::res
end
end interface
interface
function psb_hip_mv_axybz(n,alpha,x,y,beta,z) bind(c,name="psb_hip_mv_axybz") result(res)
import::c_ptr
integer(4),value::n
real(8),value::alpha
type(c_ptr),value::x
type(c_ptr),value::y
real(8),value::beta
type(c_ptr),value::z
integer(4)::res
end
end interface
interface
function psb_hip_mv_diag_inv(n,d) bind(c,name="psb_hip_mv_diag_inv") result(res)
import::c_ptr
integer(4),value::n
type(c_ptr),value::d
integer(4)::res
end
end interface
interface
function psb_hip_fix_coo_d(nr,nc,nz,ia,ja,val,dupl,nzout) bind(c,name="psb_hip_fix_coo_d") result(res)
integer(4),value::nr
integer(4),value::nc
integer(8),value::nz
integer(4)::ia(1_8:*)
integer(4)::ja(1_8:*)
real(8)::val(1_8:*)
integer(4),value::dupl
integer(8)::nzout
integer(4)::res
end
end interface
interface
function psb_hip_hll_offsets(nr,nz,ia,hksz,irn,idisp,hkoffs,isz) bind(c,name="psb_hip_hll_offsets") result(res)
integer(4),value::nr
integer(8),value::nz
integer(4)::ia(1_8:*)
integer(4),value::hksz
integer(4)::irn(1_8:*)
integer(4)::idisp(1_8:*)
integer(4)::hkoffs(1_8:*)
integer(8)::isz
integer(4)::res
end
end interface
interface
function psb_hip_hll_alloc(dmat,hksz,rows,nzeros,allocsize,elem_type,first_index) bind(c,name="psb_hip_hll_alloc") result(res)
import::c_ptr
type(c_ptr)::dmat
integer(4),value::hksz
integer(4),value::rows
integer(4),value::nzeros
integer(4),value::allocsize
integer(4),value::elem_type
integer(4),value::first_index
integer(4)::res
end
end interface
interface
subroutine psb_hip_hll_free(dmat) bind(c,name="psb_hip_hll_free")
import::c_ptr
type(c_ptr),value::dmat
end
end interface
interface
function psb_hip_hll_copy_coo(nr,nc,nza,hksz,noffs,isz,irn,hoffs,idisp,ja,val,dmat) bind(c,name="psb_hip_hll_copy_coo") result(res)
import::c_ptr
integer(4),value::nr
integer(4),value::nc
integer(4),value::nza
integer(4),value::hksz
integer(4),value::noffs
integer(4),value::isz
integer(4)::irn(1_8:*)
integer(4)::hoffs(1_8:*)
integer(4)::idisp(1_8:*)
integer(4)::ja(1_8:*)
real(8)::val(1_8:*)
type(c_ptr),value::dmat
integer(4)::res
end
end interface
interface
function psb_hip_hll_spmv(dmat,alpha,x,beta,y) bind(c,name="psb_hip_hll_spmv") result(res)
import::c_ptr
type(c_ptr),value::dmat
real(8),value::alpha
type(c_ptr),value::x
real(8),value::beta
type(c_ptr),value::y
integer(4)::res
end
end interface
interface
function psb_hip_hll_get_diag(dmat,d) bind(c,name="psb_hip_hll_get_diag") result(res)
import::c_ptr
type(c_ptr),value::dmat
type(c_ptr),value::d
integer(4)::res
end
end interface
interface
function psb_hip_ell_spmv(dmat,alpha,x,beta,y) bind(c,name="psb_hip_ell_spmv") result(res)
import::c_ptr
type(c_ptr),value::dmat
real(8),value::alpha
type(c_ptr),value::x
real(8),value::beta
type(c_ptr),value::y
integer(4)::res
end
end interface
interface
function psb_hip_csr_spmv(dmat,alpha,x,beta,y) bind(c,name="psb_hip_csr_spmv") result(res)
import::c_ptr
type(c_ptr),value::dmat
real(8),value::alpha
type(c_ptr),value::x
real(8),value::beta
type(c_ptr),value::y
integer(4)::res
end
end interface
interface
function psb_hip_krylov_fused(method,fmt,dmat,diag,b,x,eps,itmax,istop,check_every,iter,err,hist) bind(c,name="psb_hip_krylov_fused") result(res)
import::c_ptr
integer(4),value::method
integer(4),value::fmt
type(c_ptr),value::dmat
type(c_ptr),value::diag
type(c_ptr),value::b
type(c_ptr),value::x
real(8),value::eps
integer(4),value::itmax
integer(4),value::istop
integer(4),value::check_every
integer(4)::iter
real(8)::err
type(c_ptr),value::hist
integer(4)::res
end
end interface
interface
function psb_hip_krylov_mixed(method,dmat,diag,b,x,eps,outer_itmax,inner_eps,inner_itmax,check_every,variant,outer_its,inner_its,err) bind(c,name="psb_hip_krylov_mixed") result(res)
import::c_ptr
integer(4),value::method
type(c_ptr),value::dmat
type(c_ptr),value::diag
type(c_ptr),value::b
type(c_ptr),value::x
real(8),value::eps
integer(4),value::outer_itmax
real(8),value::inner_eps
integer(4),value::inner_itmax
integer(4),value::check_every
integer(4),value::variant
integer(4)::outer_its
integer(4)::inner_its
real(8)::err
integer(4)::res
end
end interface
interface
function psb_hip_batch_create(batch,method,max_systems,arena_bytes,check_every,delta,inner_eps,use_graph) bind(c,name="psb_hip_batch_create") result(res)
import::c_ptr
type(c_ptr)::batch
integer(4),value::method
integer(4),value::max_systems
integer(8),value::arena_bytes
integer(4),value::check_every
real(8),value::delta
real(8),value::inner_eps
integer(4),value::use_graph
integer(4)::res
end
end interface
interface
function psb_hip_batch_free(batch) bind(c,name="psb_hip_batch_free") result(res)
import::c_ptr
type(c_ptr),value::batch
integer(4)::res
end
end interface
interface
function psb_hip_batch_register(batch,dmat,diag,mat_id) bind(c,name="psb_hip_batch_register") result(res)
import::c_ptr
type(c_ptr),value::batch
type(c_ptr),value::dmat
type(c_ptr),value::diag
integer(4)::mat_id
integer(4)::res
end
end interface
interface
function psb_hip_batch_footprint(batch,mat_id) bind(c,name="psb_hip_batch_footprint") result(res)
import::c_ptr
type(c_ptr),value::batch
integer(4),value::mat_id
integer(8)::res
end
end interface
interface
function psb_hip_batch_can_admit(batch,mat_id) bind(c,name="psb_hip_batch_can_admit") result(res)
import::c_ptr
type(c_ptr),value::batch
integer(4),value::mat_id
integer(4)::res
end
end interface
interface
function psb_hip_batch_submit(batch,mat_id,b,x0,eps,itmax,sys_id) bind(c,name="psb_hip_batch_submit") result(res)
import::c_ptr
type(c_ptr),value::batch
integer(4),value::mat_id
type(c_ptr),value::b
type(c_ptr),value::x0
real(8),value::eps
integer(4),value::itmax
integer(4)::sys_id
integer(4)::res
end
end interface
interface
function psb_hip_batch_step(batch) bind(c,name="psb_hip_batch_step") result(res)
import::c_ptr
type(c_ptr),value::batch
integer(4)::res
end
end interface
interface
function psb_hip_batch_poll(batch,wait_last,max_out,ids,iters,outer,relres,status,nout,nlive) bind(c,name="psb_hip_batch_poll") result(res)
import::c_ptr
type(c_ptr),value::batch
integer(4),value::wait_last
integer(4),value::max_out
integer(4)::ids(1_8:*)
integer(4)::iters(1_8:*)
integer(4)::outer(1_8:*)
real(8)::relres(1_8:*)
integer(4)::status(1_8:*)
integer(4)::nout
integer(4)::nlive
integer(4)::res
end
end interface
interface
function psb_hip_batch_result(batch,sys_id,x) bind(c,name="psb_hip_batch_result") result(res)
import::c_ptr
type(c_ptr),value::batch
integer(4),value::sys_id
type(c_ptr),value::x
integer(4)::res
end
end interface
interface
function psb_hip_ilu0_fact(n,irp,ja,val,lirp,lja,lval,uirp,uja,uval,d,bad_row) bind(c,name="psb_hip_ilu0_fact") result(res)
integer(4),value::n
integer(4)::irp(1_8:*)
integer(4)::ja(1_8:*)
real(8)::val(1_8:*)
integer(4)::lirp(1_8:*)
integer(4)::lja(1_8:*)
real(8)::lval(1_8:*)
integer(4)::uirp(1_8:*)
integer(4)::uja(1_8:*)
real(8)::uval(1_8:*)
real(8)::d(1_8:*)
integer(4)::bad_row
integer(4)::res
end
end interface
interface
function psb_hip_ilu_alloc(ilu,n,lirp,lja,lval,uirp,uja,uval,d) bind(c,name="psb_hip_ilu_alloc") result(res)
import::c_ptr
type(c_ptr)::ilu
integer(4),value::n
integer(4)::lirp(1_8:*)
integer(4)::lja(1_8:*)
real(8)::lval(1_8:*)
integer(4)::uirp(1_8:*)
integer(4)::uja(1_8:*)
real(8)::uval(1_8:*)
real(8)::d(1_8:*)
integer(4)::res
end
end interface
interface
function psb_hip_ilu_free(ilu) bind(c,name="psb_hip_ilu_free") result(res)
import::c_ptr
type(c_ptr),value::ilu
integer(4)::res
end
end interface
interface
function psb_hip_ilu_apply(ilu,alpha,x,beta,y) bind(c,name="psb_hip_ilu_apply") result(res)
import::c_ptr
type(c_ptr),value::ilu
real(8),value::alpha
type(c_ptr),value::x
real(8),value::beta
type(c_ptr),value::y
integer(4)::res
end
end interface
interface
function psb_hip_hll_tune(dmat,x,y,reps,best,us) bind(c,name="psb_hip_hll_tune") result(res)
import::c_ptr
type(c_ptr),value::dmat
type(c_ptr),value::x
type(c_ptr),value::y
integer(4),value::reps
integer(4)::best
real(4)::us(1_8:4_8)
integer(4)::res
end
end interface
end
