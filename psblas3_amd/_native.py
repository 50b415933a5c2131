"""ctypes binding of libpsb_hip.so (the C ABI declared in include/psb_hip.h).

The library is built in-tree by ``psblas3_amd.build`` (hipcc, gfx950) into
``psblas3_amd/lib/libpsb_hip.so``.  There is no fallback: if the library is
missing, every entry point raises ``PsbError``.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpsb_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "psb_hip.h")

# return codes and element types (include/psb_hip.h)
SUCCESS, UNSUPPORTED, INVALID, OUTOFMEM, RUNTIME, NOT_INIT, BUSY = 0, 1, 2, 3, 4, 5, 6
TYPE_INT, TYPE_FLOAT, TYPE_DOUBLE = 0, 1, 2
INS_OVERWRITE, INS_ADD = 0, 1

_ERRNAMES = {UNSUPPORTED: "unsupported", INVALID: "invalid argument", OUTOFMEM: "out of memory",
             RUNTIME: "HIP runtime error", NOT_INIT: "not initialised",
             BUSY: "no free slot or arena room"}


class PsbError(RuntimeError):
    """Raised for a non-zero return code (the psb_errpush + error-handler path
    of the reference, base/modules/psb_error_mod.F90:391-404)."""

    def __init__(self, info, name, detail=""):
        self.info = info
        msg = f"{name}: info={info} ({_ERRNAMES.get(info, 'error')})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


_lib = None

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
u32 = C.c_uint32
u64 = C.c_uint64
f64 = C.c_double
f32 = C.c_float
pvp = C.POINTER(C.c_void_p)
pi32 = C.POINTER(C.c_int)
pi64 = C.POINTER(C.c_int64)
pf64 = C.POINTER(C.c_double)

# name -> (restype, argtypes)
_SIGS = {
    "psb_hip_init": (i32, [i32]),
    "psb_hip_exit": (i32, []),
    "psb_hip_get_device_count": (i32, []),
    "psb_hip_get_device": (i32, []),
    "psb_hip_set_device": (i32, [i32]),
    "psb_hip_get_multiprocessors": (i32, []),
    "psb_hip_get_memory_bus_width": (i32, []),
    "psb_hip_get_memory_clock_rate": (i32, []),
    "psb_hip_get_warp_size": (i32, []),
    "psb_hip_get_max_threads_per_mp": (i32, []),
    "psb_hip_get_max_registers_per_block": (i32, []),
    "psb_hip_device_has_uva": (i32, []),
    "psb_hip_get_device_name": (i32, [C.c_char_p, i32]),
    "psb_hip_memory_peak_bandwidth": (f64, []),
    "psb_hip_device_sync": (i32, []),
    "psb_hip_get_stream": (vp, []),
    "psb_hip_last_error": (C.c_char_p, []),
    "psb_hip_timer_start": (i32, []),
    "psb_hip_timer_stop": (f64, []),
    "psb_hip_stream_copy": (i32, [vp, vp, i64]),
    "psb_hip_stream_triad": (i32, [vp, vp, vp, f64, i64]),
    "psb_hip_stream_read": (i32, [vp, i64, vp]),
    "psb_hip_stream_mix": (i32, [vp, i32, i64, vp]),
    "psb_hip_alloc_buffer": (i32, [pvp, i64]),
    "psb_hip_free_buffer": (i32, [vp]),
    "psb_hip_write_buffer": (i32, [vp, vp, i64]),
    "psb_hip_read_buffer": (i32, [vp, vp, i64]),
    "psb_hip_mv_alloc": (i32, [pvp, i32, i32, i32]),
    "psb_hip_mv_free": (None, [vp]),
    "psb_hip_mv_size": (i32, [vp]),
    "psb_hip_mv_count": (i32, [vp]),
    "psb_hip_mv_pitch": (i32, [vp]),
    "psb_hip_mv_data": (vp, [vp]),
    "psb_hip_mv_write": (i32, [vp, vp]),
    "psb_hip_mv_read": (i32, [vp, vp]),
    "psb_hip_mv_write_r2": (i32, [vp, vp, i32]),
    "psb_hip_mv_read_r2": (i32, [vp, vp, i32]),
    "psb_hip_mv_write_first": (i32, [vp, i32, i32, vp]),
    "psb_hip_mv_read_first": (i32, [vp, i32, i32, vp]),
    "psb_hip_mv_setscal_d": (i32, [f64, i32, i32, i32, vp]),
    "psb_hip_mv_setscal_s": (i32, [f32, i32, i32, i32, vp]),
    "psb_hip_mv_setscal_i": (i32, [i32, i32, i32, i32, vp]),
    "psb_hip_mv_geins": (i32, [i32, vp, vp, i32, i32, vp]),
    "psb_hip_mv_igath": (i32, [vp, i32, i32, i32, vp, i32, vp, i32]),
    "psb_hip_mv_iscat": (i32, [vp, i32, i32, i32, vp, i32, vp, i32, f64]),
    "psb_hip_mv_scal": (i32, [f64, vp]),
    "psb_hip_mv_dot": (i32, [pf64, i32, vp, vp]),
    "psb_hip_mv_nrm2": (i32, [pf64, i32, vp]),
    "psb_hip_mv_amax": (i32, [pf64, i32, vp]),
    "psb_hip_mv_asum": (i32, [pf64, i32, vp]),
    "psb_hip_mv_axpby": (i32, [i32, f64, vp, f64, vp]),
    "psb_hip_mv_axpbyz": (i32, [i32, f64, vp, f64, vp, vp]),
    "psb_hip_mv_upd_xyz": (i32, [i32, f64, f64, f64, f64, vp, vp, vp]),
    "psb_hip_mv_xyzw": (i32, [i32, f64, f64, f64, f64, f64, f64, vp, vp, vp, vp]),
    "psb_hip_mv_axy": (i32, [i32, f64, vp, vp]),
    "psb_hip_mv_axybz": (i32, [i32, f64, vp, vp, f64, vp]),
    "psb_hip_mv_abs": (i32, [i32, f64, vp, vp]),
    "psb_hip_mv_diag_inv": (i32, [i32, vp]),
    "psb_hip_fix_coo_d": (i32, [i32, i32, i64, vp, vp, vp, i32, pi64]),
    "psb_hip_hll_offsets": (i32, [i32, i64, vp, i32, vp, vp, vp, pi64]),
    "psb_hip_gen_pde3d": (i64, [i32, vp, vp, f64, i32, i32, vp, vp, vp, vp]),
    "psb_hip_gen_aniso2d": (i64, [i32, f64, vp, vp, vp]),
    "psb_hip_gen_powerlaw": (i64, [i32, i32, u64, vp, vp, vp]),
    "psb_hip_fill_uniform_host": (None, [u64, u32, i64, vp]),
    "psb_hip_mv_fill_uniform": (i32, [vp, u64, u32]),
    "psb_hip_hll_alloc": (i32, [pvp, i32, i32, i32, i32, i32, i32]),
    "psb_hip_hll_free": (None, [vp]),
    "psb_hip_hll_write": (i32, [vp, vp, vp, vp, vp, vp]),
    "psb_hip_hll_read": (i32, [vp, vp, vp, vp, vp, vp]),
    "psb_hip_hll_get_params": (i32, [vp, pi32, pi32, pi32, pi32, pi32, pi32, pi32]),
    "psb_hip_hll_copy_coo": (i32, [i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp]),
    "psb_hip_hll_sort_rows": (i32, [i32, i32, i32, vp, vp, vp, vp, pi64]),
    "psb_hip_hll_copy_coo_sorted": (i32, [i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp,
                                          vp]),
    "psb_hip_hll_read_perm": (i32, [vp, vp]),
    "psb_hip_hll_units": (i32, [vp, pi32, pi32]),
    "psb_hip_hll_set_perm": (i32, [vp, vp, vp]),
    "psb_hip_host_hll_spmv": (i32, [i32, i32, vp, vp, vp, vp, vp, i32, f64, vp, f64, vp]),
    "psb_hip_host_ell_spmv": (i32, [i32, i32, vp, vp, vp, i32, f64, vp, f64, vp]),
    "psb_hip_host_csr_spmv": (i32, [i32, vp, vp, vp, i32, f64, vp, f64, vp]),
    "psb_hip_host_hll_fill": (i32, [i32, i32, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp]),
    "psb_hip_hll_spmv": (i32, [vp, f64, vp, f64, vp]),
    "psb_hip_hll_make_single": (i32, [vp]),
    "psb_hip_hll_spmv_single": (i32, [vp, f64, vp, f64, vp]),
    "psb_hip_ell_index16": (i32, [vp]),
    "psb_hip_hll_index16": (i32, [vp]),
    "psb_hip_hll_set_index16": (i32, [i32]),
    "psb_hip_hll_set_slices": (i32, [i32]),
    "psb_hip_hll_slices": (i32, [vp]),
    "psb_hip_hll_set_plane": (i32, [vp, i32, i32]),
    "psb_hip_hll_set_variant": (i32, [i32]),
    "psb_hip_hll_tune": (i32, [vp, vp, vp, i32, pi32, C.POINTER(f32)]),
    "psb_hip_hll_get_diag": (i32, [vp, vp]),
    "psb_hip_ell_alloc": (i32, [pvp, i32, i32, i32, i32, i32]),
    "psb_hip_ell_free": (None, [vp]),
    "psb_hip_ell_pitch": (i32, [vp]),
    "psb_hip_ell_max_row_size": (i32, [vp]),
    "psb_hip_ell_write": (i32, [vp, vp, vp, i32, vp, vp]),
    "psb_hip_ell_read": (i32, [vp, vp, vp, i32, vp, vp]),
    "psb_hip_ell_copy_coo": (i32, [i32, i32, i32, vp, vp, vp, vp, vp]),
    "psb_hip_ell_spmv": (i32, [vp, f64, vp, f64, vp]),
    "psb_hip_ell_get_diag": (i32, [vp, vp]),
    "psb_hip_csr_alloc": (i32, [pvp, i32, i32, i32, i32, i32]),
    "psb_hip_csr_free": (None, [vp]),
    "psb_hip_csr_write": (i32, [vp, vp, vp, vp]),
    "psb_hip_csr_read": (i32, [vp, vp, vp, vp]),
    "psb_hip_csr_spmv": (i32, [vp, f64, vp, f64, vp]),
    "psb_hip_csr_get_analysis": (i32, [vp, pi32, pi32, pi32]),
    "psb_hip_csr_set_variant": (i32, [i32]),
    "psb_hip_csr_spsv": (i32, [vp, i32, i32, i32, f64, vp, f64, vp]),
    "psb_hip_csr_set_slices": (i32, [i32]),
    "psb_hip_tmat_alloc": (i32, [pvp, i32, i32, vp, vp, vp, i32]),
    "psb_hip_tmat_free": (None, [vp]),
    "psb_hip_tmat_spmv": (i32, [vp, f64, vp, f64, vp]),
    "psb_hip_csr_slices": (i32, [vp]),
    "psb_hip_hll_scals": (i32, [vp, f64]),
    "psb_hip_hll_scal": (i32, [vp, vp, i32]),
    "psb_hip_ell_scals": (i32, [vp, f64]),
    "psb_hip_ell_scal": (i32, [vp, vp, i32]),
    "psb_hip_csr_scals": (i32, [vp, f64]),
    "psb_hip_csr_scal": (i32, [vp, vp, i32]),
    "psb_hip_csr_spsv_levels": (i32, [vp, pi32, pi32]),
    "psb_hip_hll_set_xcd_chunk": (i32, [i32]),
    "psb_hip_csr_get_diag": (i32, [vp, vp]),
    "psb_hip_krylov_fused": (i32, [i32, i32, vp, vp, vp, vp, f64, i32, i32, i32, pi32, pf64, vp]),
    "psb_hip_krylov_fused_noalloc": (i32, [i32, i32, vp, vp, vp, vp, f64, i32, i32, i32, pi32, pf64,
                                           vp]),
    "psb_hip_krylov_mixed": (i32, [i32, vp, vp, vp, vp, f64, i32, f64, i32, i32, i32, pi32,
                                   pi32, pf64]),
    "psb_hip_mm_read_coo": (i32, [C.c_char_p, i64, pi32, pi32, pi64, vp, vp, vp]),
    "psb_hip_mm_read_vec": (i32, [C.c_char_p, i64, pi64, vp]),
    "psb_hip_mm_write_coo": (i32, [C.c_char_p, C.c_char_p, i32, i32, i64, vp, vp, vp]),
    "psb_hip_mm_write_vec": (i32, [C.c_char_p, C.c_char_p, i64, vp]),
    "psb_hip_ilu0_fact": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, pi32]),
    "psb_hip_ilu0_fact_s": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, pi32]),
    "psb_hip_ilu_alloc": (i32, [pvp, i32, vp, vp, vp, vp, vp, vp, vp]),
    "psb_hip_ilu_free": (i32, [vp]),
    "psb_hip_ilu_levels": (i32, [vp, pi32, pi32]),
    "psb_hip_ilu_segments": (i32, [vp, pi32, pi32]),
    "psb_hip_ilu_set_variant": (i32, [i32]),
    "psb_hip_ilu_apply": (i32, [vp, f64, vp, f64, vp]),
    "psb_hip_batch_create": (i32, [pvp, i32, i32, i64, i32, f64, f64, i32]),
    "psb_hip_batch_free": (i32, [vp]),
    "psb_hip_batch_register": (i32, [vp, vp, vp, pi32]),
    "psb_hip_batch_footprint": (i64, [vp, i32]),
    "psb_hip_batch_submit": (i32, [vp, i32, vp, vp, f64, i32, pi32]),
    "psb_hip_batch_can_admit": (i32, [vp, i32]),
    "psb_hip_batch_step": (i32, [vp]),
    "psb_hip_batch_poll": (i32, [vp, i32, i32, pi32, pi32, pi32, pf64, pi32, pi32, pi32]),
    "psb_hip_batch_result": (i32, [vp, i32, vp]),
}


def exported_names():
    return list(_SIGS)


def load():
    """Load libpsb_hip.so (no device work).  Raises PsbError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PsbError(NOT_INIT, "psblas3_amd",
                       f"{LIB_PATH} not built; run psblas3_amd.build.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return _lib


def call(name, *args):
    """Call an int-returning entry point; raise PsbError on a non-zero code."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != SUCCESS:
        detail = lib.psb_hip_last_error()
        raise PsbError(rc, name, detail.decode() if detail else "")
    return rc


def fn(name):
    return getattr(load(), name)
