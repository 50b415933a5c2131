"""numpy/ctypes front end of the CPU oracle (oracle/psb_oracle.c).

TEST INFRASTRUCTURE ONLY.  This module is the checker: tests/, the smoke
check and bench.py's cpu_baseline leg may use it; the product package
(psblas3_amd) never imports it.  Each wrapper names the reference routine the
C restatement follows (file:line citations live in psb_oracle.c).

All index arrays are int32, 1-based (PSBLAS convention).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    L.orc_uniform01.restype = C.c_double
    L.orc_uniform01.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64]
    L.orc_fill_uniform_pm1.argtypes = [C.c_uint64, C.c_uint32, C.c_int64, _f64p]
    L.orc_gen_pde3d.restype = C.c_int64
    L.orc_gen_pde3d.argtypes = [C.c_int, _f64p, _f64p, C.c_double, C.c_int, C.c_int,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_gen_aniso2d.restype = C.c_int64
    L.orc_gen_aniso2d.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_gen_powerlaw.restype = C.c_int64
    L.orc_gen_powerlaw.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_fix_coo.restype = C.c_int64
    L.orc_fix_coo.argtypes = [C.c_int, C.c_int, C.c_int64, _i32p, _i32p, _f64p, C.c_int]
    L.orc_row_counts.argtypes = [C.c_int, C.c_int64, _i32p, _i32p]
    L.orc_csr_from_coo.argtypes = [C.c_int, C.c_int64, _i32p, _i32p]
    L.orc_hll_offsets.restype = C.c_int64
    L.orc_hll_offsets.argtypes = [C.c_int, C.c_int, _i32p, _i32p]
    L.orc_hll_fill.argtypes = [C.c_int, C.c_int64, _i32p, _i32p, _f64p, C.c_int, _i32p, _i32p,
                               _i32p, _f64p, _i32p]
    L.orc_ell_nzm.restype = C.c_int
    L.orc_ell_nzm.argtypes = [C.c_int, _i32p]
    L.orc_ell_fill.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, C.c_int, _i32p, _i32p,
                               _f64p, _i32p]
    for name, fp, ct in (("orc_d_csr_csmv", _f64p, C.c_double), ("orc_s_csr_csmv", _f32p, C.c_float)):
        f = getattr(L, name)
        f.argtypes = [C.c_int, _i32p, _i32p, fp, ct, fp, ct, fp]
    for name, fp, ct in (("orc_d_ell_csmv", _f64p, C.c_double), ("orc_s_ell_csmv", _f32p, C.c_float)):
        f = getattr(L, name)
        f.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, fp, ct, fp, ct, fp]
    for name, fp, ct in (("orc_d_hll_csmv", _f64p, C.c_double), ("orc_s_hll_csmv", _f32p, C.c_float)):
        f = getattr(L, name)
        f.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, fp, ct, fp, ct, fp, C.c_int]
    L.orc_d_hll_csmm.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _f64p, C.c_int,
                                 C.c_double, _f64p, C.c_int, C.c_double, _f64p, C.c_int]
    L.orc_d_axpby.argtypes = [C.c_int, C.c_double, _f64p, C.c_double, _f64p]
    L.orc_s_axpby.argtypes = [C.c_int, C.c_float, _f32p, C.c_float, _f32p]
    L.orc_d_upd_xyz.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                _f64p, _f64p, _f64p]
    L.orc_d_xyzw.argtypes = [C.c_int] + [C.c_double] * 6 + [_f64p] * 4
    L.orc_d_mlt.argtypes = [C.c_int, C.c_double, _f64p, _f64p, C.c_double, _f64p]
    L.orc_d_dot.restype = C.c_double
    L.orc_d_dot.argtypes = [C.c_int, _f64p, _f64p]
    L.orc_s_dot.restype = C.c_float
    L.orc_s_dot.argtypes = [C.c_int, _f32p, _f32p]
    L.orc_d_nrm2.restype = C.c_double
    L.orc_d_nrm2.argtypes = [C.c_int, _f64p]
    L.orc_s_nrm2.restype = C.c_float
    L.orc_s_nrm2.argtypes = [C.c_int, _f32p]
    L.orc_d_amax.restype = C.c_double
    L.orc_d_amax.argtypes = [C.c_int, _f64p]
    L.orc_d_asum.restype = C.c_double
    L.orc_d_asum.argtypes = [C.c_int, _f64p]
    L.orc_d_diag_build.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p]
    L.orc_d_krylov_csr.restype = C.c_int
    L.orc_d_krylov_csr.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                   C.c_double, C.c_int, C.c_int, C.c_double,
                                   C.POINTER(C.c_int), C.POINTER(C.c_double), _f64p]
    L.orc_d_krylov_hll.restype = C.c_int
    L.orc_d_krylov_hll.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i32p, _f64p,
                                   _f64p, _f64p, _f64p, C.c_double, C.c_int, C.c_int, C.c_double,
                                   C.POINTER(C.c_int), C.POINTER(C.c_double), _f64p]
    L.orc_d_krylov2_csr.restype = C.c_int
    L.orc_d_krylov2_csr.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                    _f64p, C.c_double, C.c_int, C.c_int, C.c_double,
                                    C.POINTER(C.c_int), C.POINTER(C.c_double), _f64p]
    L.orc_dilu0_fact.restype = C.c_int
    L.orc_dilu0_fact.argtypes = [C.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p, _i32p, _i32p,
                                 _f64p, _f64p]
    L.orc_silu0_fact.restype = C.c_int
    L.orc_silu0_fact.argtypes = [C.c_int, _i32p, _i32p, _f32p, _i32p, _i32p, _f32p, _i32p, _i32p,
                                 _f32p, _f32p]
    L.orc_d_csr_trsv_unit.restype = None
    L.orc_d_csr_trsv_unit.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
    for name, fp, ct in (("orc_d_csr_csmv_t", _f64p, C.c_double), ("orc_s_csr_csmv_t", _f32p, C.c_float)):
        getattr(L, name).restype = None
        getattr(L, name).argtypes = [C.c_int, C.c_int, _i32p, _i32p, fp, ct, fp, ct, fp]
    for name, fp, ct in (("orc_d_csr_cssv", _f64p, C.c_double), ("orc_s_csr_cssv", _f32p, C.c_float)):
        getattr(L, name).restype = None
        getattr(L, name).argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, ct, _i32p, _i32p, fp, fp,
                                     ct, fp]
    L.orc_d_bjac_apply.restype = None
    L.orc_d_bjac_apply.argtypes = [C.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p, _f64p, _f64p,
                                   _f64p]
    L.orc_d_krylov_ilu_csr.restype = C.c_int
    L.orc_d_krylov_ilu_csr.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, _i32p,
                                       _i32p, _f64p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                       C.c_double, C.c_int, C.c_int, C.c_double,
                                       C.POINTER(C.c_int), C.POINTER(C.c_double), _f64p]
    L.orc_num_threads.restype = C.c_int


# ---------------------------------------------------------------- rng
def uniform_pm1(seed, n, stream=0):
    out = np.empty(n, dtype=np.float64)
    lib().orc_fill_uniform_pm1(seed, stream, n, out)
    return out


# ---------------------------------------------------------------- generators
class COO:
    """Row/column/value triplets, 1-based (psb_d_coo_sparse_mat)."""

    def __init__(self, nr, nc, ia, ja, val):
        self.nr, self.nc = int(nr), int(nc)
        self.ia, self.ja, self.val = ia, ja, val

    @property
    def nz(self):
        return len(self.val)

    def copy(self):
        return COO(self.nr, self.nc, self.ia.copy(), self.ja.copy(), self.val.copy())


def gen_pde3d(L, a=(1.0, 1.0, 1.0), b=(0.0, 0.0, 0.0), c=0.0, hscale=True, gflag=False):
    """7-point operator in the reference assembly order (psb_d_pde3d.F90:478-560).
    Returns (COO, rhs) with rhs the g-boundary right-hand side (zero if not gflag)."""
    av = np.ascontiguousarray(a, dtype=np.float64)
    bv = np.ascontiguousarray(b, dtype=np.float64)
    nz = lib().orc_gen_pde3d(L, av, bv, c, int(hscale), int(gflag), None, None, None, None)
    n = L ** 3
    ia = np.empty(nz, np.int32); ja = np.empty(nz, np.int32); val = np.empty(nz)
    rhs = np.zeros(n)
    lib().orc_gen_pde3d(L, av, bv, c, int(hscale), int(gflag), ia.ctypes.data, ja.ctypes.data,
                        val.ctypes.data, rhs.ctypes.data)
    return COO(n, n, ia, ja, val), rhs


def gen_convdiff(L, conv=1.0 / np.sqrt(3.0)):
    """Convection-diffusion: a_i = 1, b_i = conv, h = 1/(L+1) (the stencil of the
    L=40 CPU-runnable case and of the L=200 sweep)."""
    return gen_pde3d(L, (1.0, 1.0, 1.0), (conv, conv, conv), 0.0, True, False)[0]


def gen_poisson3d(L):
    """SPD Poisson, diagonal 6 and off-diagonals -1."""
    return gen_pde3d(L, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), 0.0, False, False)[0]


def gen_aniso2d(L, eps):
    nz = lib().orc_gen_aniso2d(L, eps, None, None, None)
    ia = np.empty(nz, np.int32); ja = np.empty(nz, np.int32); val = np.empty(nz)
    lib().orc_gen_aniso2d(L, eps, ia.ctypes.data, ja.ctypes.data, val.ctypes.data)
    return COO(L * L, L * L, ia, ja, val)


def gen_powerlaw(n, cap=4095, seed=1):
    nz = lib().orc_gen_powerlaw(n, cap, seed, None, None, None)
    ia = np.empty(nz, np.int32); ja = np.empty(nz, np.int32); val = np.empty(nz)
    lib().orc_gen_powerlaw(n, cap, seed, ia.ctypes.data, ja.ctypes.data, val.ctypes.data)
    return COO(n, n, ia, ja, val)


# ---------------------------------------------------------------- conversions
def fix_coo(coo, dupl=1):
    """psb_d_fix_coo (row-major, dupl=add by default). Returns a new sorted COO."""
    c = coo.copy()
    nz = lib().orc_fix_coo(c.nr, c.nc, c.nz, c.ia, c.ja, c.val, dupl)
    if nz < 0:
        raise ValueError("duplicate entry with dupl=err")
    return COO(c.nr, c.nc, c.ia[:nz].copy(), c.ja[:nz].copy(), c.val[:nz].copy())


def row_counts(coo):
    irn = np.zeros(coo.nr, np.int32)
    lib().orc_row_counts(coo.nr, coo.nz, coo.ia, irn)
    return irn


def csr_from_coo(coo):
    """psb_d_cp_csr_from_coo on a fixed COO: returns (irp, ja, val), 1-based."""
    irp = np.empty(coo.nr + 1, np.int32)
    lib().orc_csr_from_coo(coo.nr, coo.nz, coo.ia, irp)
    return irp, coo.ja.copy(), coo.val.copy()


class HLL:
    def __init__(self, nr, nc, hksz, hkoffs, irn, ja, val, idiag):
        self.nr, self.nc, self.hksz = nr, nc, hksz
        self.hkoffs, self.irn, self.ja, self.val, self.idiag = hkoffs, irn, ja, val, idiag


def hll_from_coo(coo, hksz=32):
    """psi_d_convert_hll_from_coo on a fixed COO.  idiag is 0 where a row has
    no diagonal (the reference leaves those entries unset)."""
    irn = row_counts(coo)
    noffs = (coo.nr + hksz - 1) // hksz
    hkoffs = np.empty(noffs + 1, np.int32)
    isz = lib().orc_hll_offsets(coo.nr, hksz, irn, hkoffs)
    ja = np.zeros(max(isz, 1), np.int32)
    val = np.zeros(max(isz, 1))
    idiag = np.zeros(coo.nr, np.int32)
    if coo.nz > 0:
        lib().orc_hll_fill(coo.nr, coo.nz, coo.ia, coo.ja, coo.val, hksz, hkoffs, irn, ja, val,
                           idiag)
    return HLL(coo.nr, coo.nc, hksz, hkoffs, irn, ja[:isz], val[:isz], idiag)


class ELL:
    def __init__(self, nr, nc, ldv, nzm, irn, ja, val, idiag):
        self.nr, self.nc, self.ldv, self.nzm = nr, nc, ldv, nzm
        self.irn, self.ja, self.val, self.idiag = irn, ja, val, idiag


def ell_from_coo(coo, pitch_mult=1):
    """psi_d_convert_ell_from_coo; ldv = ceil(nr/pitch_mult)*pitch_mult.
    Rows nr+1..ldv are zero (ja 0, val 0)."""
    irn = row_counts(coo)
    nzm = lib().orc_ell_nzm(coo.nr, irn)
    ldv = ((coo.nr + pitch_mult - 1) // pitch_mult) * pitch_mult
    ja = np.zeros(ldv * max(nzm, 1), np.int32)
    val = np.zeros(ldv * max(nzm, 1))
    idiag = np.zeros(coo.nr, np.int32)
    if coo.nz > 0:
        lib().orc_ell_fill(coo.nr, coo.ia, coo.ja, coo.val, ldv, nzm, irn, ja, val, idiag)
    return ELL(coo.nr, coo.nc, ldv, nzm, irn, ja[:ldv * nzm], val[:ldv * nzm], idiag)


# ---------------------------------------------------------------- SpMV
def csr_spmv(irp, ja, val, x, alpha=1.0, beta=0.0, y=None):
    nr = len(irp) - 1
    single = val.dtype == np.float32
    dt = np.float32 if single else np.float64
    y = np.zeros(nr, dt) if y is None else np.ascontiguousarray(y, dt).copy()
    f = lib().orc_s_csr_csmv if single else lib().orc_d_csr_csmv
    f(nr, irp, ja, np.ascontiguousarray(val, dt), alpha, np.ascontiguousarray(x, dt), beta, y)
    return y


def hll_spmv(h, x, alpha=1.0, beta=0.0, y=None, quirk_empty_hack=True, val=None):
    val = h.val if val is None else val
    single = val.dtype == np.float32
    dt = np.float32 if single else np.float64
    y = np.zeros(h.nr, dt) if y is None else np.ascontiguousarray(y, dt).copy()
    f = lib().orc_s_hll_csmv if single else lib().orc_d_hll_csmv
    f(h.nr, h.hksz, h.hkoffs, h.irn, h.ja, np.ascontiguousarray(val, dt), alpha,
      np.ascontiguousarray(x, dt), beta, y, int(quirk_empty_hack))
    return y


def ell_spmv(e, x, alpha=1.0, beta=0.0, y=None, val=None):
    val = e.val if val is None else val
    single = val.dtype == np.float32
    dt = np.float32 if single else np.float64
    y = np.zeros(e.nr, dt) if y is None else np.ascontiguousarray(y, dt).copy()
    f = lib().orc_s_ell_csmv if single else lib().orc_d_ell_csmv
    f(e.nr, e.ldv, e.nzm, e.ja, np.ascontiguousarray(val, dt), alpha,
      np.ascontiguousarray(x, dt), beta, y)
    return y


def hll_spmm(h, x, alpha=1.0, beta=0.0, y=None):
    """x, y: (n, k) arrays; handled column-major as the reference does."""
    k = x.shape[1]
    xf = np.asfortranarray(x, dtype=np.float64)
    yf = np.zeros((h.nr, k), order="F") if y is None else np.asfortranarray(y, dtype=np.float64).copy()
    xr = xf.ravel(order="F")
    yr = yf.ravel(order="F")
    lib().orc_d_hll_csmm(h.nr, h.hksz, h.hkoffs, h.irn, h.ja, h.val, k, alpha, xr, xf.shape[0],
                         beta, yr, yf.shape[0])
    return yr.reshape((h.nr, k), order="F")


# ---------------------------------------------------------------- BLAS-1
def axpby(alpha, x, beta, y):
    y = np.ascontiguousarray(y).copy()
    if y.dtype == np.float32:
        lib().orc_s_axpby(len(y), alpha, np.ascontiguousarray(x, np.float32), beta, y)
    else:
        lib().orc_d_axpby(len(y), alpha, np.ascontiguousarray(x, np.float64), beta, y)
    return y


def upd_xyz(alpha, beta, gamma, delta, x, y, z):
    y = y.copy(); z = z.copy()
    lib().orc_d_upd_xyz(len(y), alpha, beta, gamma, delta, np.ascontiguousarray(x), y, z)
    return y, z


def xyzw(a, b, c, d, e, f, x, y, z, w):
    y = y.copy(); z = z.copy(); w = w.copy()
    lib().orc_d_xyzw(len(y), a, b, c, d, e, f, np.ascontiguousarray(x), y, z, w)
    return y, z, w


def mlt(alpha, x, y, beta, z):
    z = z.copy()
    lib().orc_d_mlt(len(z), alpha, np.ascontiguousarray(x), np.ascontiguousarray(y), beta, z)
    return z


def dot(x, y):
    if x.dtype == np.float32:
        return float(lib().orc_s_dot(len(x), x, y))
    return float(lib().orc_d_dot(len(x), np.ascontiguousarray(x), np.ascontiguousarray(y)))


def nrm2(x):
    if x.dtype == np.float32:
        return float(lib().orc_s_nrm2(len(x), x))
    return float(lib().orc_d_nrm2(len(x), np.ascontiguousarray(x)))


def amax(x):
    return float(lib().orc_d_amax(len(x), np.ascontiguousarray(x, np.float64)))


def asum(x):
    return float(lib().orc_d_asum(len(x), np.ascontiguousarray(x, np.float64)))


def diag_build(irp, ja, val):
    d = np.empty(len(irp) - 1)
    lib().orc_d_diag_build(len(irp) - 1, irp, ja, val, d)
    return d


KRYLOV_METHODS = {"CG": 0, "BICGSTAB": 1, "FCG": 2, "GCR": 3, "RGMRES": 4, "GMRES": 4,
                  "BICGSTABL": 5, "CGS": 6, "BICG": 7}
# default restart length / l when irst is absent (psb_dgcr.f90:207-216,
# psb_drgmres.f90:201-210, psb_dcgstabl.f90:190-199)
KRYLOV_IRST = {3: 10, 4: 10, 5: 1}


def krylov(method, a, b, x0=None, eps=1e-9, itmax=1000, istop=2, d=None, ani=0.0, irst=None):
    """psb_krylov with the DIAG preconditioner.  a: (irp, ja, val) CSR tuple or
    an HLL (CG / BiCGSTAB only).  method: CG, BICGSTAB, FCG, GCR, RGMRES,
    BICGSTABL, CGS, BICG (unknown names: BiCGSTAB, as the reference).  Returns (x,
    iter, err, hist)."""
    meth = KRYLOV_METHODS.get(method.upper(), 1)
    n = len(b)
    x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    if meth >= 2:
        irp, ja, val = a
        if d is None:
            d = diag_build(irp, ja, val)
        nl = KRYLOV_IRST.get(meth, 0) if irst is None else int(irst)
        hist = np.zeros(itmax + nl + 2)
        it = C.c_int(0)
        err = C.c_double(0)
        lib().orc_d_krylov2_csr(meth, nl, len(irp) - 1, irp, ja, val, d,
                                np.ascontiguousarray(b, np.float64), x, eps, itmax, istop, ani,
                                C.byref(it), C.byref(err), hist)
        return x, it.value, err.value, hist[1:it.value + 1].copy()
    hist = np.zeros(itmax + 2)
    it = C.c_int(0)
    err = C.c_double(0)
    b = np.ascontiguousarray(b, np.float64)
    if isinstance(a, HLL):
        if d is None:
            raise ValueError("pass the Jacobi diagonal d for an HLL operator")
        lib().orc_d_krylov_hll(meth, a.nr, a.hksz, a.hkoffs, a.irn, a.ja, a.val, d, b, x,
                               eps, itmax, istop, ani, C.byref(it), C.byref(err), hist)
    else:
        irp, ja, val = a
        if d is None:
            d = diag_build(irp, ja, val)
        lib().orc_d_krylov_csr(meth, len(irp) - 1, irp, ja, val, d, b, x, eps, itmax, istop, ani,
                               C.byref(it), C.byref(err), hist)
    return x, it.value, err.value, hist[1:it.value + 1].copy()


# ---------------------------------------------------------------- BJAC / ILU(0)
class ILU0:
    """ILU(0) factors (psb_dilu0_fact): L strict lower (unit), U strict upper
    scaled by the row's inverse pivot, d inverse pivots; CSR 1-based."""

    def __init__(self, n, lirp, lja, lval, uirp, uja, uval, d):
        self.n = n
        self.lirp, self.lja, self.lval = lirp, lja, lval
        self.uirp, self.uja, self.uval = uirp, uja, uval
        self.d = d


def ilu0_fact(irp, ja, val):
    """psb_[ds]ilu0_fact; the dtype of val picks the precision."""
    n = len(irp) - 1
    nz = len(val)
    dt = np.float32 if np.asarray(val).dtype == np.float32 else np.float64
    lirp = np.zeros(n + 1, np.int32); uirp = np.zeros(n + 1, np.int32)
    lja = np.zeros(max(nz, 1), np.int32); uja = np.zeros(max(nz, 1), np.int32)
    lval = np.zeros(max(nz, 1), dt); uval = np.zeros(max(nz, 1), dt); d = np.zeros(n, dt)
    fn = lib().orc_silu0_fact if dt == np.float32 else lib().orc_dilu0_fact
    rc = fn(n, np.ascontiguousarray(irp, np.int32), np.ascontiguousarray(ja, np.int32),
            np.ascontiguousarray(val, dt), lirp, lja, lval, uirp, uja, uval, d)
    if rc < 0:
        raise ZeroDivisionError(f"ILU(0) pivot of row {-rc} too small")
    nl, nu = lirp[n] - 1, uirp[n] - 1
    return ILU0(n, lirp, lja[:nl].copy(), lval[:nl].copy(), uirp, uja[:nu].copy(),
                uval[:nu].copy(), d)


def trsv_unit(irp, ja, val, x, lower=True):
    y = np.zeros(len(irp) - 1)
    lib().orc_d_csr_trsv_unit(len(irp) - 1, 1 if lower else 0, irp, ja, val,
                              np.ascontiguousarray(x, np.float64), y)
    return y


def csr_spmv_t(irp, ja, val, x, alpha=1.0, beta=0.0, y=None, nc=None):
    """psb_[ds]_csr_csmv with trans 'T': alpha * A^T x + beta * y, A with
    len(irp) - 1 rows and nc columns (default: the largest column index)."""
    single = np.asarray(val).dtype == np.float32
    dt = np.float32 if single else np.float64
    nr = len(irp) - 1
    nc = int(np.max(ja)) if nc is None and len(ja) else (nc or 0)
    out = np.zeros(nc, dt) if y is None else np.array(y, dt, copy=True)
    fn = lib().orc_s_csr_csmv_t if single else lib().orc_d_csr_csmv_t
    fn(nr, nc, np.ascontiguousarray(irp, np.int32), np.ascontiguousarray(ja, np.int32),
       np.ascontiguousarray(val, dt), alpha, np.ascontiguousarray(x, dt), beta, out)
    return out


def csr_cssv(irp, ja, val, x, lower=True, unit=False, trans="N", alpha=1.0, beta=0.0, y=None):
    """psb_[ds]_csr_cssv: alpha * T^-1 x + beta * y for a triangular CSR T
    (diagonal last in a lower row, first in an upper row, absent if unit);
    the dtype of val picks the precision."""
    single = np.asarray(val).dtype == np.float32
    dt = np.float32 if single else np.float64
    n = len(irp) - 1
    out = np.zeros(n, dt) if y is None else np.array(y, dt, copy=True)
    fn = lib().orc_s_csr_cssv if single else lib().orc_d_csr_cssv
    fn(n, 1 if lower else 0, 1 if unit else 0, 0 if trans.upper() == "N" else 1, alpha,
       np.ascontiguousarray(irp, np.int32), np.ascontiguousarray(ja, np.int32),
       np.ascontiguousarray(val, dt), np.ascontiguousarray(x, dt), beta, out)
    return out


def bjac_apply(f, x):
    """psb_d_bjac_apply_vect (ILU): y = U^-1 (d .* L^-1 x)."""
    y = np.zeros(f.n)
    lib().orc_d_bjac_apply(f.n, f.lirp, f.lja, f.lval, f.uirp, f.uja, f.uval, f.d,
                           np.ascontiguousarray(x, np.float64), y)
    return y


def krylov_bjac(method, csr, b, x0=None, eps=1e-9, itmax=1000, istop=2, irst=None, f=None):
    """psb_krylov with the BJAC / ILU(0) preconditioner (factors f, or
    computed here).  Returns (x, iter, err, hist)."""
    irp, ja, val = csr
    if f is None:
        f = ilu0_fact(irp, ja, val)
    meth = KRYLOV_METHODS.get(method.upper(), 1)
    n = len(b)
    x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    nl = KRYLOV_IRST.get(meth, 0) if irst is None else int(irst)
    hist = np.zeros(itmax + nl + 2)
    it = C.c_int(0)
    err = C.c_double(0)
    lib().orc_d_krylov_ilu_csr(meth, nl, n, irp, ja, val, f.lirp, f.lja, f.lval, f.uirp, f.uja,
                               f.uval, f.d, np.ascontiguousarray(b, np.float64), x, eps, itmax,
                               istop, 0.0, C.byref(it), C.byref(err), hist)
    return x, it.value, err.value, hist[1:it.value + 1].copy()


def num_threads():
    return lib().orc_num_threads()
