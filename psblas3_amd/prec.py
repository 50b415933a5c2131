"""Preconditioners: NONE, DIAG (Jacobi) and BJAC (block Jacobi with ILU(0);
one block at one process), psb_dprec_type (prec/psb_d_prec_type.f90,
prec/psb_d_diagprec.f90, prec/psb_d_bjacprec.f90).

DIAG build: d(i) = 1/a_ii with a zero diagonal replaced by 1
(psb_d_diag_precbld, psb_d_diagprec_impl.f90:229-293), computed on the
device from the matrix's own diagonal.  Apply: y = alpha*d.*x + beta*y
through the vector's mlt (:116), i.e. psb_hip_mv_axybz.

BJAC build (psb_d_bjac_precbld, psb_d_bjacprec_impl.f90:490-700, ILU(0),
the default factorization): the matrix is read back as host CSR and factored
by psb_hip_ilu0_fact (psb_dilu0_fact); the factors go to the device levelled
for the triangular solves.  Apply (psb_d_bjac_apply_vect, :166-173):
y = alpha * U^-1 (d .* L^-1 x) + beta * y on the device (psb_hip_ilu_apply).
"""
import ctypes as C

import numpy as np

from . import _native as N
from .vect import psb_d_hip_vect, psb_d_vect_type


class psb_dprec_type:
    def __init__(self, ptype="DIAG"):
        self.ptype = ptype.upper()
        if self.ptype not in ("DIAG", "NONE", "JACOBI", "BJAC"):
            raise N.PsbError(N.UNSUPPORTED, "psb_precinit", f"preconditioner {ptype}")
        if self.ptype == "JACOBI":
            self.ptype = "DIAG"
        self.dv = None
        self.ilu = None
        self._tri = None
        self.levels = None
        self.segments = None

    def build(self, a, desc_a=None, vmold=None):
        """psb_precbld(a, desc, prec, info, vmold=)."""
        if self.ptype == "NONE":
            return self
        if self.ptype == "BJAC":
            return self._build_ilu0(a)
        mold = vmold or (a.a and _vmold_for(a.a))
        d = mold.mold()
        d.all(a.get_ncols())
        a.a.get_diag(d)
        N.call("psb_hip_mv_diag_inv", a.get_nrows(), d.dev)
        d.set_dev()
        self.dv = d
        return self

    def apply(self, x, y, desc_a=None, trans="N", alpha=1.0, beta=0.0):
        """prec%apply(x, y, desc, info): y = alpha*M^{-1}x + beta*y."""
        xv = x.v if isinstance(x, psb_d_vect_type) else x
        yv = y.v if isinstance(y, psb_d_vect_type) else y
        xv.ensure_dev()
        if yv.is_host():
            yv.sync()
        if self.ptype == "NONE":
            # psb_d_null_prec apply: y = alpha*x + beta*y
            N.call("psb_hip_mv_axpby", yv.get_nrows(), float(alpha), xv.dev, float(beta), yv.dev)
        elif self.ptype == "BJAC":
            t = str(trans).upper()
            if t not in ("N", "T", "C"):
                raise N.PsbError(N.INVALID, "psb_bjac_apply", f"trans {trans!r}")
            if t == "N" and self.ilu is not None:
                N.call("psb_hip_ilu_apply", self.ilu, float(alpha), xv.dev, float(beta), yv.dev)
            else:
                self._apply_bjac_csr(xv, yv, alpha, beta, t)
        else:
            N.call("psb_hip_mv_axybz", yv.get_nrows(), float(alpha), self.dv.dev, xv.dev,
                   float(beta), yv.dev)
        yv.set_dev()

    def _apply_bjac_csr(self, xv, yv, alpha, beta, trans):
        """psb_d_bjac_apply_vect / psb_s_bjac_apply_vect (prec/impl/
        psb_d_bjacprec_impl.f90) through the factors as unit-triangular CSR
        matrices solved on the device (built on first use): op 'N'
        w = d .* (L^-1 x) (psb_spsm scale 'L'), y = alpha U^-1 w + beta y; op
        'T' / 'C' w = d .* (U^-T x), y = alpha L^-T w + beta y.  FP32 factors
        always go this way; FP64 ones for the transposed ops."""
        from .mat import (psb_d_coo_sparse_mat, psb_d_hip_csrg_sparse_mat,
                          psb_s_hip_csrg_sparse_mat)
        from .vect import psb_s_hip_vect
        if self._tri is None:
            n, lirp, lja, lval, uirp, uja, uval, d = self._factors
            single = d.dtype == np.float32
            mtype = psb_s_hip_csrg_sparse_mat if single else psb_d_hip_csrg_sparse_mat
            vtype = psb_s_hip_vect if single else psb_d_hip_vect
            tri = []
            for irp, ja, val, lower in ((lirp, lja, lval, True), (uirp, uja, uval, False)):
                ia = np.repeat(np.arange(1, n + 1, dtype=np.int32), np.diff(irp))
                m = mtype().cp_from_coo(psb_d_coo_sparse_mat(n, n, ia, ja, val).fix())
                m.set_triangle()
                m.set_lower(lower)
                m.set_unit()
                tri.append(m)
            self._tri = (tri[0], tri[1], vtype().bld(d[:n]), vtype().all(n))
        lt, ut, dv, w = self._tri
        first, second = (lt, ut) if trans == "N" else (ut, lt)
        first.inner_vect_sv(1.0, xv, 0.0, w, trans=trans)
        w.mlt_v(dv)
        second.inner_vect_sv(alpha, w, beta, yv, trans=trans)

    def _build_ilu0(self, a):
        inner = a.a if hasattr(a, "a") else a
        single = inner._dtype == np.float32
        dt = np.float32 if single else np.float64
        irp, ja, val = inner.to_csr()
        val = np.ascontiguousarray(val, dt)
        n, nz = len(irp) - 1, max(len(val), 1)
        lirp = np.zeros(n + 1, np.int32)
        uirp = np.zeros(n + 1, np.int32)
        lja, uja = np.zeros(nz, np.int32), np.zeros(nz, np.int32)
        lval, uval, d = np.zeros(nz, dt), np.zeros(nz, dt), np.zeros(max(n, 1), dt)
        bad = C.c_int(0)
        rc = N.fn("psb_hip_ilu0_fact_s" if single else "psb_hip_ilu0_fact")(
            n, irp.ctypes.data, ja.ctypes.data, val.ctypes.data, lirp.ctypes.data,
            lja.ctypes.data, lval.ctypes.data, uirp.ctypes.data, uja.ctypes.data,
            uval.ctypes.data, d.ctypes.data, C.byref(bad))
        if rc != N.SUCCESS:
            raise N.PsbError(rc, "psb_silu0_fact" if single else "psb_dilu0_fact",
                             f"pivot of row {bad.value} too small")
        if single:   # FP32 factors: applied through the CSR solves
            nl, nu = int(lirp[n]) - 1, int(uirp[n]) - 1
            self._factors = (n, lirp, lja[:nl].copy(), lval[:nl].copy(), uirp,
                             uja[:nu].copy(), uval[:nu].copy(), d)
            self._tri = None
            self.levels = self.segments = None
            return self
        h = C.c_void_p()
        N.call("psb_hip_ilu_alloc", C.byref(h), n, lirp.ctypes.data, lja.ctypes.data,
               lval.ctypes.data, uirp.ctypes.data, uja.ctypes.data, uval.ctypes.data,
               d.ctypes.data)
        self.ilu = h
        nl, nu = int(lirp[n]) - 1, int(uirp[n]) - 1
        self._factors = (n, lirp, lja[:nl].copy(), lval[:nl].copy(), uirp, uja[:nu].copy(),
                         uval[:nu].copy(), d)
        self._tri = None
        lo, up = C.c_int(0), C.c_int(0)
        N.call("psb_hip_ilu_levels", h, C.byref(lo), C.byref(up))
        self.levels = (lo.value, up.value)
        N.call("psb_hip_ilu_segments", h, C.byref(lo), C.byref(up))
        self.segments = (lo.value, up.value)
        return self

    def free(self):
        if self.dv is not None:
            self.dv.free()
            self.dv = None
        if self.ilu is not None:
            N.call("psb_hip_ilu_free", self.ilu)
            self.ilu = None
        if getattr(self, "_tri", None) is not None:
            for obj in self._tri:
                obj.free()
            self._tri = None


def _vmold_for(a):
    from .vect import psb_s_hip_vect
    import numpy as np
    return psb_s_hip_vect() if a._dtype == np.float32 else psb_d_hip_vect()


def psb_precinit(ptype):
    return psb_dprec_type(ptype)
