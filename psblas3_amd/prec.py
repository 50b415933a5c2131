"""Preconditioners: NONE and DIAG (Jacobi), psb_dprec_type
(prec/psb_d_prec_type.f90, prec/psb_d_diagprec.f90,
prec/impl/psb_d_diagprec_impl.f90).

DIAG build: d(i) = 1/a_ii with a zero diagonal replaced by 1
(psb_d_diag_precbld, psb_d_diagprec_impl.f90:229-293), computed on the
device from the matrix's own diagonal.  Apply: y = alpha*d.*x + beta*y
through the vector's mlt (:116), i.e. psb_hip_mv_axybz.
"""
from . import _native as N
from .vect import psb_d_hip_vect, psb_d_vect_type


class psb_dprec_type:
    def __init__(self, ptype="DIAG"):
        self.ptype = ptype.upper()
        if self.ptype not in ("DIAG", "NONE", "JACOBI"):
            raise N.PsbError(N.UNSUPPORTED, "psb_precinit", f"preconditioner {ptype}")
        if self.ptype == "JACOBI":
            self.ptype = "DIAG"
        self.dv = None

    def build(self, a, desc_a=None, vmold=None):
        """psb_precbld(a, desc, prec, info, vmold=)."""
        if self.ptype == "NONE":
            return self
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
        else:
            N.call("psb_hip_mv_axybz", yv.get_nrows(), float(alpha), self.dv.dev, xv.dev,
                   float(beta), yv.dev)
        yv.set_dev()

    def free(self):
        if self.dv is not None:
            self.dv.free()
            self.dv = None


def _vmold_for(a):
    from .vect import psb_s_hip_vect
    import numpy as np
    return psb_s_hip_vect() if a._dtype == np.float32 else psb_d_hip_vect()


def psb_precinit(ptype):
    return psb_dprec_type(ptype)
