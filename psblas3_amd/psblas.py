"""Parallel-BLAS entry points and assembly tools, single process (np = 1).

Mirrors base/psblas (psb_spmm, psb_geaxpby, psb_gedot, psb_genrm2, ...) and
base/tools (psb_cdall, psb_spall/spins/spasb, psb_geall/geasb): same names,
argument order and meaning.  With one process the descriptor is trivial,
halo exchanges are empty and the global reductions (psb_sum after gedot,
base/psblas/psb_ddot.f90:150) are identities.  Errors raise PsbError with
the reference's routine name.
"""
import numpy as np

from . import _native as N
from .mat import DUPL_ADD, _DUPL, psb_d_coo_sparse_mat, psb_d_hip_hlg_sparse_mat, psb_dspmat_type
from .vect import psb_d_hip_vect, psb_d_vect_type, psb_hip_base_vect


class psb_desc_type:
    """Communication descriptor at np = 1 (psb_desc_type, REPL/BLOCK
    trivial): local rows = local columns = global rows."""

    def __init__(self, n):
        self.n = int(n)

    def get_local_rows(self):
        return self.n

    def get_local_cols(self):
        return self.n

    def get_global_rows(self):
        return self.n

    def get_context(self):
        return None


def psb_cdall(ctxt=None, nl=None, n=None):
    """psb_cdall(ctxt, desc, info, nl=) for one process."""
    size = nl if nl is not None else n
    return psb_desc_type(size)


def psb_cdasb(desc):
    return desc


# ------------------------------------------------------------------ matrices
def psb_spall(desc, nnz=0, ncols=None):
    """psb_spall: a matrix in build state."""
    a = psb_dspmat_type()
    a.m = desc.get_local_rows()
    a.n = desc.get_local_cols() if ncols is None else ncols
    a.coo = psb_d_coo_sparse_mat(a.m, a.n)
    a._parts = []
    return a


def psb_spins(nz, ia, ja, val, a, desc=None):
    """psb_spins: append nz 1-based triplets (psb_d_coo_sparse_mat csput)."""
    if a.coo is None:
        raise N.PsbError(N.INVALID, "psb_spins", "matrix not in build state")
    ia = np.asarray(ia[:nz], np.int32)
    ja = np.asarray(ja[:nz], np.int32)
    if nz and (ia.min() < 1 or ia.max() > a.m or ja.min() < 1 or ja.max() > a.n):
        raise N.PsbError(N.INVALID, "psb_spins", "index out of range")
    a._parts.append((ia, ja, np.asarray(val[:nz], np.float64)))


def psb_spasb(a, desc=None, mold=None, afmt=None, dupl="add", keep_coo=False):
    """psb_spasb(a, desc, info, afmt=, mold=, dupl=): fix the COO and convert
    it into the mold's format (base/tools/psb_dspasb.f90:144, cscnv(mold=))."""
    if a.coo is None:
        raise N.PsbError(N.INVALID, "psb_spasb", "matrix not in build state")
    if a._parts:
        ia = np.concatenate([p[0] for p in a._parts])
        ja = np.concatenate([p[1] for p in a._parts])
        val = np.concatenate([p[2] for p in a._parts])
        a.coo = psb_d_coo_sparse_mat(a.m, a.n, ia, ja, val)
        a._parts = []
    a.coo.fix(_DUPL[dupl] if isinstance(dupl, str) else dupl)
    if mold is None:
        mold = _fmt_mold(afmt or "HLG")
    a.a = mold.mold()
    a.a.cp_from_coo(a.coo)
    a.a._triangle, a.a._upper, a.a._unit = a._tri
    if not keep_coo:
        a.coo = None
    return a


def psb_cscnv(a, mold):
    """a%cscnv(info, mold=): needs the COO kept at assembly."""
    if a.coo is None:
        raise N.PsbError(N.INVALID, "psb_cscnv", "assemble with keep_coo=True to convert")
    old = a.a
    a.a = mold.mold()
    a.a.cp_from_coo(a.coo)
    a.a._triangle, a.a._upper, a.a._unit = a._tri
    if old is not None:
        old.free()
    return a


def _fmt_mold(afmt):
    from . import mat
    table = {"HLG": mat.psb_d_hip_hlg_sparse_mat, "ELG": mat.psb_d_hip_elg_sparse_mat,
             "CSRG": mat.psb_d_hip_csrg_sparse_mat}
    if afmt.upper() not in table:
        raise N.PsbError(N.UNSUPPORTED, "psb_spasb", f"format {afmt}")
    return table[afmt.upper()]()


def psb_spfree(a, desc=None):
    a.free()


def matrix_from_coo(nr, nc, ia, ja, val, mold=None, dupl="add"):
    """Convenience: psb_spall + psb_spins + psb_spasb in one call."""
    desc = psb_desc_type(nr)
    a = psb_spall(desc, len(val), ncols=nc)
    psb_spins(len(val), ia, ja, val, a, desc)
    return psb_spasb(a, desc, mold=mold or psb_d_hip_hlg_sparse_mat(), dupl=dupl)


# ------------------------------------------------------------------ vectors
def psb_geall(desc, n=None, mold=None):
    """psb_geall + psb_geasb(mold=): a zero vector of the mold's type."""
    v = (mold or psb_d_hip_vect()).mold()
    v.all(desc.get_local_cols() if n is None else n)
    return psb_d_vect_type(v)


def psb_geasb(x, desc=None, mold=None):
    """psb_geasb(x, desc, info, mold=): move the values to the mold's storage."""
    if mold is not None and type(x.v) is not type(mold):
        vals = x.v.get_vect()
        nv = mold.mold()
        nv.bld(vals.astype(nv._dtype))
        x.v.free()
        x.v = nv
    return x


def vect_from(values, mold=None):
    """psb_geall + psb_geins + psb_geasb of a host array."""
    v = (mold or psb_d_hip_vect()).mold()
    v.bld(values)
    return psb_d_vect_type(v)


def psb_gefree(x, desc=None):
    x.free()


def _dev(x):
    v = x.v if isinstance(x, psb_d_vect_type) else x
    if not isinstance(v, psb_hip_base_vect):
        raise N.PsbError(N.INVALID, "psblas", "operand is not a device vector")
    v.ensure_dev()
    return v


def _kind(*vs):
    t = vs[0]._dtype
    for v in vs[1:]:
        if v._dtype != t:
            raise N.PsbError(N.INVALID, "psblas", "operands of different precision")


# ------------------------------------------------------------------ BLAS
def psb_spmm(alpha, a, x, beta, y, desc_a=None, trans="N"):
    """psb_spmm (base/psblas/psb_dspmm.f90:57-310): y = alpha*op(A)*x + beta*y;
    op 'T' / 'C' runs on the transpose the storage builds on the device (on
    the host copy when an operand is a host vector)."""
    t = str(trans).upper()
    if t not in ("N", "T", "C"):
        raise N.PsbError(N.INVALID, "psb_spmm", f"trans {trans!r}")
    # the storage plug-ins: device vectors take the device product, any other
    # vector type the matrix's host copy (cuda/impl/psb_d_cuda_hlg_vect_mv.F90:77-108)
    xv = x.v if isinstance(x, psb_d_vect_type) else x
    yv = y.v if isinstance(y, psb_d_vect_type) else y
    if t == "N":
        a.a.vect_mv(alpha, xv, beta, yv)
    else:                       # 'T' and 'C' are the same product for real data
        a.a.vect_mv_t(alpha, xv, beta, yv)


def psb_spsm(alpha, a, x, beta, y, desc_a=None, trans="N", scale="U", choice=None, diag=None):
    """psb_spsm / psb_dspsv_vect (base/psblas/psb_dspsm.f90): y = alpha *
    op(T)^-1 x + beta * y with the triangular matrix T (its storage flagged
    by set_triangle / set_lower / set_upper / set_unit), and with a diagonal
    d (diag) the scaled forms of psb_d_base_vect_cssv
    (base/serial/impl/psb_d_base_mat_impl.F90:2015-2130): scale 'R' solves
    with d .* x, scale 'L' gives alpha * d .* (op(T)^-1 x) + beta * y (the
    product d .* t, then y = t + beta * y); without diag the scale is
    ignored, with diag a scale other than 'L' / 'R' is an error (info 31).
    Every step runs on the device."""
    xv, yv = _dev(x), _dev(y)
    t = str(trans).upper()
    if diag is None:
        a.a.inner_vect_sv(alpha, xv, beta, yv, trans=t)
        return
    dv = _dev(diag)
    _kind(xv, yv, dv)
    sc = "L" if scale is None else str(scale).upper()
    nar, nac = a.get_nrows(), a.get_ncols()
    if sc not in ("L", "R"):
        raise N.PsbError(31, "psb_spsm", f"scale {scale!r} with a diagonal")
    if dv.get_nrows() < (nac if sc == "R" else nar):
        raise N.PsbError(N.INVALID, "psb_spsm", "diagonal too short")
    if sc == "R":                       # tmp = d .* x, then the solve
        tmp = xv.mold()
        tmp.all(nac)
        tmp.mlt_v_2(1.0, dv, xv, 0.0)
        a.a.inner_vect_sv(alpha, tmp, beta, yv, trans=t)
        tmp.free()
    elif beta == 0.0:                   # the solve, then y = y .* d
        a.a.inner_vect_sv(alpha, xv, 0.0, yv, trans=t)
        yv.mlt_v(dv)
    else:                               # tmp = d .* (alpha T^-1 x); y = tmp + beta y
        tmp = yv.mold()
        tmp.all(nar)
        a.a.inner_vect_sv(alpha, xv, 0.0, tmp, trans=t)
        tmp.mlt_v(dv)
        yv.axpby_v(nar, 1.0, tmp, beta)
        tmp.free()


def psb_geaxpby(alpha, x, beta, y, desc_a=None):
    """psb_geaxpby (base/psblas/psb_daxpby.f90:51-131): y = alpha*x + beta*y."""
    xv, yv = _dev(x), _dev(y)
    _kind(xv, yv)
    n = yv.get_nrows()
    N.call("psb_hip_mv_axpby", n, float(alpha), xv.dev, float(beta), yv.dev)
    yv.set_dev()


def psb_geaxpbyz(alpha, x, beta, y, z, desc_a=None):
    """psb_geaxpby with a separate output: z = alpha*x + beta*y."""
    xv, yv, zv = _dev(x), _dev(y), _dev(z)
    _kind(xv, yv, zv)
    N.call("psb_hip_mv_axpbyz", zv.get_nrows(), float(alpha), xv.dev, float(beta), yv.dev, zv.dev)
    zv.set_dev()


def psb_upd_xyz(alpha, beta, gamma, delta, x, y, z, desc_a=None):
    """psb_upd_xyz: y = alpha*x + beta*y; z = gamma*y + delta*z."""
    xv, yv, zv = _dev(x), _dev(y), _dev(z)
    _kind(xv, yv, zv)
    N.call("psb_hip_mv_upd_xyz", yv.get_nrows(), float(alpha), float(beta), float(gamma),
           float(delta), xv.dev, yv.dev, zv.dev)
    yv.set_dev()
    zv.set_dev()


def psb_xyzw(a, b, c, d, e, f, x, y, z, w, desc_a=None):
    """psb_abgdxyz / xyzw: y = a*x+b*y; z = c*y+d*z; w = e*z+f*w."""
    xv, yv, zv, wv = _dev(x), _dev(y), _dev(z), _dev(w)
    _kind(xv, yv, zv, wv)
    N.call("psb_hip_mv_xyzw", yv.get_nrows(), *[float(t) for t in (a, b, c, d, e, f)],
           xv.dev, yv.dev, zv.dev, wv.dev)
    for v in (yv, zv, wv):
        v.set_dev()


def psb_gedot(x, y, desc_a=None):
    """psb_gedot (base/psblas/psb_ddot.f90:51-159); the result leaves the
    device (one synchronisation), psb_sum is the identity at np = 1."""
    import ctypes as C
    xv, yv = _dev(x), _dev(y)
    _kind(xv, yv)
    res = (C.c_double * max(xv.count, 1))()
    N.call("psb_hip_mv_dot", res, xv.get_nrows(), xv.dev, yv.dev)
    return res[0] if xv.count == 1 else np.array(res[:xv.count])


def _reduce(name, x):
    import ctypes as C
    xv = _dev(x)
    res = (C.c_double * max(xv.count, 1))()
    N.call(name, res, xv.get_nrows(), xv.dev)
    return res[0] if xv.count == 1 else np.array(res[:xv.count])


def psb_genrm2(x, desc_a=None):
    """psb_genrm2 (base/psblas/psb_dnrm2.f90:281-374)."""
    return _reduce("psb_hip_mv_nrm2", x)


def psb_geamax(x, desc_a=None):
    """psb_geamax: max |x_i|."""
    return _reduce("psb_hip_mv_amax", x)


def psb_geasum(x, desc_a=None):
    """psb_geasum: sum |x_i|."""
    return _reduce("psb_hip_mv_asum", x)


def psb_gemlt(alpha, x, y, beta, z, desc_a=None):
    """psb_gemlt / z%mlt(alpha, x, y, beta): z = alpha*y.*x + beta*z."""
    xv, yv, zv = _dev(x), _dev(y), _dev(z)
    _kind(xv, yv, zv)
    N.call("psb_hip_mv_axybz", zv.get_nrows(), float(alpha), xv.dev, yv.dev, float(beta), zv.dev)
    zv.set_dev()


def psb_geabs(x, y, desc_a=None):
    """psb_geabs: y = |x|."""
    xv, yv = _dev(x), _dev(y)
    N.call("psb_hip_mv_abs", yv.get_nrows(), 1.0, xv.dev, yv.dev)
    yv.set_dev()


def psb_gescal(alpha, x, desc_a=None):
    xv = _dev(x)
    N.call("psb_hip_mv_scal", float(alpha), xv.dev)
    xv.set_dev()
