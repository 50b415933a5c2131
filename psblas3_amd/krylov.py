"""Krylov solvers over the device plug-ins: psb_krylov (krylov/psb_dkrylov.f90:
81-227), CG (krylov/psb_dcg.F90:98-333), BiCGSTAB (krylov/psb_dcgstab.f90:
96-393), FCG (krylov/psb_dfcg.F90), GCR (krylov/psb_dgcr.f90), RGMRES
(krylov/psb_drgmres.f90), BiCGSTAB(l) (krylov/psb_dcgstabl.f90), CGS
(krylov/psb_dcgs.f90) and BiCG (krylov/psb_dbicg.f90) with the
stopping tests of krylov/psb_d_krylov_conv_mod.f90:205-356.

These routines issue exactly the reference's sequence of psb_spmm /
psb_geaxpby / psb_gedot / psb_genrm2 / prec%apply calls, each running on the
device; every scalar (dot, nrm2) returns to the host as in the reference.
psb_krylov_fused and psb_krylov_mixed below call the device-resident
solvers of libpsb_hip (one synchronisation per several iterations).
"""
import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _native as N
from .psblas import (psb_geall, psb_geamax, psb_geaxpby, psb_gedot, psb_genrm2, psb_gescal,
                     psb_spmm)
from .vect import psb_d_vect_type


@dataclass
class psb_itconv_type:
    """stopdat of psb_d_krylov_conv_mod."""
    istop: int = 2
    itmax: int = 1000
    itrace: int = 0
    eps: float = 0.0
    ani: float = 0.0
    bni: float = 0.0
    bn2: float = 0.0
    r0n2: float = 0.0
    errnum: float = 0.0
    errden: float = 1.0
    history: List[float] = field(default_factory=list)


def psb_spnrmi(a, desc_a=None):
    """||A||_inf, needed by istop = 1 (read back once)."""
    import numpy as np
    a_ = a.a
    if hasattr(a_, "to_host") and a_.get_fmt() == "CSRG":
        irp, ja, val = a_.to_host()
        rows = np.repeat(np.arange(a_.nr), np.diff(irp))
        return float(np.bincount(rows, weights=np.abs(val), minlength=a_.nr).max())
    if a_.get_fmt() == "HLG":
        irp, ja, val = a_.to_csr()
        rows = np.repeat(np.arange(a_.nr), np.diff(irp))
        return float(np.bincount(rows, weights=np.abs(val), minlength=a_.nr).max())
    irn, ja, val, idiag = a_.to_host()
    v = np.abs(val).reshape(a_.nzm, a_.nr)
    return float(v.sum(axis=0).max())


def psb_init_conv(methdname, istop, itrace, itmax, a, x, b, eps, desc_a):
    """psb_d_init_conv_vect (psb_d_krylov_conv_mod.f90:205-277)."""
    s = psb_itconv_type(istop=istop, itmax=itmax, itrace=itrace, eps=eps)
    if istop == 1:
        s.ani = psb_spnrmi(a, desc_a)
        s.bni = psb_geamax(b, desc_a)
    elif istop == 2:
        s.bn2 = psb_genrm2(b, desc_a)
    elif istop == 3:
        r = psb_geall(desc_a, mold=x.v)
        psb_geaxpby(1.0, b, 0.0, r, desc_a)
        psb_spmm(-1.0, a, x, 1.0, r, desc_a)
        s.r0n2 = psb_genrm2(r, desc_a)
        r.free()
    else:
        raise N.PsbError(N.INVALID, "psb_init_conv", f"invalid istop {istop}")
    s.errnum, s.errden = 0.0, 1.0
    if itrace > 0:
        print(f"{methdname + ':':18s} Iteration   Error Estimate       Tolerance")
    return s


def psb_check_conv(methdname, it, x, r, desc_a, s, record=True):
    """psb_d_check_conv_vect (psb_d_krylov_conv_mod.f90:279-356)."""
    if s.istop == 1:
        rni = psb_geamax(r, desc_a)
        xni = psb_geamax(x, desc_a)
        s.errnum, s.errden = rni, s.ani * xni + s.bni
    elif s.istop == 2:
        s.errnum, s.errden = psb_genrm2(r, desc_a), s.bn2
    else:
        s.errnum, s.errden = psb_genrm2(r, desc_a), s.r0n2
    if s.errden == 0.0:
        res = s.errnum <= s.eps
    else:
        res = s.errnum <= s.eps * s.errden
    if record:
        s.history.append(s.errnum / s.errden if s.errden > 0 else s.errnum)
    res = res or (s.itmax <= it)
    if s.itrace > 0 and (it % s.itrace == 0 or res):
        e = s.errnum / s.errden if s.errden > 0 else s.errnum
        print(f"{methdname + ':':18s} {it:4d}  {e:15.9e}  {s.eps:15.9e}")
    return res


def psb_end_conv(s, itx):
    """psb_d_end_conv: err = errnum/errden (or errnum), iter = itx."""
    err = s.errnum / s.errden if s.errden > 0 else s.errnum
    return itx, err


def psb_dcg_vect(a, prec, b, x, eps, desc_a, itmax=1000, itrace=0, istop=2):
    """psb_dcg_vect (krylov/psb_dcg.F90:98-333)."""
    methd = "CG"
    p, q, r, z = (psb_geall(desc_a, mold=x.v) for _ in range(4))
    itx = 0
    s = None
    done = False
    try:
        while not done:  # restart
            if itx >= itmax:
                break
            it = 0
            psb_geaxpby(1.0, b, 0.0, r, desc_a)
            psb_spmm(-1.0, a, x, 1.0, r, desc_a)
            rho = 0.0
            s = psb_init_conv(methd, istop, itrace, itmax, a, x, b, eps, desc_a)
            while True:  # iteration
                it += 1
                itx += 1
                prec.apply(r, z, desc_a)
                rho_old = rho
                rho = psb_gedot(r, z, desc_a)
                if it == 1:
                    psb_geaxpby(1.0, z, 0.0, p, desc_a)
                else:
                    if rho_old == 0.0:
                        break
                    beta = rho / rho_old
                    psb_geaxpby(1.0, z, beta, p, desc_a)
                psb_spmm(1.0, a, p, 0.0, q, desc_a)
                sigma = psb_gedot(p, q, desc_a)
                if sigma == 0.0:
                    break
                alpha = rho / sigma
                psb_geaxpby(alpha, p, 1.0, x, desc_a)
                psb_geaxpby(-alpha, q, 1.0, r, desc_a)
                if psb_check_conv(methd, itx, x, r, desc_a, s):
                    done = True
                    break
    finally:
        for v in (p, q, r, z):
            v.free()
    if s is None:
        return itx, 0.0, []
    it_, err = psb_end_conv(s, itx)
    return it_, err, s.history


def psb_dcgstab_vect(a, prec, b, x, eps, desc_a, itmax=1000, itrace=0, istop=2):
    """psb_dcgstab_vect (krylov/psb_dcgstab.f90:96-393)."""
    methd = "BiCGSTAB"
    q, r, p, v, sv, t, z, f = (psb_geall(desc_a, mold=x.v) for _ in range(8))
    itx = 0
    s = psb_init_conv(methd, istop, itrace, itmax, a, x, b, eps, desc_a)
    alpha = omega = rho = 0.0
    try:
        done = False
        while not done:  # restart
            if itx >= itmax:
                break
            it = 0
            psb_geaxpby(1.0, b, 0.0, r, desc_a)
            psb_spmm(-1.0, a, x, 1.0, r, desc_a)
            psb_geaxpby(1.0, r, 0.0, q, desc_a)
            if psb_check_conv(methd, itx, x, r, desc_a, s, record=False):
                break
            rho = 0.0
            while True:
                it += 1
                itx += 1
                rho_old = rho
                rho = psb_gedot(q, r, desc_a)
                if rho == 0.0:
                    break
                if it == 1:
                    psb_geaxpby(1.0, r, 0.0, p, desc_a)
                else:
                    beta = (rho / rho_old) * (alpha / omega)
                    psb_geaxpby(-omega, v, 1.0, p, desc_a)
                    psb_geaxpby(1.0, r, beta, p, desc_a)
                prec.apply(p, f, desc_a)
                psb_spmm(1.0, a, f, 0.0, v, desc_a)
                sigma = psb_gedot(q, v, desc_a)
                if sigma == 0.0:
                    break
                alpha = rho / sigma
                psb_geaxpby(1.0, r, 0.0, sv, desc_a)
                psb_geaxpby(-alpha, v, 1.0, sv, desc_a)
                prec.apply(sv, z, desc_a)
                psb_spmm(1.0, a, z, 0.0, t, desc_a)
                sigma = psb_gedot(t, t, desc_a)
                if sigma == 0.0:
                    break
                tau = psb_gedot(t, sv, desc_a)
                omega = tau / sigma
                if omega == 0.0:
                    break
                psb_geaxpby(alpha, f, 1.0, x, desc_a)
                psb_geaxpby(omega, z, 1.0, x, desc_a)
                psb_geaxpby(1.0, sv, 0.0, r, desc_a)
                psb_geaxpby(-omega, t, 1.0, r, desc_a)
                if psb_check_conv(methd, itx, x, r, desc_a, s):
                    done = True
                    break
    finally:
        for w in (q, r, p, v, sv, t, z, f):
            w.free()
    it_, err = psb_end_conv(s, itx)
    return it_, err, s.history


def psb_dcgs_vect(a, prec, b, x, eps, desc_a, itmax=1000, itrace=0, istop=2):
    """psb_dcgs_vect (krylov/psb_dcgs.f90:100-400): CGS; a zero rho or sigma
    restarts from the true residual."""
    methd = "CGS"
    q, qt, r, rt, p, v, uv, z, f, sv = (psb_geall(desc_a, mold=x.v) for _ in range(10))
    itx = 0
    s = psb_init_conv(methd, istop, itrace, itmax, a, x, b, eps, desc_a)
    try:
        done = False
        while not done:  # restart
            if itx >= itmax:
                break
            it = 0
            psb_geaxpby(1.0, b, 0.0, r, desc_a)
            psb_spmm(-1.0, a, x, 1.0, r, desc_a)
            psb_geaxpby(1.0, r, 0.0, rt, desc_a)
            if psb_check_conv(methd, itx, x, r, desc_a, s, record=False):
                break
            rho = 0.0
            while True:
                it += 1
                itx += 1
                rho_old = rho
                rho = psb_gedot(rt, r, desc_a)
                if rho == 0.0:
                    break
                if it == 1:
                    psb_geaxpby(1.0, r, 0.0, uv, desc_a)
                    psb_geaxpby(1.0, r, 0.0, p, desc_a)
                else:
                    beta = rho / rho_old
                    psb_geaxpby(1.0, r, 0.0, uv, desc_a)
                    psb_geaxpby(beta, q, 1.0, uv, desc_a)
                    psb_geaxpby(1.0, q, beta, p, desc_a)
                    psb_geaxpby(1.0, uv, beta, p, desc_a)
                prec.apply(p, f, desc_a)
                psb_spmm(1.0, a, f, 0.0, v, desc_a)
                sigma = psb_gedot(rt, v, desc_a)
                if sigma == 0.0:
                    break
                alpha = rho / sigma
                psb_geaxpby(1.0, uv, 0.0, q, desc_a)
                psb_geaxpby(-alpha, v, 1.0, q, desc_a)
                psb_geaxpby(1.0, uv, 0.0, sv, desc_a)
                psb_geaxpby(1.0, q, 1.0, sv, desc_a)
                prec.apply(sv, z, desc_a)
                psb_geaxpby(alpha, z, 1.0, x, desc_a)
                psb_spmm(1.0, a, z, 0.0, qt, desc_a)
                psb_geaxpby(-alpha, qt, 1.0, r, desc_a)
                if psb_check_conv(methd, itx, x, r, desc_a, s):
                    done = True
                    break
    finally:
        for w in (q, qt, r, rt, p, v, uv, z, f, sv):
            w.free()
    it_, err = psb_end_conv(s, itx)
    return it_, err, s.history


def psb_dbicg_vect(a, prec, b, x, eps, desc_a, itmax=1000, itrace=0, istop=2):
    """psb_dbicg_vect (krylov/psb_dbicg.f90:100-330): BiCG with the shadow
    residual rt = r0; the products with A^T (psb_spmm trans 'C') and the
    preconditioner's transposed apply run on the device."""
    methd = "BICG"
    q, qt, r, rt, p, pt, z, zt = (psb_geall(desc_a, mold=x.v) for _ in range(8))
    itx = 0
    s = psb_init_conv(methd, istop, itrace, itmax, a, x, b, eps, desc_a)
    try:
        done = False
        while not done:  # restart
            if itx >= itmax:
                break
            it = 0
            psb_geaxpby(1.0, b, 0.0, r, desc_a)
            psb_spmm(-1.0, a, x, 1.0, r, desc_a)
            psb_geaxpby(1.0, r, 0.0, rt, desc_a)
            rho = 0.0
            if psb_check_conv(methd, itx, x, r, desc_a, s, record=False):
                break
            while True:
                it += 1
                itx += 1
                prec.apply(r, z, desc_a)
                prec.apply(rt, zt, desc_a, trans="C")
                rho_old = rho
                rho = psb_gedot(rt, z, desc_a)
                if rho == 0.0:
                    break
                if it == 1:
                    psb_geaxpby(1.0, z, 0.0, p, desc_a)
                    psb_geaxpby(1.0, zt, 0.0, pt, desc_a)
                else:
                    beta = rho / rho_old
                    psb_geaxpby(1.0, z, beta, p, desc_a)
                    psb_geaxpby(1.0, zt, beta, pt, desc_a)
                psb_spmm(1.0, a, p, 0.0, q, desc_a)
                psb_spmm(1.0, a, pt, 0.0, qt, desc_a, trans="C")
                sigma = psb_gedot(pt, q, desc_a)
                if sigma == 0.0:
                    break
                alpha = rho / sigma
                psb_geaxpby(alpha, p, 1.0, x, desc_a)
                psb_geaxpby(-alpha, q, 1.0, r, desc_a)
                psb_geaxpby(-alpha, qt, 1.0, rt, desc_a)
                if psb_check_conv(methd, itx, x, r, desc_a, s):
                    done = True
                    break
    finally:
        for w in (q, qt, r, rt, p, pt, z, zt):
            w.free()
    it_, err = psb_end_conv(s, itx)
    return it_, err, s.history


# ------------------------------------------------------------------ host scalars
def _trsv_upper(m, h, y):
    """dtrsm('l','u','n','n', m, 1, 1.0, H, ., y, .) in the reference BLAS loop
    order (column-oriented back substitution), 1-based h[i][k], y[k]."""
    for k in range(m, 0, -1):
        if y[k] != 0.0:
            y[k] = y[k] / h[k][k]
            for i in range(1, k):
                y[i] = y[i] - y[k] * h[i][k]


def _drotg(a, b):
    """Reference BLAS drotg: returns (c, s)."""
    roe = a if abs(a) > abs(b) else b
    scale = abs(a) + abs(b)
    if scale == 0.0:
        return 1.0, 0.0
    r = scale * math.sqrt((a / scale) ** 2 + (b / scale) ** 2)
    if roe < 0.0:
        r = -r
    return a / r, b / r


def _drot(x, y, c, s):
    """Reference BLAS drot on one pair: returns (c x + s y, c y - s x)."""
    return c * x + s * y, c * y - s * x


def _zero(v, desc_a):
    psb_geaxpby(0.0, v, 0.0, v, desc_a)


# ------------------------------------------------------------------ more methods
def psb_dfcg_vect(a, prec, b, x, eps, desc_a, itmax=1000, itrace=0, istop=2):
    """psb_dfcg_vect (krylov/psb_dfcg.F90:98-348): flexible CG."""
    methd = "FCG"
    w, v, d, q, r = (psb_geall(desc_a, mold=x.v) for _ in range(5))
    s = psb_init_conv(methd, istop, itrace, itmax, a, x, b, eps, desc_a)
    itx = 0
    try:
        done = False
        while not done:  # restart
            if itx >= itmax:
                break
            psb_geaxpby(1.0, b, 0.0, r, desc_a)
            psb_spmm(-1.0, a, x, 1.0, r, desc_a)
            if psb_check_conv(methd, itx, x, r, desc_a, s, record=False):
                break
            prec.apply(r, v, desc_a)
            psb_spmm(1.0, a, v, 0.0, w, desc_a)
            alpha = psb_gedot(r, v, desc_a)
            beta = psb_gedot(w, v, desc_a)
            if beta == 0.0:
                break
            psb_geaxpby(1.0, v, 0.0, d, desc_a)
            psb_geaxpby(1.0, w, 0.0, q, desc_a)
            delta = beta
            theta = alpha / delta
            psb_geaxpby(theta, d, 1.0, x, desc_a)
            psb_geaxpby(-theta, q, 1.0, r, desc_a)
            while True:  # iteration
                itx += 1
                if psb_check_conv(methd, itx, x, r, desc_a, s):
                    done = True
                    break
                prec.apply(r, v, desc_a)
                psb_spmm(1.0, a, v, 0.0, w, desc_a)
                alpha = psb_gedot(r, v, desc_a)
                beta = psb_gedot(w, v, desc_a)
                gamma = psb_gedot(q, v, desc_a)
                theta = gamma / delta
                psb_geaxpby(1.0, v, -theta, d, desc_a)
                psb_geaxpby(1.0, w, -theta, q, desc_a)
                delta = beta - (gamma * gamma) / delta
                if delta == 0.0:
                    break
                theta = alpha / delta
                psb_geaxpby(theta, d, 1.0, x, desc_a)
                psb_geaxpby(-theta, q, 1.0, r, desc_a)
    finally:
        for t in (w, v, d, q, r):
            t.free()
    it_, err = psb_end_conv(s, itx)
    return it_, err, s.history


def psb_dgcr_vect(a, prec, b, x, eps, desc_a, itmax=1000, itrace=0, istop=2, irst=10):
    """psb_dgcr_vect (krylov/psb_dgcr.f90:109-370), restart length irst.  The
    recurrence and stop tests are the reference's; x = x0 + every cycle's
    correction (the reference zeroes x before the first cycle's correction
    and again before the final one)."""
    methd = "GCR"
    nl = int(irst)
    if nl <= 0:
        raise N.PsbError(N.INVALID, "psb_dgcr", f"irst {irst}")
    r = psb_geall(desc_a, mold=x.v)
    z = [None] + [psb_geall(desc_a, mold=x.v) for _ in range(nl + 1)]
    c = [None] + [psb_geall(desc_a, mold=x.v) for _ in range(nl + 1)]
    cs = [None] + [psb_geall(desc_a, mold=x.v) for _ in range(nl + 1)]
    s = psb_init_conv(methd, istop, itrace, itmax, a, x, b, eps, desc_a)
    itx = 0
    try:
        while True:  # restart
            if itx >= itmax:
                break
            h = [[0.0] * (nl + 2) for _ in range(nl + 2)]
            alpha = [0.0] * (nl + 2)
            it = j = 0
            conv = False
            psb_geaxpby(1.0, b, 0.0, r, desc_a)
            psb_spmm(-1.0, a, x, 1.0, r, desc_a)
            if psb_check_conv(methd, itx, x, r, desc_a, s, record=False):
                break
            while True:  # iteration
                itx += 1
                it += 1
                j = it
                prec.apply(r, z[j], desc_a)
                psb_spmm(1.0, a, z[j], 0.0, c[1], desc_a)
                for i in range(1, j):
                    h[i][j] = psb_gedot(cs[i], c[i], desc_a)
                    psb_geaxpby(1.0, c[i], 0.0, c[i + 1], desc_a)
                    psb_geaxpby(-h[i][j], cs[i], 1.0, c[i + 1], desc_a)
                h[j][j] = psb_genrm2(c[j], desc_a)
                if h[j][j] == 0.0:
                    # A z(j) lies in the span of the earlier directions: z(j)
                    # adds nothing; the cycle ends on the j-1 before it (the
                    # reference divides by zero here, psb_dgcr.f90:292)
                    j -= 1
                    break
                psb_geaxpby(1.0 / h[j][j], c[j], 0.0, cs[j], desc_a)
                alpha[j] = psb_gedot(cs[j], r, desc_a)
                psb_geaxpby(-alpha[j], cs[j], 1.0, r, desc_a)
                conv = psb_check_conv(methd, itx, x, r, desc_a, s)
                if conv or j >= nl:
                    break
            _trsv_upper(j, h, alpha)
            for i in range(1, j + 1):
                psb_geaxpby(alpha[i], z[i], 1.0, x, desc_a)
            if conv:
                break
    finally:
        for t in [r] + z[1:] + c[1:] + cs[1:]:
            t.free()
    it_, err = psb_end_conv(s, itx)
    return it_, err, s.history


def psb_drgmres_vect(a, prec, b, x, eps, desc_a, itmax=1000, itrace=0, istop=2, irst=10):
    """psb_drgmres_vect (krylov/psb_drgmres.f90:109-502): restarted GMRES with
    right preconditioning, Givens rotations on the host.  istop 1 (normwise
    backward error, infinity norm) or 2 (||r||/||b||), as the reference
    accepts (:182); anything else is psb_err_invalid_istop_.  istop 1
    rebuilds xt = x + M^-1 V y from the partial least-squares solve at every
    inner iteration and tests ||b - A xt||_inf / (||A||_inf ||xt||_inf +
    ||b||_inf) (:387-405); the reference accumulates the rebuild into xt
    without clearing it (it clears w1 instead, :392), which only equals the
    rebuild at the first inner iteration; here xt is rebuilt from zero.  At an
    exact breakdown (h(i+1,i) = 0: the Krylov space is invariant) v(i+1) is
    not normalised; the rotation then zeroes the residual estimate and the
    solve ends there (the reference divides by zero and never uses the
    result)."""
    methd = "RGMRES"
    nl = int(irst)
    if nl <= 0:
        raise N.PsbError(N.INVALID, "psb_drgmres", f"irst {irst}")
    if istop not in (1, 2):
        raise N.PsbError(N.INVALID, "psb_drgmres", f"invalid istop {istop}")
    v = [None] + [psb_geall(desc_a, mold=x.v) for _ in range(nl + 1)]
    w = psb_geall(desc_a, mold=x.v)
    w1 = psb_geall(desc_a, mold=x.v)
    xt = psb_geall(desc_a, mold=x.v) if istop == 1 else None
    hist = []
    itx = 0
    errnum, errden = 0.0, 1.0
    try:
        ani = bni = bn2 = 0.0
        if istop == 1:
            ani = psb_spnrmi(a, desc_a)
            bni = psb_geamax(b, desc_a)
        else:
            bn2 = psb_genrm2(b, desc_a)
        while True:  # restart
            psb_geaxpby(1.0, b, 0.0, v[1], desc_a)
            psb_spmm(-1.0, a, x, 1.0, v[1], desc_a)
            rs = [0.0] * (nl + 2)
            rs[1] = psb_genrm2(v[1], desc_a)
            scal = 1.0 / rs[1] if rs[1] != 0.0 else math.inf
            if istop == 1:
                errnum = psb_geamax(v[1], desc_a)
                errden = ani * psb_geamax(x, desc_a) + bni
            else:
                errnum = rs[1]
                errden = bn2
            if errnum <= eps * errden:
                break
            psb_gescal(scal, v[1], desc_a)
            if itx >= itmax:
                break
            h = [[0.0] * (nl + 2) for _ in range(nl + 2)]
            cc = [0.0] * (nl + 2)
            sn = [0.0] * (nl + 2)
            done = False
            i = 0
            for i in range(1, nl + 1):  # inner
                itx += 1
                prec.apply(v[i], w1, desc_a)
                psb_spmm(1.0, a, w1, 0.0, w, desc_a)
                for k in range(1, i + 1):
                    h[k][i] = psb_gedot(v[k], w, desc_a)
                    psb_geaxpby(-h[k][i], v[k], 1.0, w, desc_a)
                h[i + 1][i] = psb_genrm2(w, desc_a)
                invariant = h[i + 1][i] == 0.0
                if not invariant:
                    psb_geaxpby(1.0 / h[i + 1][i], w, 0.0, v[i + 1], desc_a)
                for k in range(2, i + 1):
                    h[k - 1][i], h[k][i] = _drot(h[k - 1][i], h[k][i], cc[k - 1], sn[k - 1])
                cc[i], sn[i] = _drotg(h[i][i], h[i + 1][i])
                h[i][i], h[i + 1][i] = _drot(h[i][i], h[i + 1][i], cc[i], sn[i])
                h[i + 1][i] = 0.0
                rs[i], rs[i + 1] = _drot(rs[i], rs[i + 1], cc[i], sn[i])
                if istop == 1:
                    # xt = x + M^-1 (V y), y the least-squares solution so far
                    rst = list(rs)
                    _trsv_upper(i, h, rst)
                    _zero(w1, desc_a)
                    for k in range(1, i + 1):
                        psb_geaxpby(rst[k], v[k], 1.0, w1, desc_a)
                    prec.apply(w1, xt, desc_a)
                    psb_geaxpby(1.0, x, 1.0, xt, desc_a)
                    psb_geaxpby(1.0, b, 0.0, w1, desc_a)
                    psb_spmm(-1.0, a, xt, 1.0, w1, desc_a)
                    errnum = psb_geamax(w1, desc_a)
                    errden = ani * psb_geamax(xt, desc_a) + bni
                else:
                    errnum = abs(rs[i + 1])
                hist.append(errnum / errden if errden > 0 else errnum)
                if itrace > 0 and itx % itrace == 0:
                    print(f"{methd + ':':18s} {itx:9d} {errnum / errden:20.6e} {eps:20.6e}")
                if errnum <= eps * errden or invariant:
                    done = True
                    break
            if istop == 1:
                psb_geaxpby(1.0, xt, 0.0, x, desc_a)
            else:
                m = i if done else nl
                _trsv_upper(m, h, rs)
                _zero(w1, desc_a)
                for k in range(1, m + 1):
                    psb_geaxpby(rs[k], v[k], 1.0, w1, desc_a)
                prec.apply(w1, w, desc_a)
                psb_geaxpby(1.0, w, 1.0, x, desc_a)
            if done:
                break
    finally:
        for t in v[1:] + [w, w1] + ([xt] if xt is not None else []):
            t.free()
    return itx, (errnum / errden if errden > 0 else errnum), hist


def psb_dcgstabl_vect(a, prec, b, x, eps, desc_a, itmax=1000, itrace=0, istop=2, irst=1):
    """psb_dcgstabl_vect (krylov/psb_dcgstabl.f90:106-421): BiCGSTAB(l), l =
    irst, left preconditioning (the stop test sees the preconditioned
    residual, as the reference's)."""
    methd = "BiCGSTAB(L)"
    nl = int(irst)
    if nl <= 0:
        raise N.PsbError(N.INVALID, "psb_dcgstabl", f"irst {irst}")
    r = psb_geall(desc_a, mold=x.v)
    rt0 = psb_geall(desc_a, mold=x.v)
    rh = [psb_geall(desc_a, mold=x.v) for _ in range(nl + 1)]
    uh = [psb_geall(desc_a, mold=x.v) for _ in range(nl + 1)]
    s = psb_init_conv(methd, istop, itrace, itmax, a, x, b, eps, desc_a)
    itx = 0
    try:
        done = False
        while not done:  # restart
            if itx >= itmax:
                break
            psb_geaxpby(1.0, b, 0.0, r, desc_a)
            psb_spmm(-1.0, a, x, 1.0, r, desc_a)
            prec.apply(r, r, desc_a)
            psb_geaxpby(1.0, r, 0.0, rt0, desc_a)
            psb_geaxpby(1.0, r, 0.0, rh[0], desc_a)
            psb_geaxpby(0.0, r, 0.0, uh[0], desc_a)
            rho, alpha, omega = 1.0, 0.0, 1.0
            if psb_check_conv(methd, itx, x, r, desc_a, s, record=False):
                break
            while True:  # iteration
                itx += nl
                rho = -omega * rho
                brk = False
                gamma = [0.0] * (nl + 1)
                for j in range(nl):
                    rho_old = rho
                    rho = psb_gedot(rh[j], rt0, desc_a)
                    if rho == 0.0:
                        brk = True
                        break
                    beta = alpha * rho / rho_old
                    for k in range(j + 1):
                        psb_geaxpby(1.0, rh[k], -beta, uh[k], desc_a)
                    psb_spmm(1.0, a, uh[j], 0.0, uh[j + 1], desc_a)
                    prec.apply(uh[j + 1], uh[j + 1], desc_a)
                    gamma[j] = psb_gedot(uh[j + 1], rt0, desc_a)
                    if gamma[j] == 0.0:
                        brk = True
                        break
                    alpha = rho / gamma[j]
                    for k in range(j + 1):
                        psb_geaxpby(-alpha, uh[k + 1], 1.0, rh[k], desc_a)
                    psb_geaxpby(alpha, uh[0], 1.0, x, desc_a)
                    psb_spmm(1.0, a, rh[j], 0.0, rh[j + 1], desc_a)
                    prec.apply(rh[j + 1], rh[j + 1], desc_a)
                if brk:
                    break
                taum = [[0.0] * (nl + 1) for _ in range(nl + 1)]
                sigma = [0.0] * (nl + 1)
                gamma1 = [0.0] * (nl + 1)
                gamma2 = [0.0] * (nl + 1)
                for j in range(1, nl + 1):
                    for i in range(1, j):
                        taum[i][j] = psb_gedot(rh[i], rh[j], desc_a)
                        taum[i][j] = taum[i][j] / sigma[i]
                        psb_geaxpby(-taum[i][j], rh[i], 1.0, rh[j], desc_a)
                    sigma[j] = psb_gedot(rh[j], rh[j], desc_a)
                    gamma1[j] = psb_gedot(rh[0], rh[j], desc_a)
                    gamma1[j] = gamma1[j] / sigma[j]
                gamma[nl] = gamma1[nl]
                omega = gamma[nl]
                for j in range(nl - 1, 0, -1):
                    gamma[j] = gamma1[j]
                    for i in range(j + 1, nl + 1):
                        gamma[j] = gamma[j] - taum[j][i] * gamma[i]
                for j in range(1, nl):
                    gamma2[j] = gamma[j + 1]
                    for i in range(j + 1, nl):
                        gamma2[j] = gamma2[j] + taum[j][i] * gamma[i + 1]
                psb_geaxpby(gamma[1], rh[0], 1.0, x, desc_a)
                psb_geaxpby(-gamma1[nl], rh[nl], 1.0, rh[0], desc_a)
                psb_geaxpby(-gamma[nl], uh[nl], 1.0, uh[0], desc_a)
                for j in range(1, nl):
                    psb_geaxpby(-gamma[j], uh[j], 1.0, uh[0], desc_a)
                    psb_geaxpby(gamma2[j], rh[j], 1.0, x, desc_a)
                    psb_geaxpby(-gamma1[j], rh[j], 1.0, rh[0], desc_a)
                if psb_check_conv(methd, itx, x, rh[0], desc_a, s):
                    done = True
                    break
    finally:
        for t in [r, rt0] + rh + uh:
            t.free()
    it_, err = psb_end_conv(s, itx)
    return it_, err, s.history


def psb_krylov(method, a, prec, b, x, eps, desc_a, itmax=1000, itrace=0, istop=2,
               irst=None):
    """psb_krylov dispatcher (krylov/psb_dkrylov.f90:81-227): CG, FCG, GCR,
    CGS, BICG, BICGSTAB, RGMRES (or GMRES), BICGSTABL; unknown names fall back to
    BiCGSTAB as the reference's default branch does.  irst: restart length
    (GCR, RGMRES; default 10) or l (BICGSTABL; default 1).  Returns
    (iter, err, history); x is updated in place."""
    if not isinstance(x, psb_d_vect_type) or not isinstance(b, psb_d_vect_type):
        raise N.PsbError(N.INVALID, "psb_krylov", "b and x must be psb_d_vect_type")
    m = method.upper()
    if m == "CG":
        return psb_dcg_vect(a, prec, b, x, eps, desc_a, itmax, itrace, istop)
    if m == "FCG":
        return psb_dfcg_vect(a, prec, b, x, eps, desc_a, itmax, itrace, istop)
    if m == "GCR":
        return psb_dgcr_vect(a, prec, b, x, eps, desc_a, itmax, itrace, istop,
                             10 if irst is None else irst)
    if m in ("RGMRES", "GMRES"):
        return psb_drgmres_vect(a, prec, b, x, eps, desc_a, itmax, itrace, istop,
                                10 if irst is None else irst)
    if m == "BICGSTABL":
        return psb_dcgstabl_vect(a, prec, b, x, eps, desc_a, itmax, itrace, istop,
                                 1 if irst is None else irst)
    if m == "CGS":
        return psb_dcgs_vect(a, prec, b, x, eps, desc_a, itmax, itrace, istop)
    if m == "BICG":
        return psb_dbicg_vect(a, prec, b, x, eps, desc_a, itmax, itrace, istop)
    return psb_dcgstab_vect(a, prec, b, x, eps, desc_a, itmax, itrace, istop)


_FMT = {"HLG": 0, "ELG": 1, "CSRG": 2}


def _diag_only(prec, name):
    """The device-resident solvers apply DIAG (or no) preconditioning."""
    if prec is not None and prec.ptype not in ("DIAG", "NONE"):
        raise N.PsbError(N.UNSUPPORTED, name, f"preconditioner {prec.ptype} (DIAG or NONE)")


def psb_krylov_fused(method, a, prec, b, x, eps, desc_a=None, itmax=1000, istop=2,
                     check_every=8, history=False, noalloc=False):
    """Device-resident CG / BiCGSTAB (same recurrences as psb_krylov, scalars
    kept on the device, one host poll every check_every iterations).
    noalloc: run through psb_hip_krylov_fused_noalloc, whose buffers come
    from the library's static 64 MiB scratch array (no allocation per call;
    FP64 CG up to about 2M rows, BiCGSTAB about 1M).
    Returns (iter, err, history)."""
    import ctypes as C
    import numpy as np
    if not isinstance(x, psb_d_vect_type) or not isinstance(b, psb_d_vect_type):
        raise N.PsbError(N.INVALID, "psb_krylov_fused", "b and x must be psb_d_vect_type")
    b.v.ensure_dev()
    x.v.ensure_dev()
    _diag_only(prec, "psb_krylov_fused")
    diag = None
    if prec is not None and prec.ptype == "DIAG":
        diag = prec.dv.dev
    it, err = C.c_int(0), C.c_double(0.0)
    hist = np.zeros(max(itmax, 1)) if history else None
    meth = 0 if method.upper() == "CG" else 1
    entry = "psb_hip_krylov_fused_noalloc" if noalloc else "psb_hip_krylov_fused"
    N.call(entry, meth, _FMT[a.a.get_fmt()], a.a.dev, diag, b.v.dev, x.v.dev,
           float(eps), int(itmax), int(istop), int(check_every), C.byref(it), C.byref(err),
           hist.ctypes.data if history else None)
    x.v.set_dev()
    h = list(hist[:it.value]) if history else []
    return it.value, err.value, h


def psb_krylov_mixed(method, a, prec, b, x, eps, desc_a=None, outer_itmax=50, inner_eps=1e-4,
                     inner_itmax=5000, check_every=8, variant="restart"):
    """FP32-inner / FP64-outer iterative refinement (HLG matrices).  Builds the
    FP32 value copy on first use.  variant "restart": each refinement cycle
    restarts the inner FP32 solve on r/||r||; "reliable" (CG): one FP32
    recurrence with FP64 residual replacement whenever the residual drops by
    inner_eps.  Returns (outer_its, inner_its, err)."""
    import ctypes as C
    if a.a.get_fmt() != "HLG" or a.a._dtype != np.float64:
        raise N.PsbError(N.UNSUPPORTED, "psb_krylov_mixed", "needs a double HLG matrix")
    if not getattr(a.a, "_has_single", False):
        N.call("psb_hip_hll_make_single", a.a.dev)
        a.a._has_single = True
    b.v.ensure_dev()
    x.v.ensure_dev()
    _diag_only(prec, "psb_krylov_mixed")
    diag = prec.dv.dev if (prec is not None and prec.ptype == "DIAG") else None
    oi, ii, err = C.c_int(0), C.c_int(0), C.c_double(0.0)
    meth = 0 if method.upper() == "CG" else 1
    var = {"restart": 0, "reliable": 1}[variant]
    N.call("psb_hip_krylov_mixed", meth, a.a.dev, diag, b.v.dev, x.v.dev, float(eps),
           int(outer_itmax), float(inner_eps), int(inner_itmax), int(check_every), var,
           C.byref(oi), C.byref(ii), C.byref(err))
    x.v.set_dev()
    return oi.value, ii.value, err.value
