"""BJAC / ILU(0) (psb_dilu0_fact, psb_d_bjac_apply_vect): the oracle pinned by
the defining property of ILU(0) and by exact cases, and the library's host
factorization checked against it bit for bit (no GPU needed)."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from psblas3_amd import _native as N


def _csr(orc, coo):
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    A = sp.csr_matrix((val, ja - 1, irp - 1), shape=(coo.nr, coo.nc)).toarray()
    return (irp, ja, val), A


def _dense_factors(f):
    n = f.n
    L = np.eye(n)
    for i in range(n):
        for k in range(f.lirp[i] - 1, f.lirp[i + 1] - 1):
            L[i, f.lja[k] - 1] = f.lval[k]
    U = np.diag(1.0 / f.d)                       # pivots on the diagonal
    for i in range(n):
        for k in range(f.uirp[i] - 1, f.uirp[i + 1] - 1):
            U[i, f.uja[k] - 1] = f.uval[k] / f.d[i]
    return L, U


def _cases(orc):
    return {"convdiff": orc.gen_convdiff(6, conv=3.0),
            "aniso": orc.gen_aniso2d(9, 0.05),
            "poisson": orc.gen_poisson3d(5)}


@pytest.mark.parametrize("name", ["convdiff", "aniso", "poisson"])
def test_ilu0_matches_a_on_its_pattern(orc, name):
    csr, A = _csr(orc, _cases(orc)[name])
    f = orc.ilu0_fact(*csr)
    L, U = _dense_factors(f)
    P = L @ U
    mask = A != 0
    assert np.max(np.abs(P - A)[mask]) <= 1e-13 * np.max(np.abs(A))
    # the factors keep A's pattern: no fill
    assert np.all((L != 0)[np.tril(np.ones_like(A), -1) > 0] <= mask[np.tril(np.ones_like(A), -1) > 0])


def test_ilu0_tridiagonal_is_exact_lu(orc):
    n = 50
    rng = np.random.default_rng(1)
    ia, ja, val = [], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0 + 0.1 * rng.standard_normal()), (i, 4.0), (i + 1, -1.2)):
            if 0 <= j < n:
                ia.append(i + 1); ja.append(j + 1); val.append(v)
    coo = orc.COO(n, n, np.array(ia, np.int32), np.array(ja, np.int32), np.array(val))
    csr, A = _csr(orc, coo)
    f = orc.ilu0_fact(*csr)
    x = orc.uniform_pm1(3, n)
    y = orc.bjac_apply(f, A @ x)
    assert np.linalg.norm(y - x) <= 1e-13 * np.linalg.norm(x)


def test_bjac_apply_equals_dense_solves(orc):
    csr, A = _csr(orc, _cases(orc)["convdiff"])
    f = orc.ilu0_fact(*csr)
    L, U = _dense_factors(f)
    x = orc.uniform_pm1(4, f.n)
    w = sla.solve_triangular(L, x, lower=True, unit_diagonal=True)
    U1 = U * f.d[:, None]                        # unit upper: U scaled by the inverse pivots
    y = sla.solve_triangular(U1, f.d * w, lower=False, unit_diagonal=True)
    assert np.allclose(orc.bjac_apply(f, x), y, rtol=1e-12, atol=1e-14)
    # the oracle triangular solves alone
    assert np.allclose(orc.trsv_unit(f.lirp, f.lja, f.lval, x), w, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name", ["convdiff", "aniso", "poisson"])
def test_host_ilu0_fact_bitexact(orc, name):
    (irp, ja, val), _ = _csr(orc, _cases(orc)[name])
    f = orc.ilu0_fact(irp, ja, val)
    n, nz = len(irp) - 1, len(val)
    lirp, uirp = np.zeros(n + 1, np.int32), np.zeros(n + 1, np.int32)
    lja, uja = np.zeros(nz, np.int32), np.zeros(nz, np.int32)
    lval, uval, d = np.zeros(nz), np.zeros(nz), np.zeros(n)
    bad = C.c_int(0)
    rc = N.fn("psb_hip_ilu0_fact")(n, irp.ctypes.data, ja.ctypes.data, val.ctypes.data,
                                   lirp.ctypes.data, lja.ctypes.data, lval.ctypes.data,
                                   uirp.ctypes.data, uja.ctypes.data, uval.ctypes.data,
                                   d.ctypes.data, C.byref(bad))
    assert rc == 0
    nl, nu = lirp[n] - 1, uirp[n] - 1
    assert np.array_equal(lirp, f.lirp) and np.array_equal(uirp, f.uirp)
    assert np.array_equal(lja[:nl], f.lja) and np.array_equal(uja[:nu], f.uja)
    for a_, b_ in ((lval[:nl], f.lval), (uval[:nu], f.uval), (d, f.d)):
        assert np.array_equal(a_.view(np.int64), b_.view(np.int64))


def test_ilu0_small_pivot(orc):
    # [[1, 1], [1, 1]]: the second pivot is 0
    irp = np.array([1, 3, 5], np.int32)
    ja = np.array([1, 2, 1, 2], np.int32)
    val = np.ones(4)
    with pytest.raises(ZeroDivisionError):
        orc.ilu0_fact(irp, ja, val)
    lirp, uirp = np.zeros(3, np.int32), np.zeros(3, np.int32)
    lja, uja = np.zeros(4, np.int32), np.zeros(4, np.int32)
    lval, uval, d = np.zeros(4), np.zeros(4), np.zeros(2)
    bad = C.c_int(0)
    rc = N.fn("psb_hip_ilu0_fact")(2, irp.ctypes.data, ja.ctypes.data, val.ctypes.data,
                                   lirp.ctypes.data, lja.ctypes.data, lval.ctypes.data,
                                   uirp.ctypes.data, uja.ctypes.data, uval.ctypes.data,
                                   d.ctypes.data, C.byref(bad))
    assert rc == N.INVALID and bad.value == 2


@pytest.mark.parametrize("method", ["CG", "BICGSTAB", "RGMRES", "BICGSTABL"])
def test_bjac_solvers_converge(orc, method):
    kind = "poisson" if method == "CG" else "convdiff"
    csr, A = _csr(orc, _cases(orc)[kind])
    xt = orc.uniform_pm1(5, A.shape[0])
    b = A @ xt
    x, it, err, _ = orc.krylov_bjac(method, csr, b, eps=1e-10, itmax=1000,
                                    irst=2 if method == "BICGSTABL" else None)
    _, itd, _, _ = orc.krylov(method, csr, b, eps=1e-10, itmax=1000,
                              irst=2 if method == "BICGSTABL" else None)
    assert err <= 1e-10 and it < itd
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-6 * np.linalg.norm(xt)


# ---------------------------------------------------------------- single precision
@pytest.mark.parametrize("name", ["convdiff", "aniso", "poisson"])
def test_ilu0_single_precision(orc, name):
    """psb_silu0_fact: the same elimination in FP32.  Pinned by the ILU(0)
    property at FP32 accuracy, by the FP64 factors (agreement to FP32
    rounding) and, for the library's host code, bit for bit."""
    (irp, ja, val), A = _csr(orc, _cases(orc)[name])
    v32 = val.astype(np.float32)
    f = orc.ilu0_fact(irp, ja, v32)
    assert f.lval.dtype == np.float32 and f.d.dtype == np.float32
    L, U = _dense_factors(f)
    A32 = sp.csr_matrix((v32.astype(np.float64), ja - 1, irp - 1), shape=A.shape).toarray()
    mask = A32 != 0
    assert np.max(np.abs(L @ U - A32)[mask]) <= 1e-5 * np.max(np.abs(A32))
    f64 = orc.ilu0_fact(irp, ja, v32.astype(np.float64))
    assert np.allclose(f.lval, f64.lval, rtol=1e-4, atol=1e-6)
    assert np.allclose(f.d, f64.d, rtol=1e-4)
    n, nz = len(irp) - 1, len(val)
    lirp, uirp = np.zeros(n + 1, np.int32), np.zeros(n + 1, np.int32)
    lja, uja = np.zeros(nz, np.int32), np.zeros(nz, np.int32)
    lval, uval, d = (np.zeros(nz, np.float32), np.zeros(nz, np.float32),
                     np.zeros(n, np.float32))
    bad = C.c_int(0)
    rc = N.fn("psb_hip_ilu0_fact_s")(n, irp.ctypes.data, ja.ctypes.data, v32.ctypes.data,
                                     lirp.ctypes.data, lja.ctypes.data, lval.ctypes.data,
                                     uirp.ctypes.data, uja.ctypes.data, uval.ctypes.data,
                                     d.ctypes.data, C.byref(bad))
    assert rc == 0
    nl, nu = lirp[n] - 1, uirp[n] - 1
    assert np.array_equal(lja[:nl], f.lja) and np.array_equal(uja[:nu], f.uja)
    for a_, b_ in ((lval[:nl], f.lval), (uval[:nu], f.uval), (d, f.d)):
        assert np.array_equal(a_.view(np.int32), b_.view(np.int32))


def test_ilu0_single_tridiagonal_against_fp64(orc):
    """The (1, 2, 1) tridiagonal matrix, where ILU(0) is the exact LU: its
    pivots are (k+1)/k, so the FP32 inverse pivots equal the FP64 ones to
    FP32 rounding, the first (1/2) exactly."""
    n = 40
    ia, ja, val = [], [], []
    for i in range(n):
        for j, v in ((i - 1, 1.0), (i, 2.0), (i + 1, 1.0)):
            if 0 <= j < n:
                ia.append(i + 1); ja.append(j + 1); val.append(v)
    (irp, jj, vv), _ = _csr(orc, orc.COO(n, n, np.array(ia, np.int32), np.array(ja, np.int32),
                                         np.array(val)))
    f32 = orc.ilu0_fact(irp, jj, vv.astype(np.float32))
    f64 = orc.ilu0_fact(irp, jj, vv)
    assert f32.d[0] == np.float32(0.5) and f64.d[0] == 0.5
    assert np.allclose(f64.d, np.arange(1, n + 1) / np.arange(2, n + 2), rtol=1e-12)
    assert np.allclose(f32.d, f64.d, rtol=1e-6)
