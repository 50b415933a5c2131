"""Pins the oracle's SpMV, SpMM and BLAS-1.

SpMV: the torture known answers (test/torture/psb_{d,s}_mvsv_tester.f90,
extracted to tests/golden/torture_usmv.json), dense brute force,
scipy.sparse, and an exact integer case.  BLAS-1: the constant-vector known
answers of test/kernel/vecoperation.f90:233-360 and math.fsum error bounds.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _torture_cases():
    with open(os.path.join(GOLDEN, "torture_usmv.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _torture_cases(), ids=lambda c: c["name"])
def test_torture_usmv(orc, case):
    coo = orc.fix_coo(orc.COO(case["m"], case["k"], np.array(case["ia"], np.int32),
                              np.array(case["ja"], np.int32), np.array(case["va"])))
    dt = np.float32 if case["kind"] == "s" else np.float64
    x = np.array(case["x"], dt)
    y0 = np.array(case["bcy"], dt)
    irp, ja, val = orc.csr_from_coo(coo)
    y = orc.csr_spmv(irp, ja, val.astype(dt), x, case["alpha"], case["beta"], y0)
    assert y.tolist() == case["cy"]
    for hksz in (1, 2, 32):
        h = orc.hll_from_coo(coo, hksz)
        # the torture answers require every row to be written, which the
        # CPU HLL's empty-hack skip (psb_d_hll_csmv.f90:222-226) violates for
        # hksz = 1 on an empty row: checked with the skip off, as the device
        # kernels do
        y = orc.hll_spmv(h, x, case["alpha"], case["beta"], y0, val=h.val.astype(dt),
                         quirk_empty_hack=False)
        assert y.tolist() == case["cy"], hksz
    e = orc.ell_from_coo(coo)
    y = orc.ell_spmv(e, x, case["alpha"], case["beta"], y0, val=e.val.astype(dt))
    assert y.tolist() == case["cy"]


def _bound(a_abs, x, nnz_row, eps):
    return 4.0 * nnz_row * eps * (a_abs @ np.abs(x))


@pytest.mark.parametrize("fmt", ["csr", "hll32", "hll64", "ell"])
def test_spmv_vs_dense(orc, fmt):
    c = orc.fix_coo(orc.gen_powerlaw(400, cap=60, seed=3))
    d = np.zeros((c.nr, c.nc))
    np.add.at(d, (c.ia - 1, c.ja - 1), c.val)
    x = orc.uniform_pm1(7, c.nc)
    y0 = orc.uniform_pm1(8, c.nr)
    alpha, beta = -1.5, 0.25
    if fmt == "csr":
        irp, ja, val = orc.csr_from_coo(c)
        y = orc.csr_spmv(irp, ja, val, x, alpha, beta, y0)
    elif fmt == "ell":
        y = orc.ell_spmv(orc.ell_from_coo(c), x, alpha, beta, y0)
    else:
        y = orc.hll_spmv(orc.hll_from_coo(c, int(fmt[3:])), x, alpha, beta, y0)
    ref = alpha * (d @ x) + beta * y0
    nnz_row = np.bincount(c.ia - 1, minlength=c.nr)
    tol = 4 * (nnz_row + 2) * np.finfo(float).eps * (np.abs(alpha) * (np.abs(d) @ np.abs(x))
                                                     + abs(beta) * np.abs(y0))
    assert np.all(np.abs(y - ref) <= tol)


def test_spmv_vs_scipy_stencil(orc):
    c = orc.fix_coo(orc.gen_convdiff(20))
    irp, ja, val = orc.csr_from_coo(c)
    A = sp.csr_matrix((val, ja - 1, irp - 1), shape=(c.nr, c.nc))
    x = orc.uniform_pm1(0, c.nc)
    y = orc.csr_spmv(irp, ja, val, x)
    # same summation order (ascending columns) -> scipy's CSR kernel agrees to rounding
    assert np.allclose(y, A @ x, rtol=1e-13, atol=1e-9 * np.abs(val).max())
    yh = orc.hll_spmv(orc.hll_from_coo(c, 32), x)
    assert np.array_equal(yh, y)  # identical per-row order -> identical bits


def test_spmv_exact_integer_poisson(orc):
    """SPD Poisson (6,-1) times ones: row sums 6 - #neighbours, exact."""
    L = 12
    c = orc.fix_coo(orc.gen_poisson3d(L))
    h = orc.hll_from_coo(c, 32)
    y = orc.hll_spmv(h, np.ones(c.nc))
    assert np.array_equal(y, 7.0 - h.irn)


def test_torture_contradicts_empty_hack_skip(orc):
    """With hksz = 1 the reference HLL skips the empty row of the torture
    case d_usmv_2_n_am3_bm0 (beta = 0) and leaves y(2) = 3, not the expected 0."""
    case = [c for c in _torture_cases() if c["name"] == "d_usmv_2_n_am3_bm0_ix1_iy1"][0]
    coo = orc.fix_coo(orc.COO(case["m"], case["k"], np.array(case["ia"], np.int32),
                              np.array(case["ja"], np.int32), np.array(case["va"])))
    h = orc.hll_from_coo(coo, 1)
    y = orc.hll_spmv(h, np.array(case["x"]), case["alpha"], case["beta"],
                     np.array(case["bcy"]), quirk_empty_hack=True)
    assert y.tolist() != case["cy"] and y[1] == case["bcy"][1]


def test_hll_empty_hack_quirk(orc):
    """The CPU HLL skips a full hack whose rows are all empty, leaving y
    untouched (psb_d_hll_csmv.f90:222-226); the quirk flag turns it off."""
    ia = np.array([1, 2], np.int32)
    ja = np.array([1, 2], np.int32)
    c = orc.fix_coo(orc.COO(8, 8, ia, ja, np.array([2.0, 3.0])))
    h = orc.hll_from_coo(c, 4)
    x = np.ones(8)
    y0 = np.full(8, 10.0)
    yq = orc.hll_spmv(h, x, 1.0, 0.5, y0, quirk_empty_hack=True)
    yc = orc.hll_spmv(h, x, 1.0, 0.5, y0, quirk_empty_hack=False)
    assert yq[4:].tolist() == [10.0] * 4
    assert yc[4:].tolist() == [5.0] * 4
    assert yq[:4].tolist() == yc[:4].tolist() == [7.0, 8.0, 5.0, 5.0]


def test_hll_csmm_matches_columns(orc):
    c = orc.fix_coo(orc.gen_convdiff(9))
    h = orc.hll_from_coo(c, 32)
    X = np.stack([orc.uniform_pm1(s, c.nc) for s in range(5)], axis=1)
    Y0 = np.stack([orc.uniform_pm1(10 + s, c.nr) for s in range(5)], axis=1)
    Y = orc.hll_spmm(h, X, 2.0, -1.0, Y0)
    for k in range(5):
        yk = orc.hll_spmv(h, X[:, k], 2.0, -1.0, Y0[:, k])
        assert np.allclose(Y[:, k], yk, rtol=1e-14, atol=1e-10)


# ------------------------------------------------------------------ BLAS-1
def test_vecoperation_known_answers(orc):
    n = 1000
    one, two = np.ones(n), np.full(n, 2.0)
    assert np.all(orc.axpby(1.0, one, 1.0, np.full(n, -2.0)) == -1.0)
    assert np.all(orc.axpby(-1.0, one, 1.0, two) == 1.0)
    assert np.all(orc.axpby(0.5, two, 1.0, np.full(n, -2.0)) == -1.0)
    assert np.all(orc.mlt(1.0, two, np.full(n, -0.5), 0.0, np.zeros(n)) == -1.0)
    assert orc.dot(two, np.full(n, 0.5)) == n
    assert orc.amax(np.full(n, -0.5)) == 0.5
    assert orc.asum(np.full(n, -0.5)) == n * 0.5
    assert orc.nrm2(np.full(n, 3.0)) == pytest.approx(3.0 * math.sqrt(n), rel=1e-15)


def test_dot_nrm2_error_bounds(orc):
    x = orc.uniform_pm1(11, 100000)
    y = orc.uniform_pm1(12, 100000)
    exact = math.fsum((x * y).tolist())
    bound = len(x) * np.finfo(float).eps * np.sum(np.abs(x * y))
    assert abs(orc.dot(x, y) - exact) <= bound
    ex2 = math.sqrt(math.fsum((x * x).tolist()))
    assert abs(orc.nrm2(x) - ex2) <= 4 * len(x) * np.finfo(float).eps * ex2
    big = np.full(10, 1e200)
    assert orc.nrm2(big) == pytest.approx(1e200 * math.sqrt(10), rel=1e-14)  # scaled: no overflow


def test_axpby_branches(orc):
    x = orc.uniform_pm1(1, 50)
    y = orc.uniform_pm1(2, 50)
    assert np.array_equal(orc.axpby(0.0, x, 0.0, np.full(50, np.nan)), np.zeros(50))
    assert np.array_equal(orc.axpby(2.0, x, 0.0, np.full(50, np.nan)), 2.0 * x)
    assert np.array_equal(orc.axpby(0.0, x, 3.0, y), 3.0 * y)
    assert np.array_equal(orc.axpby(2.0, x, 3.0, y), 2.0 * x + 3.0 * y)


def test_upd_xyz_xyzw(orc):
    x, y, z, w = (orc.uniform_pm1(s, 40) for s in range(4))
    y1, z1 = orc.upd_xyz(2.0, 3.0, 0.5, -1.0, x, y, z)
    assert np.array_equal(y1, 2.0 * x + 3.0 * y)
    assert np.array_equal(z1, 0.5 * y1 + -1.0 * z)
    y2, z2 = orc.upd_xyz(0.0, 0.0, 0.5, 2.0, x, y, z)   # gamma term dropped
    assert np.array_equal(y2, np.zeros(40)) and np.array_equal(z2, 2.0 * z)
    y3, z3, w3 = orc.xyzw(1.0, 2.0, 3.0, 4.0, 5.0, 6.0, x, y, z, w)
    assert np.array_equal(y3, 1.0 * x + 2.0 * y)
    assert np.array_equal(z3, 3.0 * y3 + 4.0 * z)
    assert np.array_equal(w3, 5.0 * z3 + 6.0 * w)
    y4, z4, w4 = orc.xyzw(1.0, 0.0, 3.0, 4.0, 5.0, 6.0, x, y, z, w)   # a zero -> no-op
    assert np.array_equal(y4, y) and np.array_equal(w4, w)


# ---------------------------------------------------------------- transposed product
def _torture_t_cases():
    with open(os.path.join(GOLDEN, "torture_usmv_t.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _torture_t_cases(), ids=lambda c: c["name"])
def test_torture_usmv_transposed(orc, case):
    """The 32 torture usmv cases with trans 't' and 'c' (same arithmetic for
    real data), through the oracle's transposed CSR product."""
    coo = orc.fix_coo(orc.COO(case["m"], case["k"], np.array(case["ia"], np.int32),
                              np.array(case["ja"], np.int32), np.array(case["va"])))
    dt = np.float32 if case["kind"] == "s" else np.float64
    irp, ja, val = orc.csr_from_coo(coo)
    y = orc.csr_spmv_t(irp, ja, val.astype(dt), np.array(case["x"], dt), case["alpha"],
                       case["beta"], np.array(case["bcy"], dt), nc=case["k"])
    assert y.tolist() == case["cy"]


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0), (2.5, -0.5), (0.0, 3.0)])
def test_csr_transposed_product(orc, alpha, beta):
    """Against scipy's A^T x within the per-entry bound (the accumulation
    into y makes the error grow with the column count), and exactly on an
    integer case."""
    coo = orc.gen_powerlaw(3000, cap=300, seed=4)
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    A = sp.csr_matrix((val, ja - 1, irp - 1), shape=(coo.nr, coo.nc))
    x = orc.uniform_pm1(5, coo.nr)
    y0 = orc.uniform_pm1(6, coo.nc)
    y = orc.csr_spmv_t(irp, ja, val, x, alpha, beta, y0, nc=coo.nc)
    ref = alpha * (A.T @ x) + beta * y0
    cnt = np.bincount(ja - 1, minlength=coo.nc)
    bound = 4 * (cnt + 2) * np.finfo(float).eps * (abs(alpha) * (abs(A).T @ np.abs(x))
                                                    + abs(beta) * np.abs(y0))
    assert np.all(np.abs(y - ref) <= bound + 1e-300)
    # integers: exact
    vi = np.round(val * 4)
    xi = np.round(x * 8)
    yi = orc.csr_spmv_t(irp, ja, vi, xi, 1.0, 0.0, None, nc=coo.nc)
    assert np.array_equal(yi, sp.csr_matrix((vi, ja - 1, irp - 1), shape=A.shape).T @ xi)
