"""The oracle's CSR triangular solve (psb_d_csr_cssv, base/serial/impl/
psb_d_csr_impl.F90:782-1039) pinned by the reference's torture known answers
(test/torture/psb_{d,s}_mvsv_tester.f90, every ussv case, extracted to
tests/golden/torture_ussv.json), dense triangular solves (scipy) for every
triangle / diagonal / trans combination, exact integer solutions, and the
alpha / beta rules of the reference's wrapper."""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

from helpers import dense, tri_csr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _ussv_cases():
    with open(os.path.join(GOLDEN, "torture_ussv.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _ussv_cases(), ids=lambda c: c["name"])
def test_torture_ussv(orc, case):
    coo = orc.fix_coo(orc.COO(case["m"], case["k"], np.array(case["ia"], np.int32),
                              np.array(case["ja"], np.int32), np.array(case["va"])))
    dt = np.float32 if case["kind"] == "s" else np.float64
    irp, ja, val = orc.csr_from_coo(coo)
    y = orc.csr_cssv(irp, ja, val.astype(dt), np.array(case["x"], dt), lower=case["lower"],
                     unit=case["unit"], trans=case["trans"], alpha=case["alpha"],
                     beta=case["beta"], y=np.array(case["bcy"], dt))
    assert y.tolist() == case["cy"]


@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_cssv_dense(orc, lower, unit, trans, dt):
    n = 60
    irp, ja, val = tri_csr(orc, n, lower, unit, seed=7)
    x = orc.uniform_pm1(3, n).astype(dt)
    y = orc.csr_cssv(irp, ja, val.astype(dt), x, lower=lower, unit=unit, trans=trans)
    A = dense(irp, ja, val, n)
    if unit:
        A = A + np.eye(n)
    ref = sla.solve_triangular(A, x.astype(np.float64), lower=lower, unit_diagonal=unit,
                               trans=0 if trans == "N" else 1)
    tol = 1e-12 if dt == np.float64 else 2e-5
    assert y.dtype == dt
    assert np.allclose(y, ref, rtol=tol, atol=tol * np.abs(ref).max())


@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_cssv_integer_exact(orc, lower, unit, trans):
    """Integer triangle with diagonal +-1 and integer solution: every
    intermediate is an exactly representable integer, so the solve must
    return x_true exactly."""
    n = 40
    irp, ja, val = tri_csr(orc, n, lower, unit, seed=11, density=0.1, integer=True)
    A = dense(irp, ja, val, n) + (np.eye(n) if unit else 0)
    xt = np.random.default_rng(5).integers(-3, 4, n).astype(float)
    b = (A if trans == "N" else A.T) @ xt
    assert np.abs(b).max() < 2 ** 40
    y = orc.csr_cssv(irp, ja, val, b, lower=lower, unit=unit, trans=trans)
    assert np.array_equal(y, xt)


def test_cssv_alpha_beta_rules(orc):
    """alpha == 0 gives beta*y without solving; beta == 0 scales the solution
    in place (alpha == -1 by negation); otherwise alpha*sol + beta*y through
    the library axpby, each product and the sum rounded separately."""
    n = 50
    irp, ja, val = tri_csr(orc, n, True, False, seed=3)
    x = orc.uniform_pm1(1, n)
    y0 = orc.uniform_pm1(2, n)
    sol = orc.csr_cssv(irp, ja, val, x)
    nanx = np.full(n, np.nan)
    assert np.array_equal(orc.csr_cssv(irp, ja, val, nanx, alpha=0.0, beta=0.5, y=y0), 0.5 * y0)
    assert np.array_equal(orc.csr_cssv(irp, ja, val, nanx, alpha=0.0, beta=0.0, y=y0),
                          np.zeros(n))
    assert np.array_equal(orc.csr_cssv(irp, ja, val, x, alpha=-1.0), -sol)
    assert np.array_equal(orc.csr_cssv(irp, ja, val, x, alpha=2.5), 2.5 * sol)
    assert np.array_equal(orc.csr_cssv(irp, ja, val, x, alpha=-3.0, beta=0.5, y=y0),
                          -3.0 * sol + 0.5 * y0)
    assert np.array_equal(orc.csr_cssv(irp, ja, val, x, alpha=1.0, beta=1.0, y=y0), sol + y0)


def test_cssv_c_equals_t(orc):
    irp, ja, val = tri_csr(orc, 45, False, False, seed=9)
    x = orc.uniform_pm1(4, 45)
    a = orc.csr_cssv(irp, ja, val, x, lower=False, trans="T")
    b = orc.csr_cssv(irp, ja, val, x, lower=False, trans="C")
    assert np.array_equal(a.view(np.int64), b.view(np.int64))
