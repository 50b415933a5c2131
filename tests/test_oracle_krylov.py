"""Pins the oracle's CG / BiCGSTAB + DIAG (krylov/psb_dcg.F90,
krylov/psb_dcgstab.f90) against the reference's own recorded run and against
scipy solutions."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import bicgstab_spread

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_reference_pde3d_run(orc):
    """psb_d_pde3d L=40, BiCGSTAB + DIAG, CSR, eps 1e-6, istop 2: the compiled
    reference took 80 iterations to 9.197e-7; the restatement must land within
    +-2 iterations and reach the same tolerance."""
    ref = json.load(open(os.path.join(GOLDEN, "reference_runs.json")))["pde3d_L40_bicgstab_diag"]
    a = ref["coefficients"]["a"]
    coo, rhs = orc.gen_pde3d(ref["idim"], a=a, b=(0, 0, 0), c=0.0, hscale=True, gflag=True)
    f = orc.fix_coo(coo)
    assert (f.nr, f.nz) == (ref["n"], ref["nnz"])
    x, it, err, hist = orc.krylov("BICGSTAB", orc.csr_from_coo(f), rhs, eps=ref["eps"],
                                  itmax=1000, istop=ref["istop"])
    assert abs(it - ref["iterations"]) <= 2
    assert err <= ref["eps"]
    assert err == pytest.approx(ref["final_error"], rel=0.25)


def _poisson(orc, L):
    f = orc.fix_coo(orc.gen_poisson3d(L))
    irp, ja, val = orc.csr_from_coo(f)
    return (irp, ja, val), sp.csr_matrix((val, ja - 1, irp - 1), shape=(f.nr, f.nc))


@pytest.mark.parametrize("method", ["CG", "BICGSTAB"])
def test_solution_vs_direct(orc, method):
    csr, A = _poisson(orc, 10)
    xt = orc.uniform_pm1(2, A.shape[0])
    b = A @ xt
    x, it, err, hist = orc.krylov(method, csr, b, eps=1e-12, itmax=1000)
    assert err <= 1e-12
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    # istop=2 error estimate is ||b - A x|| / ||b||
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) == pytest.approx(err, rel=1e-3, abs=1e-15)


def test_cg_energy_norm_monotone(orc):
    """CG minimises the A-norm of the error: it decreases every iteration."""
    csr, A = _poisson(orc, 8)
    xt = orc.uniform_pm1(3, A.shape[0])
    b = A @ xt
    errs = []
    for k in range(1, 25):
        x, it, _, _ = orc.krylov("CG", csr, b, eps=0.0, itmax=k)
        e = x - xt
        errs.append(float(e @ (A @ e)))
    assert all(e2 < e1 for e1, e2 in zip(errs, errs[1:]))


def test_cg_matches_scipy_iterations(orc):
    """Jacobi-preconditioned CG: iteration count equals scipy's cg with the
    same M = D^-1 and relative tolerance within +-2."""
    L = 14
    f = orc.fix_coo(orc.gen_convdiff(L))   # nonsymmetric -> use the SPD Poisson instead
    csr, A = _poisson(orc, L)
    b = A @ np.ones(A.shape[0])
    x, it, err, hist = orc.krylov("CG", csr, b, eps=1e-9, itmax=1000)
    d = 1.0 / A.diagonal()
    count = [0]

    def cb(_):
        count[0] += 1
    spla.cg(A, b, rtol=1e-9, atol=0.0, M=sp.diags(d), maxiter=1000, callback=cb)
    assert abs(it - count[0]) <= 2
    del f


def test_istop_variants(orc):
    csr, A = _poisson(orc, 6)
    b = A @ np.ones(A.shape[0])
    for istop in (1, 2, 3):
        ani = float(abs(A).sum(axis=1).max())
        x, it, err, _ = orc.krylov("BICGSTAB", csr, b, eps=1e-10, istop=istop, ani=ani)
        assert err <= 1e-10 and it > 0


def test_bicgstab_rounding_spread_pde3d(orc):
    """BiCGSTAB on the pde3d sample is sensitive to rounding: perturbing b by a
    few ulps moves the iteration count by more than +-2, and the compiled
    reference's 80 lies inside that spread.  GPU BiCGSTAB parity is therefore
    checked against this spread, CG against +-2."""
    ref = json.load(open(os.path.join(GOLDEN, "reference_runs.json")))["pde3d_L40_bicgstab_diag"]
    coo, rhs = orc.gen_pde3d(40, a=ref["coefficients"]["a"], b=(0, 0, 0), c=0.0, hscale=True,
                             gflag=True)
    csr = orc.csr_from_coo(orc.fix_coo(coo))
    lo, hi = bicgstab_spread(orc, csr, rhs, ref["eps"])
    assert lo <= ref["iterations"] <= hi
    assert hi - lo > 4
