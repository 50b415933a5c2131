"""GPU parity of psb_krylov (CG, BiCGSTAB with DIAG) through the device
plug-ins: iteration counts within +-2 of the oracle / the reference run and
the same final relative residual."""
import json
import os

import numpy as np
import pytest

import psblas3_amd as P

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _setup(coo, mold):
    a = P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val, mold=mold)
    desc = P.psb_cdall(nl=coo.nr)
    prec = P.psb_dprec_type("DIAG").build(a, desc)
    return a, desc, prec


@pytest.mark.parametrize("fmt", ["HLG", "CSRG", "ELG"])
def test_reference_pde3d_run(gpu, orc, fmt):
    ref = json.load(open(os.path.join(GOLDEN, "reference_runs.json")))["pde3d_L40_bicgstab_diag"]
    print("case pde3d L=40 bicgstab", fmt, flush=True)
    coo, rhs = orc.gen_pde3d(ref["idim"], a=ref["coefficients"]["a"], b=(0, 0, 0), c=0.0,
                             hscale=True, gflag=True)
    mold = {"HLG": P.psb_d_hip_hlg_sparse_mat(), "CSRG": P.psb_d_hip_csrg_sparse_mat(),
            "ELG": P.psb_d_hip_elg_sparse_mat()}[fmt]
    a, desc, prec = _setup(coo, mold)
    b = P.vect_from(rhs)
    x = P.vect_from(np.zeros(coo.nr))
    it, err, hist = P.psb_krylov("BICGSTAB", a, prec, b, x, ref["eps"], desc, itmax=1000, istop=2)
    assert abs(it - ref["iterations"]) <= 2, it
    assert err <= ref["eps"]
    xo, ito, erro, _ = orc.krylov("BICGSTAB", orc.csr_from_coo(orc.fix_coo(coo)), rhs,
                                  eps=ref["eps"], itmax=1000, istop=2)
    assert abs(it - ito) <= 2
    xs = x.get_vect()
    assert np.linalg.norm(xs - xo) <= 1e-4 * np.linalg.norm(xo)


@pytest.mark.parametrize("method", ["CG", "BICGSTAB"])
def test_convdiff_L40_vs_oracle(gpu, orc, method):
    """The CPU-runnable workload: convection-diffusion L=40, b = A*ones,
    x0 = 0, eps 1e-9 (CG runs on the SPD Poisson instead)."""
    print("case L=40", method, flush=True)
    coo = orc.gen_convdiff(40) if method == "BICGSTAB" else orc.gen_poisson3d(40)
    f = orc.fix_coo(coo)
    csr = orc.csr_from_coo(f)
    bh = orc.csr_spmv(*csr, np.ones(f.nc))
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat(32))
    b = P.vect_from(bh)
    x = P.vect_from(np.zeros(f.nr))
    it, err, hist = P.psb_krylov(method, a, prec, b, x, 1e-9, desc, itmax=1000, istop=2)
    xo, ito, erro, histo = orc.krylov(method, csr, bh, eps=1e-9, itmax=1000, istop=2)
    assert abs(it - ito) <= 2, (it, ito)
    assert err <= 1e-9
    assert np.allclose(x.get_vect(), 1.0, atol=1e-5)
    # residual histories track each other until round-off separates them
    k = min(len(hist), len(histo)) // 2
    assert np.allclose(hist[:k], histo[:k], rtol=1e-6)


def test_istop_1_and_3(gpu, orc):
    print("case istop 1/3", flush=True)
    coo = orc.gen_poisson3d(12)
    f = orc.fix_coo(coo)
    csr = orc.csr_from_coo(f)
    bh = orc.csr_spmv(*csr, np.ones(f.nc))
    a, desc, prec = _setup(coo, P.psb_d_hip_csrg_sparse_mat())
    for istop in (1, 3):
        x = P.vect_from(np.zeros(f.nr))
        it, err, _ = P.psb_krylov("CG", a, prec, P.vect_from(bh), x, 1e-8, desc, istop=istop)
        ani = float(np.abs(orc.csr_spmv(csr[0], csr[1], np.abs(csr[2]), np.ones(f.nc))).max())
        _, ito, _, _ = orc.krylov("CG", csr, bh, eps=1e-8, istop=istop, ani=ani)
        assert abs(it - ito) <= 2 and err <= 1e-8


def test_itmax_and_zero_rhs(gpu, orc):
    print("case itmax / zero rhs", flush=True)
    coo = orc.gen_poisson3d(10)
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat())
    x = P.vect_from(np.zeros(coo.nr))
    b = P.vect_from(orc.uniform_pm1(1, coo.nr))
    it, err, _ = P.psb_krylov("CG", a, prec, b, x, 1e-14, desc, itmax=5)
    assert it == 5 and err > 1e-14
    x = P.vect_from(np.zeros(coo.nr))
    it, err, _ = P.psb_krylov("BICGSTAB", a, prec, P.vect_from(np.zeros(coo.nr)), x, 1e-9, desc)
    assert it == 0 and np.all(x.get_vect() == 0)
