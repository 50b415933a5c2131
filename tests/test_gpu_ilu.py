"""BJAC / ILU(0) on the device: the levelled triangular solves equal the
oracle's sequential solves bit for bit (same per-row summation order, no
fused multiply-add), and psb_krylov with BJAC follows the oracle."""
import numpy as np
import pytest

import psblas3_amd as P

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def _setup(orc, coo, ptype="BJAC"):
    a = P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val,
                          mold=P.psb_d_hip_hlg_sparse_mat(64))
    desc = P.psb_cdall(nl=coo.nr)
    prec = P.psb_dprec_type(ptype).build(a, desc)
    csr = orc.csr_from_coo(orc.fix_coo(coo))
    return a, desc, prec, csr


CASES = {"convdiff12": lambda o: o.gen_convdiff(12, conv=3.0),
         "aniso33": lambda o: o.gen_aniso2d(33, 0.02),
         "poisson10": lambda o: o.gen_poisson3d(10)}


@pytest.mark.parametrize("name", list(CASES))
def test_ilu_apply_bitexact(gpu, orc, name):
    _say(f"case BJAC/ILU(0) apply {name}")
    coo = CASES[name](orc)
    a, desc, prec, csr = _setup(orc, coo)
    f = orc.ilu0_fact(*csr)
    x = orc.uniform_pm1(6, coo.nr)
    y = P.vect_from(np.zeros(coo.nr))
    prec.apply(P.vect_from(x), y, desc)
    ref = orc.bjac_apply(f, x)
    assert np.array_equal(y.get_vect().view(np.int64), ref.view(np.int64))
    # alpha / beta: y = 2 M^-1 x - 0.5 y0 (the axpby of the CPU path)
    y0 = orc.uniform_pm1(7, coo.nr)
    y2 = P.vect_from(y0.copy())
    prec.apply(P.vect_from(x), y2, desc, alpha=2.0, beta=-0.5)
    assert np.array_equal(y2.get_vect(), orc.axpby(2.0, ref, -0.5, y0.copy()))
    if name == "convdiff12":
        assert prec.levels == (34, 34)          # 3 (L - 1) + 1 wavefronts each way
    prec.free()
    a.free()


@pytest.mark.parametrize("method,kind", [("BICGSTAB", "convdiff12"), ("CG", "poisson10"),
                                         ("RGMRES", "convdiff12"), ("BICGSTABL", "aniso33")])
def test_krylov_with_bjac(gpu, orc, method, kind):
    _say(f"case psb_krylov {method} + BJAC on {kind}")
    coo = CASES[kind](orc)
    a, desc, prec, csr = _setup(orc, coo)
    xt = orc.uniform_pm1(8, coo.nr)
    bh = orc.csr_spmv(*csr, xt)
    x = P.vect_from(np.zeros(coo.nr))
    irst = 2 if method == "BICGSTABL" else None
    it, err, _ = P.psb_krylov(method, a, prec, P.vect_from(bh), x, 1e-9, desc, itmax=1000, irst=irst)
    xo, ito, erro, _ = orc.krylov_bjac(method, csr, bh, eps=1e-9, itmax=1000, irst=irst)
    _say(f"  device {it} its / oracle {ito} its")
    assert err <= 1e-9
    assert abs(it - ito) <= (2 * (irst or 1))
    assert np.linalg.norm(x.get_vect() - xo) <= 1e-6 * np.linalg.norm(xo)
    prec.free()
    a.free()


def test_device_resident_solvers_refuse_bjac(gpu, orc):
    _say("case fused / mixed / batch refuse BJAC")
    coo = orc.gen_poisson3d(6)
    a, desc, prec, csr = _setup(orc, coo)
    b = P.vect_from(np.ones(coo.nr))
    x = P.vect_from(np.zeros(coo.nr))
    with pytest.raises(P.PsbError):
        P.psb_krylov_fused("CG", a, prec, b, x, 1e-9, desc)
    with pytest.raises(P.PsbError):
        P.psb_krylov_mixed("CG", a, prec, b, x, 1e-9, desc)
    eng = P.psb_hip_batch("CG", max_systems=2, arena_bytes=1 << 20)
    with pytest.raises(P.PsbError):
        eng.register(a, prec)
    eng.free()
    prec.free()
    a.free()
