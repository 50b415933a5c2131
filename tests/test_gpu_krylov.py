"""GPU parity of the Krylov path: psb_krylov (the reference's call sequence
over the device plug-ins), the fused device-resident solvers and the
FP32-inner / FP64-outer refinement.

Tolerances: CG iteration counts within +-2 of the oracle.  BiCGSTAB is
chaotic in rounding on these problems (tests/test_oracle_krylov.py shows a
spread of >4 iterations from ulp-level perturbations of b, with the compiled
reference's own count inside it), so its count must fall inside the oracle's
sampled rounding spread, widened by a quarter of its width plus 2.  Every solve must reach the same stopping
tolerance and agree with the oracle's solution.
"""
import json
import os

import numpy as np
import pytest

import psblas3_amd as P
from helpers import bicgstab_spread, in_spread

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

MOLDS = {"HLG": lambda: P.psb_d_hip_hlg_sparse_mat(), "CSRG": lambda: P.psb_d_hip_csrg_sparse_mat(),
         "ELG": lambda: P.psb_d_hip_elg_sparse_mat()}


def _setup(coo, mold, ptype="DIAG"):
    a = P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val, mold=mold)
    desc = P.psb_cdall(nl=coo.nr)
    prec = P.psb_dprec_type(ptype).build(a, desc)
    return a, desc, prec


def _pde3d(orc):
    ref = json.load(open(os.path.join(GOLDEN, "reference_runs.json")))["pde3d_L40_bicgstab_diag"]
    coo, rhs = orc.gen_pde3d(ref["idim"], a=ref["coefficients"]["a"], b=(0, 0, 0), c=0.0,
                             hscale=True, gflag=True)
    return ref, coo, rhs


@pytest.mark.parametrize("fmt", ["HLG", "CSRG", "ELG"])
@pytest.mark.parametrize("fused", [False, True])
def test_reference_pde3d_run(gpu, orc, fmt, fused):
    ref, coo, rhs = _pde3d(orc)
    print("case pde3d L=40 bicgstab", fmt, "fused" if fused else "drop-in", flush=True)
    a, desc, prec = _setup(coo, MOLDS[fmt]())
    b = P.vect_from(rhs)
    x = P.vect_from(np.zeros(coo.nr))
    solver = P.psb_krylov_fused if fused else P.psb_krylov
    it, err, hist = solver("BICGSTAB", a, prec, b, x, ref["eps"], desc, itmax=1000, istop=2)
    csr = orc.csr_from_coo(orc.fix_coo(coo))
    spread = bicgstab_spread(orc, csr, rhs, ref["eps"])
    assert in_spread(it, spread), (it, spread)
    assert err <= ref["eps"]
    xo, ito, erro, _ = orc.krylov("BICGSTAB", csr, rhs, eps=ref["eps"], itmax=1000)
    xs = x.get_vect()
    assert np.linalg.norm(xs - xo) <= 1e-4 * np.linalg.norm(xo)
    a.free()


@pytest.mark.parametrize("fused", [False, True])
def test_cg_poisson_L40(gpu, orc, fused):
    """CG + DIAG on the SPD Poisson L=40, b = A*ones, eps 1e-9: +-2 iterations."""
    print("case cg poisson L=40", "fused" if fused else "drop-in", flush=True)
    coo = orc.gen_poisson3d(40)
    f = orc.fix_coo(coo)
    csr = orc.csr_from_coo(f)
    bh = orc.csr_spmv(*csr, np.ones(f.nc))
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat(32))
    b = P.vect_from(bh)
    x = P.vect_from(np.zeros(f.nr))
    solver = P.psb_krylov_fused if fused else P.psb_krylov
    it, err, hist = solver("CG", a, prec, b, x, 1e-9, desc, itmax=1000, istop=2)
    xo, ito, erro, histo = orc.krylov("CG", csr, bh, eps=1e-9, itmax=1000, istop=2)
    assert abs(it - ito) <= 2, (it, ito)
    assert err <= 1e-9
    assert np.allclose(x.get_vect(), 1.0, atol=1e-6)
    if not fused:
        # residual histories agree while rounding differences are small
        k = min(len(hist), len(histo)) // 2
        assert np.allclose(hist[:k], histo[:k], rtol=1e-6)
    a.free()


def test_convdiff_L40_bicgstab(gpu, orc):
    """The CPU-runnable workload: convection-diffusion L=40, b = A*ones, x0 = 0,
    BiCGSTAB + DIAG to 1e-9, drop-in and fused."""
    print("case convdiff L=40 bicgstab", flush=True)
    coo = orc.gen_convdiff(40)
    f = orc.fix_coo(coo)
    csr = orc.csr_from_coo(f)
    bh = orc.csr_spmv(*csr, np.ones(f.nc))
    spread = bicgstab_spread(orc, csr, bh, 1e-9)
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat(32))
    for solver in (P.psb_krylov, P.psb_krylov_fused):
        x = P.vect_from(np.zeros(f.nr))
        it, err, hist = solver("BICGSTAB", a, prec, P.vect_from(bh), x, 1e-9, desc, itmax=1000)
        assert in_spread(it, spread), (solver.__name__, it, spread)
        assert err <= 1e-9
        assert np.allclose(x.get_vect(), 1.0, atol=1e-6)
    a.free()


@pytest.mark.parametrize("fmt", ["HLG", "ELG", "CSRG"])
@pytest.mark.parametrize("ptype", ["DIAG", "NONE"])
def test_fused_cg_formats(gpu, orc, fmt, ptype):
    print("case fused cg", fmt, ptype, flush=True)
    coo = orc.gen_aniso2d(64, 0.05)
    f = orc.fix_coo(coo)
    csr = orc.csr_from_coo(f)
    xt = orc.uniform_pm1(2, f.nr)
    bh = orc.csr_spmv(*csr, xt)
    a, desc, prec = _setup(coo, MOLDS[fmt](), ptype)
    x = P.vect_from(np.zeros(f.nr))
    it, err, hist = P.psb_krylov_fused("CG", a, prec, P.vect_from(bh), x, 1e-10, desc,
                                       itmax=2000, history=True)
    d = orc.diag_build(*csr) if ptype == "DIAG" else np.ones(f.nr)
    xo, ito, erro, histo = orc.krylov("CG", csr, bh, eps=1e-10, itmax=2000, d=d)
    assert abs(it - ito) <= 2, (it, ito)
    assert err <= 1e-10 and len(hist) == it
    assert hist[-1] == pytest.approx(err)
    assert np.linalg.norm(x.get_vect() - xo) <= 1e-7 * np.linalg.norm(xo)
    # the host-visible history tracks the oracle's
    k = min(it, ito) // 2
    assert np.allclose(hist[:k], histo[:k], rtol=1e-5)
    a.free()


@pytest.mark.parametrize("check_every", [1, 3, 16])
def test_fused_stops_at_the_converged_iteration(gpu, orc, check_every):
    """Iterations queued after the stop flag are no-ops: the count and the
    iterate do not depend on how often the host polls."""
    print("case fused check_every", check_every, flush=True)
    coo = orc.gen_poisson3d(20)
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat())
    bh = orc.uniform_pm1(4, coo.nr)
    out = []
    for ce in (check_every, 1):
        x = P.vect_from(np.zeros(coo.nr))
        it, err, _ = P.psb_krylov_fused("CG", a, prec, P.vect_from(bh), x, 1e-10, desc,
                                        check_every=ce)
        out.append((it, err, x.get_vect()))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])
    a.free()


def test_fused_itmax_zero_rhs_istop3(gpu, orc):
    print("case fused itmax / zero rhs / istop 3", flush=True)
    coo = orc.gen_poisson3d(10)
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat())
    x = P.vect_from(np.zeros(coo.nr))
    it, err, _ = P.psb_krylov_fused("CG", a, prec, P.vect_from(orc.uniform_pm1(1, coo.nr)), x,
                                    1e-14, desc, itmax=5)
    assert it == 5 and err > 1e-14
    x = P.vect_from(np.zeros(coo.nr))
    it, err, _ = P.psb_krylov_fused("BICGSTAB", a, prec, P.vect_from(np.zeros(coo.nr)), x,
                                    1e-9, desc)
    assert it == 0 and np.all(x.get_vect() == 0)
    f = orc.fix_coo(coo)
    csr = orc.csr_from_coo(f)
    bh = orc.csr_spmv(*csr, np.ones(f.nc))
    x = P.vect_from(np.zeros(coo.nr))
    it, err, _ = P.psb_krylov_fused("CG", a, prec, P.vect_from(bh), x, 1e-8, desc, istop=3)
    _, ito, _, _ = orc.krylov("CG", csr, bh, eps=1e-8, istop=3)
    assert abs(it - ito) <= 2 and err <= 1e-8
    with pytest.raises(P.PsbError):
        P.psb_krylov_fused("CG", a, prec, P.vect_from(bh), x, 1e-8, desc, istop=1)
    a.free()


def test_fused_single_precision(gpu, orc):
    print("case fused cg fp32", flush=True)
    coo = orc.gen_poisson3d(24)
    a, desc, prec = _setup(coo, P.psb_s_hip_hlg_sparse_mat())
    f = orc.fix_coo(coo)
    csr = orc.csr_from_coo(f)
    bh = orc.csr_spmv(*csr, np.ones(f.nc))
    x = P.vect_from(np.zeros(f.nr, np.float32), mold=P.psb_s_hip_vect())
    b = P.vect_from(bh.astype(np.float32), mold=P.psb_s_hip_vect())
    it, err, _ = P.psb_krylov_fused("CG", a, prec, b, x, 1e-5, desc)
    _, ito, _, _ = orc.krylov("CG", csr, bh, eps=1e-5)
    assert abs(it - ito) <= 3 and err <= 1e-5
    assert np.allclose(x.get_vect(), 1.0, atol=1e-3)
    a.free()


@pytest.mark.parametrize("method,variant", [("CG", "restart"), ("BICGSTAB", "restart"),
                                            ("CG", "reliable")])
def test_mixed_precision_refinement(gpu, orc, method, variant):
    """FP32 inner (1e-4) / FP64 outer to 1e-10 on the SPD Poisson:
    the same final relative residual as FP64, and the error against the
    oracle's FP64 solution of b = A x_true stated."""
    print("case mixed", method, variant, flush=True)
    L = 48
    coo = orc.gen_poisson3d(L)
    f = orc.fix_coo(coo)
    csr = orc.csr_from_coo(f)
    xt = orc.uniform_pm1(2, f.nr)
    bh = orc.csr_spmv(*csr, xt)
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat())
    x = P.vect_from(np.zeros(f.nr))
    oi, ii, err = P.psb_krylov_mixed(method, a, prec, P.vect_from(bh), x, 1e-10, desc,
                                     variant=variant)
    assert err <= 1e-10 and oi >= 2
    xo, ito, erro, _ = orc.krylov("CG", csr, bh, eps=1e-10, itmax=2000)
    xs = x.get_vect()
    # both are 1e-10-residual solutions: their distance is bounded by the
    # conditioning times the tolerance
    assert np.linalg.norm(xs - xo) <= 1e-7 * np.linalg.norm(xo)
    res = orc.csr_spmv(*csr, xs, 1.0, 0.0) - bh
    assert np.linalg.norm(res) <= 1.01e-10 * np.linalg.norm(bh)
    if variant == "reliable":
        # the recurrence is not restarted: close to the FP64 CG iteration count
        assert ii <= 1.25 * ito, (ii, ito)
    print(f"mixed {method} {variant}: outer {oi}, inner {ii}, err {err:.2e}, "
          f"|x-x_true|/|x_true| = {np.linalg.norm(xs - xt) / np.linalg.norm(xt):.2e}", flush=True)
    a.free()


def test_dropin_itmax_and_zero_rhs(gpu, orc):
    print("case drop-in itmax / zero rhs", flush=True)
    coo = orc.gen_poisson3d(10)
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat())
    x = P.vect_from(np.zeros(coo.nr))
    b = P.vect_from(orc.uniform_pm1(1, coo.nr))
    it, err, _ = P.psb_krylov("CG", a, prec, b, x, 1e-14, desc, itmax=5)
    assert it == 5 and err > 1e-14
    x = P.vect_from(np.zeros(coo.nr))
    it, err, _ = P.psb_krylov("BICGSTAB", a, prec, P.vect_from(np.zeros(coo.nr)), x, 1e-9, desc)
    assert it == 0 and np.all(x.get_vect() == 0)
    a.free()


def test_dropin_istop_1_and_3(gpu, orc):
    print("case drop-in istop 1/3", flush=True)
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
    a.free()


# FCG / GCR / RGMRES / BiCGSTAB(l) through psb_krylov on the device, against
# the oracle's restatements.  The minimal-residual methods (GCR, RGMRES) and
# FCG on an SPD matrix are smooth in rounding: +-2 iterations.  BiCGSTAB(l)
# counts advance by l; its window is the oracle's count +- max(2 l, 15%).
# BiCG (CG's recurrence with a shadow sequence through A^T) is +-2 as well;
# CGS squares the BiCG polynomial and amplifies rounding: +- max(3, 20%).
NEW_CASES = [("FCG", None, "poisson", 14), ("GCR", 8, "convdiff", 12),
             ("RGMRES", 8, "convdiff", 12), ("RGMRES", 30, "poisson", 14),
             ("GCR", 30, "convdiff", 10), ("BICGSTABL", 2, "convdiff", 12),
             ("BICGSTABL", 4, "poisson", 14), ("BICG", None, "convdiff", 12),
             ("BICG", None, "poisson", 14), ("CGS", None, "convdiff", 12),
             ("CGS", None, "poisson", 10)]


@pytest.mark.parametrize("method,irst,kind,L", NEW_CASES)
def test_more_krylov_methods(gpu, orc, method, irst, kind, L):
    print(f"case psb_krylov {method} irst={irst} {kind} L={L}", flush=True)
    coo = orc.gen_poisson3d(L) if kind == "poisson" else orc.gen_convdiff(L, conv=4.0 / np.sqrt(3.0))
    csr = orc.csr_from_coo(orc.fix_coo(coo))
    xt = orc.uniform_pm1(12, coo.nr)
    bh = orc.csr_spmv(*csr, xt)
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat())
    b = P.vect_from(bh)
    x = P.vect_from(np.zeros(coo.nr))
    it, err, hist = P.psb_krylov(method, a, prec, b, x, 1e-9, desc, itmax=2000, irst=irst)
    xo, ito, erro, histo = orc.krylov(method, csr, bh, eps=1e-9, itmax=2000, irst=irst)
    print(f"  device {it} its err {err:.3e} / oracle {ito} its err {erro:.3e}", flush=True)
    assert err <= 1e-9
    if method == "BICGSTABL":
        assert it % irst == 0
        assert abs(it - ito) <= max(2 * irst, int(0.15 * ito))
    elif method == "CGS":
        assert abs(it - ito) <= max(3, int(0.2 * ito))
    else:
        assert abs(it - ito) <= 2
        k = min(len(hist), len(histo), 10)
        assert np.allclose(hist[:k], histo[:k], rtol=1e-6)
    xs = x.get_vect()
    assert np.linalg.norm(xs - xo) <= 1e-5 * np.linalg.norm(xo)
    a.free()


def test_krylov_unknown_method_falls_back(gpu, orc):
    print("case psb_krylov unknown method -> BiCGSTAB", flush=True)
    coo = orc.gen_poisson3d(8)
    csr = orc.csr_from_coo(orc.fix_coo(coo))
    bh = orc.csr_spmv(*csr, np.ones(coo.nr))
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat())
    x1, x2 = P.vect_from(np.zeros(coo.nr)), P.vect_from(np.zeros(coo.nr))
    r1 = P.psb_krylov("NOSUCH", a, prec, P.vect_from(bh), x1, 1e-9, desc)
    r2 = P.psb_krylov("BICGSTAB", a, prec, P.vect_from(bh), x2, 1e-9, desc)
    assert r1[0] == r2[0] and np.array_equal(x1.get_vect(), x2.get_vect())
    a.free()


def test_bicg_with_bjac(gpu, orc):
    """BiCG with BJAC / ILU(0): the shadow sequence applies the factors
    transposed (w = d .* U^-T x, then L^-T w) on the device; it converges to
    the solution in fewer iterations than with DIAG."""
    print("case psb_krylov BICG + BJAC", flush=True)
    coo = orc.gen_convdiff(12, conv=4.0 / np.sqrt(3.0))
    csr = orc.csr_from_coo(orc.fix_coo(coo))
    xt = orc.uniform_pm1(21, coo.nr)
    bh = orc.csr_spmv(*csr, xt)
    a = P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val,
                          mold=P.psb_d_hip_hlg_sparse_mat(64))
    desc = P.psb_cdall(nl=coo.nr)
    its = {}
    for pt in ("DIAG", "BJAC"):
        prec = P.psb_dprec_type(pt).build(a, desc)
        x = P.vect_from(np.zeros(coo.nr))
        it, err, _ = P.psb_krylov("BICG", a, prec, P.vect_from(bh), x, 1e-10, desc, itmax=1000)
        assert err <= 1e-10
        assert np.linalg.norm(x.get_vect() - xt) <= 1e-6 * np.linalg.norm(xt)
        its[pt] = it
        prec.free()
    print(f"  DIAG {its['DIAG']} its, BJAC {its['BJAC']} its", flush=True)
    assert its["BJAC"] < its["DIAG"]
    a.free()


@pytest.mark.parametrize("method,irst", [("CG", None), ("BICGSTAB", None), ("CGS", None),
                                         ("BICG", None), ("FCG", None), ("GCR", 10),
                                         ("RGMRES", 10), ("BICGSTABL", 2)])
def test_krylov_single_precision(gpu, orc, method, irst):
    """psb_krylov over the single-precision plug-ins (psb_s_hip_hlg_sparse_mat,
    psb_s_hip_vect, DIAG in FP32: the psb_s* solvers of krylov/): every
    method reaches 1e-5 and the FP64 solution to FP32 accuracy."""
    print(f"case psb_krylov FP32 {method}", flush=True)
    spd = method in ("CG", "FCG")
    coo = orc.gen_poisson3d(10) if spd else orc.gen_convdiff(10, conv=2.0)
    csr = orc.csr_from_coo(orc.fix_coo(coo))
    xt = orc.uniform_pm1(31, coo.nr)
    bh = orc.csr_spmv(*csr, xt)
    a = P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val,
                          mold=P.psb_s_hip_hlg_sparse_mat(64))
    desc = P.psb_cdall(nl=coo.nr)
    prec = P.psb_dprec_type("DIAG").build(a, desc)
    assert prec.dv._dtype == np.float32
    b = P.vect_from(bh.astype(np.float32), mold=P.psb_s_hip_vect())
    x = P.vect_from(np.zeros(coo.nr, np.float32), mold=P.psb_s_hip_vect())
    it, err, _ = P.psb_krylov(method, a, prec, b, x, 1e-5, desc, itmax=1000, irst=irst)
    xs = x.get_vect()
    assert xs.dtype == np.float32
    print(f"  {it} its err {err:.2e} |x - xt| / |xt| "
          f"{np.linalg.norm(xs - xt) / np.linalg.norm(xt):.2e}", flush=True)
    assert err <= 1e-5 and it < 1000
    # BiCGSTAB(l) never recomputes its residual (krylov/psb_dcgstabl.f90): in
    # FP32 the recurrence drifts from the true residual (measured 4.6% error
    # in x here while the recurrence reports 1e-5); the other methods land
    # within FP32 accuracy of the solution
    tol = 1e-1 if method == "BICGSTABL" else 1e-3
    assert np.linalg.norm(xs - xt) <= tol * np.linalg.norm(xt)
    a.free()


def test_rgmres_istop1_dropin(gpu, orc):
    """RGMRES with istop 1 (normwise backward error) through psb_krylov on the
    device against the oracle: +-2 iterations, the bound met by the returned
    x; istop 3 is rejected as the reference rejects it (psb_drgmres.f90:182)."""
    print("case psb_krylov RGMRES istop 1", flush=True)
    coo = orc.gen_convdiff(12, conv=4.0 / np.sqrt(3.0))
    csr = orc.csr_from_coo(orc.fix_coo(coo))
    bh = orc.csr_spmv(*csr, orc.uniform_pm1(15, coo.nr))
    a, desc, prec = _setup(coo, P.psb_d_hip_hlg_sparse_mat())
    x = P.vect_from(np.zeros(coo.nr))
    it, err, hist = P.psb_krylov("RGMRES", a, prec, P.vect_from(bh), x, 1e-9, desc, itmax=2000,
                                 istop=1, irst=10)
    ani = P.psb_spnrmi(a, desc)
    xo, ito, erro, histo = orc.krylov("RGMRES", csr, bh, eps=1e-9, itmax=2000, istop=1,
                                      ani=ani, irst=10)
    print(f"  device {it} its err {err:.3e} / oracle {ito} its err {erro:.3e}", flush=True)
    assert abs(it - ito) <= 2 and err <= 1e-9
    k = min(len(hist), len(histo), 10)
    assert np.allclose(hist[:k], histo[:k], rtol=1e-6)
    xs = x.get_vect()
    irp, ja, val = csr
    r = bh - orc.csr_spmv(irp, ja, val, xs)
    bwe = np.abs(r).max() / (ani * np.abs(xs).max() + np.abs(bh).max())
    assert bwe <= 1e-9
    with pytest.raises(P.PsbError):
        P.psb_krylov("RGMRES", a, prec, P.vect_from(bh), P.vect_from(np.zeros(coo.nr)), 1e-9,
                     desc, istop=3)
    a.free()


@pytest.mark.parametrize("method", ["RGMRES", "GCR"])
def test_invariant_krylov_space(gpu, orc, method):
    """A = I, b = e1: the Krylov space is invariant after one step (RGMRES:
    h(2,1) = 0); both methods end with the exact solution, no division by
    zero."""
    print(f"case {method} invariant space", flush=True)
    n = 70
    ia = np.arange(1, n + 1, dtype=np.int32)
    coo = P.psb_d_coo_sparse_mat(n, n, ia, ia.copy(), np.ones(n))
    coo.fix()
    a = P.psb_dspmat_type()
    a.m = a.n = n
    a.a = P.psb_d_hip_hlg_sparse_mat().cp_from_coo(coo)
    desc = P.psb_cdall(nl=n)
    prec = P.psb_dprec_type("DIAG").build(a, desc)
    b = np.zeros(n)
    b[0] = 1.0
    for istop in ((1, 2) if method == "RGMRES" else (2,)):
        x = P.vect_from(np.zeros(n))
        it, err, _ = P.psb_krylov(method, a, prec, P.vect_from(b), x, 1e-12, desc, itmax=20,
                                  istop=istop, irst=4)
        assert it == 1 and err == 0.0, (istop, it, err)
        assert np.array_equal(x.get_vect(), b)
    a.free()
