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


def _mixed_levels(orc):
    """n = 20000: rows < 10000 have no lower entries (a level 0 of 10000 rows,
    beyond one workgroup), rows >= 10000 depend on i - 10000 and on i - 1 in
    chains of 50 (levels of about 200 rows), and every 997th row has 12
    lower entries (levels too wide for a workgroup run); the pattern is
    symmetric so U mirrors it.  Diagonally dominant."""
    n = 20000
    rng = np.random.default_rng(17)
    pairs = set()
    for i in range(10000, n):
        pairs.add((i, i - 10000))
        if i % 50:
            pairs.add((i, i - 1))
    for i in range(10500, n, 997):
        for j in rng.choice(i - 1, 12, replace=False):
            pairs.add((i, int(j)))
    ia, ja, val = [], [], []
    deg = np.zeros(n)
    for i, j in pairs:
        for r, c in ((i, j), (j, i)):
            v = -0.5 - rng.random()
            ia.append(r + 1); ja.append(c + 1); val.append(v)
            deg[r] += abs(v)
    for i in range(n):
        ia.append(i + 1); ja.append(i + 1); val.append(deg[i] + 1.0)
    return orc.COO(n, n, np.array(ia, np.int32), np.array(ja, np.int32), np.array(val))


def _band6(orc):
    """n = 3000, lower and upper offsets 1, 2, 3, 5, 8, 13 (six entries per
    side: the widest run template), nonsymmetric values, diagonally dominant;
    every level is one row."""
    n = 3000
    rng = np.random.default_rng(23)
    ia, ja, val = [], [], []
    for i in range(n):
        s = 0.0
        for o in (1, 2, 3, 5, 8, 13):
            for j in (i - o, i + o):
                if 0 <= j < n:
                    v = rng.uniform(-1.0, 1.0)
                    ia.append(i + 1); ja.append(j + 1); val.append(v)
                    s += abs(v)
        ia.append(i + 1); ja.append(i + 1); val.append(s + 0.5)
    return orc.COO(n, n, np.array(ia, np.int32), np.array(ja, np.int32), np.array(val))


CASES = {"convdiff12": lambda o: o.gen_convdiff(12, conv=3.0),
         "aniso33": lambda o: o.gen_aniso2d(33, 0.02),
         "poisson10": lambda o: o.gen_poisson3d(10),
         "poisson40": lambda o: o.gen_poisson3d(40),
         "mixed_levels": _mixed_levels,
         "band6": _band6}


@pytest.fixture(params=[0, 1], ids=["workgroup-runs", "launch-per-level"])
def trsv_variant(request):
    from psblas3_amd import _native as N
    old = N.fn("psb_hip_ilu_set_variant")(request.param)
    yield request.param
    N.fn("psb_hip_ilu_set_variant")(old)


@pytest.mark.parametrize("name", list(CASES))
def test_ilu_apply_bitexact(gpu, orc, name, trsv_variant):
    _say(f"case BJAC/ILU(0) apply {name} variant {trsv_variant}")
    coo = CASES[name](orc)
    a, desc, prec, csr = _setup(orc, coo)
    if trsv_variant == 1:
        assert prec.segments == (0, 0)
    else:
        assert min(prec.segments) >= 1
    _say(f"  levels {prec.levels} workgroup runs {prec.segments}")
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


@pytest.mark.parametrize("name", ["convdiff12", "band6", "poisson40"])
def test_bjac_transposed_apply_bitexact(gpu, orc, name):
    """prec%apply with trans 'T' / 'C' (psb_d_bjac_apply_vect): w = d .* U^-T x
    (psb_spsm scale 'L'), then y = alpha L^-T w + beta y, equal to the same
    composition of the oracle's CSR solves bit for bit."""
    _say(f"case BJAC transposed apply {name}")
    coo = CASES[name](orc)
    a, desc, prec, csr = _setup(orc, coo)
    f = orc.ilu0_fact(*csr)
    x = orc.uniform_pm1(16, coo.nr)
    y0 = orc.uniform_pm1(17, coo.nr)
    w = orc.csr_cssv(f.uirp, f.uja, f.uval, x, lower=False, unit=True, trans="T") * f.d
    for trans, alpha, beta in (("T", 1.0, 0.0), ("C", 2.0, -0.5)):
        y = P.vect_from(y0.copy())
        prec.apply(P.vect_from(x), y, desc, trans=trans, alpha=alpha, beta=beta)
        ref = orc.csr_cssv(f.lirp, f.lja, f.lval, w, lower=True, unit=True, trans="T",
                           alpha=alpha, beta=beta, y=y0)
        assert np.array_equal(y.get_vect().view(np.int64), ref.view(np.int64)), trans
    prec.free()
    a.free()


@pytest.mark.parametrize("name", ["convdiff12", "band6"])
def test_bjac_single_precision(gpu, orc, name):
    """BJAC on an FP32 matrix: FP32 ILU(0) factors (psb_silu0_fact) applied
    through the device CSR solves, op 'N' and 'T', equal to the oracle's FP32
    composition bit for bit; BiCGSTAB with it converges faster than with DIAG."""
    _say(f"case BJAC FP32 {name}")
    coo = CASES[name](orc)
    f_ = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f_)
    a = P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val,
                          mold=P.psb_s_hip_hlg_sparse_mat(64))
    desc = P.psb_cdall(nl=coo.nr)
    prec = P.psb_dprec_type("BJAC").build(a, desc)
    f = orc.ilu0_fact(irp, ja, val.astype(np.float32))
    vm = P.psb_s_hip_vect()
    x = orc.uniform_pm1(18, coo.nr).astype(np.float32)
    y0 = orc.uniform_pm1(19, coo.nr).astype(np.float32)
    for trans in ("N", "T"):
        y = P.vect_from(y0.copy(), mold=vm)
        prec.apply(P.vect_from(x, mold=vm), y, desc, trans=trans, alpha=1.5, beta=-0.5)
        if trans == "N":
            w = orc.csr_cssv(f.lirp, f.lja, f.lval, x, lower=True, unit=True) * f.d
            ref = orc.csr_cssv(f.uirp, f.uja, f.uval, w, lower=False, unit=True, alpha=1.5,
                               beta=-0.5, y=y0)
        else:
            w = orc.csr_cssv(f.uirp, f.uja, f.uval, x, lower=False, unit=True, trans="T") * f.d
            ref = orc.csr_cssv(f.lirp, f.lja, f.lval, w, lower=True, unit=True, trans="T",
                               alpha=1.5, beta=-0.5, y=y0)
        got = y.get_vect()
        assert got.dtype == np.float32
        assert np.array_equal(got.view(np.int32), ref.view(np.int32)), trans
    xt = orc.uniform_pm1(20, coo.nr)
    bh = orc.csr_spmv(irp, ja, val, xt).astype(np.float32)
    its = {}
    for pt, pr in (("BJAC", prec), ("DIAG", P.psb_dprec_type("DIAG").build(a, desc))):
        xs = P.vect_from(np.zeros(coo.nr, np.float32), mold=vm)
        it, err, _ = P.psb_krylov("BICGSTAB", a, pr, P.vect_from(bh, mold=vm), xs, 1e-5, desc,
                                  itmax=1000)
        assert err <= 1e-5
        its[pt] = it
        pr.free()
    _say(f"  BiCGSTAB FP32 its: BJAC {its['BJAC']}, DIAG {its['DIAG']}")
    assert its["BJAC"] < its["DIAG"]
    a.free()
