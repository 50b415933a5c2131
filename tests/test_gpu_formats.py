"""GPU parity: device COO -> HLL / ELL / CSR conversions (bit-exact against
the oracle) and SpMV in FP64 / FP32 (per-row bound 4*nnz_row*eps*(|A||x|)_i).
All device work goes through the C ABI (libpsb_hip.so)."""
import numpy as np
import pytest

import psblas3_amd as P

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _coo_cases(orc):
    ia = np.array([1, 1, 3, 4, 4, 4, 5, 4], np.int32)
    ja = np.array([1, 3, 2, 4, 1, 5, 5, 4], np.int32)      # (4,4) duplicated
    val = np.array([1, 2, 3, 4, 5, 6, 7, 0.5])
    return {
        "hand5": orc.COO(5, 5, ia, ja, val),
        "convdiff_L13": orc.gen_convdiff(13),
        "convdiff_L40": orc.gen_convdiff(40),
        "powerlaw_3000": orc.gen_powerlaw(3000, cap=300, seed=1),
        "aniso_L33": orc.gen_aniso2d(33, 0.01),
    }


def _mat(coo, mold):
    return P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val, mold=mold)


@pytest.mark.parametrize("name", ["hand5", "convdiff_L13", "convdiff_L40", "powerlaw_3000",
                                  "aniso_L33"])
@pytest.mark.parametrize("hksz", [32, 64, 7])
def test_hll_conversion_bitexact(gpu, orc, name, hksz):
    _say("case", name, "hksz", hksz)
    coo = _coo_cases(orc)[name]
    a = _mat(coo, P.psb_d_hip_hlg_sparse_mat(hksz=hksz))
    hk, irn, ja, val, idiag = a.a.to_host()
    h = orc.hll_from_coo(orc.fix_coo(coo), hksz)
    assert np.array_equal(hk, h.hkoffs)
    assert np.array_equal(irn, h.irn)
    assert np.array_equal(ja, h.ja)
    assert np.array_equal(val.view(np.int64), h.val.view(np.int64))
    assert np.array_equal(idiag, h.idiag)
    a.free()


@pytest.mark.parametrize("name", ["hand5", "convdiff_L13", "powerlaw_3000"])
def test_ell_conversion_bitexact(gpu, orc, name):
    _say("case", name)
    coo = _coo_cases(orc)[name]
    a = _mat(coo, P.psb_d_hip_elg_sparse_mat())
    irn, ja, val, idiag = a.a.to_host()
    e = orc.ell_from_coo(orc.fix_coo(coo))
    assert a.a.pitch() == ((coo.nr + 31) // 32) * 32
    assert np.array_equal(irn, e.irn)
    assert np.array_equal(ja, e.ja)
    assert np.array_equal(val.view(np.int64), e.val.view(np.int64))
    assert np.array_equal(idiag, e.idiag)
    a.free()


@pytest.mark.parametrize("name", ["hand5", "convdiff_L13", "powerlaw_3000"])
def test_csr_conversion_bitexact(gpu, orc, name):
    _say("case", name)
    coo = _coo_cases(orc)[name]
    a = _mat(coo, P.psb_d_hip_csrg_sparse_mat())
    irp, ja, val = a.a.to_host()
    oirp, oja, oval = orc.csr_from_coo(orc.fix_coo(coo))
    assert np.array_equal(irp, oirp) and np.array_equal(ja, oja)
    assert np.array_equal(val.view(np.int64), oval.view(np.int64))
    a.free()


def _row_bound(coo, x, alpha, beta, y0, eps):
    nr = coo.nr
    absax = np.zeros(nr)
    np.add.at(absax, coo.ia - 1, np.abs(coo.val) * np.abs(x[coo.ja - 1]))
    nnz_row = np.bincount(coo.ia - 1, minlength=nr)
    return 4.0 * (nnz_row + 1) * eps * (abs(alpha) * absax + abs(beta) * np.abs(y0)) + 1e-300


MOLDS = {
    "hll32": lambda: P.psb_d_hip_hlg_sparse_mat(hksz=32),
    "hll64": lambda: P.psb_d_hip_hlg_sparse_mat(hksz=64),
    "ell": lambda: P.psb_d_hip_elg_sparse_mat(),
    "csr": lambda: P.psb_d_hip_csrg_sparse_mat(),
}
SMOLDS = {
    "hll32": lambda: P.psb_s_hip_hlg_sparse_mat(hksz=32),
    "hll64": lambda: P.psb_s_hip_hlg_sparse_mat(hksz=64),
    "ell": lambda: P.psb_s_hip_elg_sparse_mat(),
    "csr": lambda: P.psb_s_hip_csrg_sparse_mat(),
}


@pytest.mark.parametrize("fmt", list(MOLDS))
@pytest.mark.parametrize("name", ["hand5", "convdiff_L40", "powerlaw_3000", "aniso_L33"])
@pytest.mark.parametrize("ab", [(1.0, 0.0), (-2.5, 0.75), (0.0, 2.0), (1.0, 1.0)])
def test_spmv_fp64(gpu, orc, fmt, name, ab):
    alpha, beta = ab
    _say("case", fmt, name, alpha, beta)
    coo = _coo_cases(orc)[name]
    f = orc.fix_coo(coo)
    a = _mat(coo, MOLDS[fmt]())
    x = orc.uniform_pm1(1, coo.nc)
    y0 = orc.uniform_pm1(9, coo.nr)
    xv = P.vect_from(x)
    yv = P.vect_from(y0)
    P.psb_spmm(alpha, a, xv, beta, yv)
    y = yv.get_vect()
    irp, ja, val = orc.csr_from_coo(f)
    yref = orc.csr_spmv(irp, ja, val, x, alpha, beta, y0)
    bound = _row_bound(f, x, alpha, beta, y0, np.finfo(np.float64).eps)
    assert np.all(np.abs(y - yref) <= bound), np.max(np.abs(y - yref) - bound)
    a.free()


@pytest.mark.parametrize("fmt", list(SMOLDS))
@pytest.mark.parametrize("name", ["convdiff_L40", "powerlaw_3000"])
def test_spmv_fp32(gpu, orc, fmt, name):
    _say("case", fmt, name)
    coo = _coo_cases(orc)[name]
    f = orc.fix_coo(coo)
    a = _mat(coo, SMOLDS[fmt]())
    x = orc.uniform_pm1(1, coo.nc).astype(np.float32)
    y0 = orc.uniform_pm1(2, coo.nr).astype(np.float32)
    xv = P.vect_from(x, mold=P.psb_s_hip_vect())
    yv = P.vect_from(y0, mold=P.psb_s_hip_vect())
    P.psb_spmm(1.5, a, xv, -0.5, yv)
    y = yv.get_vect().astype(np.float64)
    irp, ja, val = orc.csr_from_coo(f)
    yref = orc.csr_spmv(irp, ja, val.astype(np.float32), x, 1.5, -0.5, y0).astype(np.float64)
    # FP32 oracle itself carries ~nnz_row*eps32 error: allow twice the bound
    bound = 2 * _row_bound(f, x.astype(np.float64), 1.5, -0.5, y0.astype(np.float64),
                           np.finfo(np.float32).eps)
    assert np.all(np.abs(y - yref) <= bound)
    a.free()


def test_spmv_poisson_exact(gpu, orc):
    """Integer entries, x = ones: every row sum is exact in any order."""
    _say("case poisson L=30 exact")
    coo = orc.gen_poisson3d(30)
    for mold in (P.psb_d_hip_hlg_sparse_mat(32), P.psb_d_hip_elg_sparse_mat(),
                 P.psb_d_hip_csrg_sparse_mat()):
        a = _mat(coo, mold)
        xv = P.vect_from(np.ones(coo.nc))
        yv = P.vect_from(np.zeros(coo.nr))
        P.psb_spmm(1.0, a, xv, 0.0, yv)
        irn = np.bincount(coo.ia - 1, minlength=coo.nr)
        assert np.array_equal(yv.get_vect(), 7.0 - irn)
        a.free()


@pytest.mark.parametrize("k,kind", [(2, "stencil"), (5, "stencil"), (8, "powerlaw"),
                                    (33, "stencil")])
def test_spmv_multicolumn(gpu, orc, k, kind):
    _say(f"case multicolumn hll spmm k={k} {kind}")
    coo = orc.gen_convdiff(17) if kind == "stencil" else orc.gen_powerlaw(3000, cap=150, seed=7)
    f = orc.fix_coo(coo)
    hk = 32 if kind == "stencil" else 64
    h = orc.hll_from_coo(f, hk)
    a = _mat(coo, P.psb_d_hip_hlg_sparse_mat(hk))
    X = np.stack([orc.uniform_pm1(s, coo.nc) for s in range(k)], axis=1)
    Y0 = np.stack([orc.uniform_pm1(20 + s, coo.nr) for s in range(k)], axis=1)
    xv = P.vect_from(X)
    yv = P.vect_from(Y0)
    P.psb_spmm(2.0, a, xv, -1.0, yv)
    Y = yv.get_vect()
    Yref = orc.hll_spmm(h, X, 2.0, -1.0, Y0)
    for col in range(k):
        bound = _row_bound(f, X[:, col], 2.0, -1.0, Y0[:, col], np.finfo(float).eps)
        assert np.all(np.abs(Y[:, col] - Yref[:, col]) <= bound)
    a.free()


@pytest.mark.parametrize("k", [2, 3, 5, 8, 9, 16, 33])
@pytest.mark.parametrize("hk,kind,idx16", [(64, "stencil", 1), (64, "stencil", 0),
                                           (32, "stencil", 1), (64, "powerlaw", 1)])
def test_spmm_columns_match_spmv(gpu, orc, k, hk, kind, idx16):
    """Every column of the multi-RHS product (two rows per lane up to 16
    columns, in groups of 8; one row per lane beyond) equals the single-vector
    product of that column bit for bit: both sum each row's slots in slot order from zero and
    apply alpha and beta the same way.  n is odd (the last lane of the pair
    kernel has one row)."""
    from psblas3_amd import _native as N
    _say(f"case spmm columns vs spmv k={k} hk={hk} {kind} idx16={idx16}")
    coo = orc.gen_convdiff(13) if kind == "stencil" else orc.gen_powerlaw(3001, cap=90, seed=5)
    assert coo.nr % 2 == 1
    old = N.fn("psb_hip_hll_set_index16")(idx16)
    try:
        a = _mat(coo, P.psb_d_hip_hlg_sparse_mat(hk))
        X = np.stack([orc.uniform_pm1(40 + s, coo.nc) for s in range(k)], axis=1)
        Y0 = np.stack([orc.uniform_pm1(80 + s, coo.nr) for s in range(k)], axis=1)
        yv = P.vect_from(Y0)
        P.psb_spmm(2.0, a, P.vect_from(X), -1.0, yv)
        Y = yv.get_vect()
        for col in range(k):
            y1 = P.vect_from(Y0[:, col].copy())
            P.psb_spmm(2.0, a, P.vect_from(X[:, col].copy()), -1.0, y1)
            assert np.array_equal(Y[:, col], y1.get_vect()), col
        a.free()
    finally:
        N.fn("psb_hip_hll_set_index16")(old)


def test_hll_variants_agree(gpu, orc):
    _say("case hll variants")
    from psblas3_amd import _native as N
    coo = orc.gen_powerlaw(5000, cap=200, seed=4)
    a = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64))
    x = P.vect_from(orc.uniform_pm1(3, coo.nc))
    outs = []
    for var in (1, 2, 5, 6):
        old = N.fn("psb_hip_hll_set_variant")(var)
        y = P.vect_from(np.zeros(coo.nr))
        P.psb_spmm(1.0, a, x, 0.0, y)
        outs.append(y.get_vect())
        N.fn("psb_hip_hll_set_variant")(old)
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)   # same per-row order -> same bits
    a.free()


def test_spmv_errors(gpu, orc):
    _say("case error paths")
    coo = orc.gen_convdiff(5)
    a = _mat(coo, P.psb_d_hip_hlg_sparse_mat(32))
    xs = P.vect_from(np.zeros(10))               # too short
    y = P.vect_from(np.zeros(coo.nr))
    with pytest.raises(P.PsbError):
        P.psb_spmm(1.0, a, xs, 0.0, y)
    xf = P.vect_from(np.zeros(coo.nc, np.float32), mold=P.psb_s_hip_vect())
    with pytest.raises(P.PsbError):
        P.psb_spmm(1.0, a, xf, 0.0, y)            # precision mismatch
    a.free()


def test_hll_autotune(gpu, orc):
    _say("case hll autotune (hksz 32/64 x 4 variants)")
    coo = orc.gen_convdiff(30)
    pc = P.psb_d_coo_sparse_mat(coo.nr, coo.nc, coo.ia, coo.ja, coo.val)
    m, rep = P.psb_hip_hll_autotune(pc, reps=5)
    assert rep["hksz"] in (32, 64) and rep["variant"] in (1, 2, 5, 6)
    assert len(rep["table"]) == 8 and all(t > 0 for t in rep["table"].values())
    assert m.hksz == rep["hksz"]
    # the tuned matrix gives the same bits as the default kernel on the same hack size
    x = P.vect_from(orc.uniform_pm1(2, coo.nc))
    y1 = P.vect_from(np.zeros(coo.nr))
    a = P.psb_dspmat_type()
    a.m = a.n = coo.nr
    a.a = m
    P.psb_spmm(1.0, a, x, 0.0, y1)
    ref = P.psb_dspmat_type()
    ref.m = ref.n = coo.nr
    ref.a = P.psb_d_hip_hlg_sparse_mat(rep["hksz"]).cp_from_coo(pc)
    y2 = P.vect_from(np.zeros(coo.nr))
    P.psb_spmm(1.0, ref, x, 0.0, y2)
    assert np.array_equal(y1.get_vect(), y2.get_vect())
    m.free()
    ref.a.free()


@pytest.mark.parametrize("kind", ["powerlaw", "stencil", "mixed_long"])
def test_csr_variants(gpu, orc, kind):
    _say(f"case csr uniform vs length-binned kernels, {kind}")
    from psblas3_amd import _native as N
    if kind == "powerlaw":
        coo = orc.gen_powerlaw(20000, cap=4095, seed=5)
    elif kind == "stencil":
        coo = orc.gen_convdiff(21)
    else:   # a few rows longer than the binned wave-per-row threshold, empty rows
        n = 3000
        rng = np.random.default_rng(3)
        rows, cols = [], []
        for r in range(n):
            k = 0 if r % 97 == 0 else (900 if r % 500 == 1 else 1 + (r % 40))
            rows += [r + 1] * k
            cols += list(rng.choice(n, size=k, replace=False) + 1)
        coo = orc.COO(n, n, np.array(rows, np.int32), np.array(cols, np.int32),
                      rng.uniform(-1, 1, len(rows)))
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    a = _mat(coo, P.psb_d_hip_csrg_sparse_mat())
    x = orc.uniform_pm1(8, coo.nc)
    y0 = orc.uniform_pm1(9, coo.nr)
    ref = orc.csr_spmv(irp, ja, val, x, 1.5, -0.5, y0)
    bound = _row_bound(f, x, 1.5, -0.5, y0, np.finfo(float).eps)
    for var in (1, 2, 3, 0):
        old = N.fn("psb_hip_csr_set_variant")(var)
        xv, yv = P.vect_from(x), P.vect_from(y0.copy())
        P.psb_spmm(1.5, a, xv, -0.5, yv)
        N.fn("psb_hip_csr_set_variant")(old)
        assert np.all(np.abs(yv.get_vect() - ref) <= bound), var
    # the row-block kernel sums separately rounded products in CSR order, as
    # the CPU loop does: with beta = 0 its rows of at most 512 entries match
    # the oracle bit for bit (longer rows go to the wave-per-row kernel)
    old = N.fn("psb_hip_csr_set_variant")(3)
    yv = P.vect_from(np.zeros(coo.nr))
    P.psb_spmm(1.5, a, P.vect_from(x), 0.0, yv)
    N.fn("psb_hip_csr_set_variant")(old)
    ref0 = orc.csr_spmv(irp, ja, val, x, 1.5, 0.0)
    short = np.diff(irp) <= 512
    got = yv.get_vect()
    assert np.array_equal(got[short].view(np.int64), ref0[short].view(np.int64))
    assert np.all(np.abs(got - ref0) <= _row_bound(f, x, 1.5, 0.0, np.zeros(coo.nr),
                                                    np.finfo(float).eps))
    a.free()


@pytest.mark.parametrize("slices", [2, 5, 16])
@pytest.mark.parametrize("single", [False, True], ids=["d", "s"])
def test_csr_column_slices(gpu, orc, slices, single):
    """The product over column slices (psb_hip_csr_set_slices) adds each row's
    slice sums: within the per-row bound for alpha / beta, also after the
    matrix (and so its slices) is scaled."""
    _say(f"case csr column slices {slices} single={single}")
    from psblas3_amd import _native as N
    coo = orc.gen_powerlaw(20000, cap=1500, seed=7)
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    dt = np.float32 if single else np.float64
    eps = np.finfo(dt).eps
    mold = P.psb_s_hip_csrg_sparse_mat() if single else P.psb_d_hip_csrg_sparse_mat()
    vm = P.psb_s_hip_vect() if single else None
    old = N.fn("psb_hip_csr_set_slices")(slices)
    try:
        a = _mat(coo, mold)
    finally:
        N.fn("psb_hip_csr_set_slices")(old)
    assert a.a.slices() == slices
    x = orc.uniform_pm1(10, coo.nc)
    y0 = orc.uniform_pm1(11, coo.nr)
    for alpha, beta in ((1.5, -0.5), (1.0, 0.0)):
        ref = orc.csr_spmv(irp, ja, val, x, alpha, beta, y0)
        bound = _row_bound(f, x, alpha, beta, y0, eps) * (2 if single else 1)
        yv = P.vect_from(y0.astype(dt), mold=vm)
        P.psb_spmm(alpha, a, P.vect_from(x.astype(dt), mold=vm), beta, yv)
        assert np.all(np.abs(yv.get_vect() - ref) <= bound), (alpha, beta)
    # scaling reaches the slices: 2 * diag(d) A
    d = orc.uniform_pm1(12, coo.nr) + 2.0
    a.a.scals(2.0)
    a.a.scal(d.astype(dt), side="L")
    scaled = 2.0 * val * d[np.repeat(np.arange(coo.nr), np.diff(irp))]
    ref = orc.csr_spmv(irp, ja, scaled, x)
    fs = orc.COO(coo.nr, coo.nc, f.ia, f.ja, np.abs(scaled))
    bound = _row_bound(fs, x, 1.0, 0.0, np.zeros(coo.nr), eps) * (8 if single else 2)
    yv = P.vect_from(np.zeros(coo.nr, dt), mold=vm)
    P.psb_spmm(1.0, a, P.vect_from(x.astype(dt), mold=vm), 0.0, yv)
    assert np.all(np.abs(yv.get_vect() - ref) <= bound)
    # variant 3 runs the row-block kernel on the whole matrix
    oldv = N.fn("psb_hip_csr_set_variant")(3)
    y3 = P.vect_from(np.zeros(coo.nr, dt), mold=vm)
    P.psb_spmm(1.0, a, P.vect_from(x.astype(dt), mold=vm), 0.0, y3)
    N.fn("psb_hip_csr_set_variant")(oldv)
    assert np.all(np.abs(y3.get_vect() - ref) <= bound)
    a.free()


def test_csr_slices_chosen_automatically(gpu, orc):
    """Column slices are built for a scattered matrix whose x exceeds the
    threshold (13.6 MB of x: 3 slices), not for a banded one of the same
    size, nor below the threshold."""
    _say("case csr automatic column slices")
    from psblas3_amd import _native as N
    assert N.fn("psb_hip_csr_set_slices")(-1) == -1
    n = 1_700_000
    coo = orc.gen_powerlaw(n, cap=1000, seed=9)
    a = _mat(coo, P.psb_d_hip_csrg_sparse_mat())
    assert a.a.slices() == 3
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    x = orc.uniform_pm1(13, n)
    yv = P.vect_from(np.zeros(n))
    P.psb_spmm(1.0, a, P.vect_from(x), 0.0, yv)
    ref = orc.csr_spmv(irp, ja, val, x)
    assert np.all(np.abs(yv.get_vect() - ref) <= _row_bound(f, x, 1.0, 0.0, np.zeros(n),
                                                             np.finfo(float).eps))
    a.free()
    s = _mat(orc.gen_convdiff(120), P.psb_d_hip_csrg_sparse_mat())      # 1.728M rows, banded
    assert s.a.slices() == 0
    s.free()
    small = _mat(orc.gen_powerlaw(200000, cap=500, seed=9), P.psb_d_hip_csrg_sparse_mat())
    assert small.a.slices() == 0
    small.free()


@pytest.mark.parametrize("fmt", ["hll64", "ell", "csr"])
def test_xcd_block_order_is_a_permutation(gpu, orc, fmt):
    """Every row is computed by one lane whatever the block order, so the
    chunked XCD orders (chunks of 1, 4, 16, 64 blocks, or one range per XCD)
    give the same bits as each other and as launch order; the multi-RHS
    product too.  Matrix large enough for many chunks per XCD."""
    from psblas3_amd import _native as N
    _say(f"case XCD block orders, {fmt}")
    coo = orc.gen_convdiff(48)                      # 110,592 rows: 216 / 432 blocks
    a = _mat(coo, MOLDS[fmt]())
    x = orc.uniform_pm1(4, coo.nc)
    xk = np.asfortranarray(np.stack([orc.uniform_pm1(5 + q, coo.nc) for q in range(3)], axis=1))
    outs = []
    old = N.fn("psb_hip_hll_set_xcd_chunk")(16)
    try:
        for chunk in (16, 0, 1, 4, 64, -1):
            N.fn("psb_hip_hll_set_xcd_chunk")(chunk)
            yv = P.vect_from(np.zeros(coo.nr))
            P.psb_spmm(1.0, a, P.vect_from(x), 0.0, yv)
            y = yv.get_vect()
            if fmt == "hll64":
                yk = P.vect_from(np.zeros((coo.nr, 3)))
                P.psb_spmm(1.0, a, P.vect_from(xk), 0.0, yk)
                y = np.concatenate([y, np.asarray(yk.get_vect()).ravel(order="F")])
            outs.append(y)
        if fmt == "hll64":   # launch order (variant 2) for the single product
            oldv = N.fn("psb_hip_hll_set_variant")(2)
            yv = P.vect_from(np.zeros(coo.nr))
            P.psb_spmm(1.0, a, P.vect_from(x), 0.0, yv)
            N.fn("psb_hip_hll_set_variant")(oldv)
            assert np.array_equal(yv.get_vect().view(np.int64), outs[0][:coo.nr].view(np.int64))
    finally:
        N.fn("psb_hip_hll_set_xcd_chunk")(old)
    for y in outs[1:]:
        assert np.array_equal(y.view(np.int64), outs[0].view(np.int64))
    irp, ja, val = orc.csr_from_coo(orc.fix_coo(coo))
    ref = orc.csr_spmv(irp, ja, val, x)
    assert np.allclose(outs[0][:coo.nr], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    a.free()


def _torture_t_cases():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "torture_usmv_t.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("fmt", ["hll32", "ell", "csr"])
@pytest.mark.parametrize("single", [False, True], ids=["d", "s"])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0), (2.5, -0.5), (0.0, 2.0),
                                        (1.0, -1.0)])
@pytest.mark.parametrize("ncols", [5200, 5201])
def test_transposed_product_bitexact(gpu, orc, fmt, single, alpha, beta, ncols):
    """psb_spmm with trans 'T' / 'C': the transpose built on the device from
    the stored entries, each y(c) updated in the CPU loop's order, equal to
    psb_d_csr_csmv's transposed branch bit for bit (rectangular matrix, empty
    rows and columns; an odd column count leaves the FP64 two-rows-per-lane
    kernel's last lane one row)."""
    _say(f"case transposed product {fmt} single={single} alpha={alpha} beta={beta} nc={ncols}")
    base = orc.gen_powerlaw(5000, cap=400, seed=12)
    keep = (base.ja % 53) != 0                       # some empty columns
    coo = orc.COO(5000, ncols, base.ia[keep], base.ja[keep], base.val[keep])
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    dt = np.float32 if single else np.float64
    mold = {"hll32": lambda: P.psb_s_hip_hlg_sparse_mat(32) if single else P.psb_d_hip_hlg_sparse_mat(32),
            "ell": lambda: P.psb_s_hip_elg_sparse_mat() if single else P.psb_d_hip_elg_sparse_mat(),
            "csr": lambda: P.psb_s_hip_csrg_sparse_mat() if single else P.psb_d_hip_csrg_sparse_mat()}[fmt]
    a = _mat(coo, mold())
    vm = P.psb_s_hip_vect() if single else None
    x = orc.uniform_pm1(13, coo.nr).astype(dt)
    y0 = orc.uniform_pm1(14, coo.nc).astype(dt)
    for op in ("T", "C"):
        yv = P.vect_from(y0.copy(), mold=vm)
        P.psb_spmm(alpha, a, P.vect_from(x, mold=vm), beta, yv, trans=op)
        ref = orc.csr_spmv_t(irp, ja, val.astype(dt), x, alpha, beta, y0, nc=coo.nc)
        got = yv.get_vect()
        assert got.dtype == dt
        assert np.array_equal(got.view(np.int32 if single else np.int64),
                              ref.view(np.int32 if single else np.int64)), op
    # the transpose follows a rewrite of the storage
    a.a.scals(2.0)
    yv = P.vect_from(y0.copy(), mold=vm)
    P.psb_spmm(alpha, a, P.vect_from(x, mold=vm), beta, yv, trans="T")
    ref = orc.csr_spmv_t(irp, ja, (2.0 * val).astype(dt) if not single else
                         (val.astype(dt) * dt(2.0)), x, alpha, beta, y0, nc=coo.nc)
    assert np.array_equal(yv.get_vect(), ref)
    a.free()


@pytest.mark.parametrize("case", _torture_t_cases(), ids=lambda c: c["name"])
def test_torture_usmv_transposed_device(gpu, orc, case):
    single = case["kind"] == "s"
    dt = np.float32 if single else np.float64
    vm = P.psb_s_hip_vect() if single else None
    coo = orc.COO(case["m"], case["k"], np.array(case["ia"], np.int32),
                  np.array(case["ja"], np.int32), np.array(case["va"]))
    a = _mat(coo, P.psb_s_hip_csrg_sparse_mat() if single else P.psb_d_hip_csrg_sparse_mat())
    y = P.vect_from(np.array(case["bcy"], dt), mold=vm)
    P.psb_spmm(case["alpha"], a, P.vect_from(np.array(case["x"], dt), mold=vm), case["beta"], y,
               trans=case["trans"])
    assert y.get_vect().tolist() == case["cy"]
    a.free()
