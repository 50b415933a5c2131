"""Row-sorted HLL (psb_hip_hll_sort_rows + psb_hip_hll_copy_coo_sorted): rows
sorted by length inside windows, hacks built over the sorted positions, each
row summed over its own entries only, wide hacks cut into work units.

Checked against the oracle and against the rows-in-order HLL:
- the stored entries, read back and un-permuted, equal the oracle's CSR bit
  for bit; perm is a bijection; per-position counts and hack offsets follow
  from the sorted counts; padding past nr is ja = nr, val = 0;
- every row not in a cut hack has the value of the rows-in-order product (the
  same slots in the same order); every row meets the per-row bound against
  the oracle's CSR product, for alpha / beta, FP64 and FP32, one and several
  columns;
- get_diag, scal (left / right) and a drop-in Krylov solve agree with the
  rows-in-order matrix.
"""
import numpy as np
import pytest

import psblas3_amd as P

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _mat(coo, mold):
    return P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val, mold=mold)


def _row_bound(f, x, alpha, beta, y0, eps):
    nr = f.nr
    absax = np.zeros(nr)
    np.add.at(absax, f.ia - 1, np.abs(f.val) * np.abs(x[f.ja - 1]))
    nnz_row = np.bincount(f.ia - 1, minlength=nr)
    return 4.0 * (nnz_row + 1) * eps * (abs(alpha) * absax + abs(beta) * np.abs(y0)) + 1e-300


CASES = {
    # rows up to 1000 entries: hacks of the sorted matrix up to 1000 slots (cut)
    "powerlaw_20000": lambda orc: orc.gen_powerlaw(20000, cap=1000, seed=5),
    "powerlaw_3000": lambda orc: orc.gen_powerlaw(3000, cap=300, seed=1),
    "convdiff_L13": lambda orc: orc.gen_convdiff(13),
}


def _cut_rows(a):
    """Rows in hacks wider than 256 slots (computed in several units)."""
    hk = a.a.hkoffs
    w = np.diff(hk) // a.a.hksz
    pos = np.arange(a.a.nr)
    cut = w[pos // a.a.hksz] > 256
    return np.sort(a.a.perm[cut] - 1)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("window,hksz", [(0, 64), (4096, 64), (64, 64), (0, 32)])
def test_sorted_conversion(gpu, orc, name, window, hksz):
    _say("case sorted conversion", name, "window", window, "hksz", hksz)
    coo = CASES[name](orc)
    f = orc.fix_coo(coo)
    a = _mat(coo, P.psb_d_hip_hlg_sparse_mat(hksz, sort_window=window))
    perm = a.a.perm
    assert np.array_equal(np.sort(perm), np.arange(1, f.nr + 1))
    from psblas3_amd import _native as N
    pd = np.zeros(f.nr, np.int32)
    N.call("psb_hip_hll_read_perm", a.a.dev, pd.ctypes.data)
    assert np.array_equal(pd, perm)
    cnt = np.bincount(f.ia - 1, minlength=f.nr)
    hk, irn, ja, val, idiag = a.a.to_host()
    assert np.array_equal(irn, cnt[perm - 1])
    # sorted inside each window
    win = f.nr if window <= 0 else -(-window // hksz) * hksz
    for w0 in range(0, f.nr, win):
        seg = irn[w0:w0 + win]
        assert np.all(seg[:-1] >= seg[1:])
    widths = np.array([irn[h * hksz:(h + 1) * hksz].max() for h in range(len(hk) - 1)])
    assert np.array_equal(np.diff(hk), hksz * widths)
    # entries, un-permuted, are the oracle's CSR
    irp, cja, cval = a.a.to_csr()
    oirp, oja, oval = orc.csr_from_coo(f)
    assert np.array_equal(irp, oirp) and np.array_equal(cja, oja)
    assert np.array_equal(cval.view(np.int64), oval.view(np.int64))
    # diagonal slot per position
    for p in range(0, f.nr, max(1, f.nr // 200)):
        r = perm[p] - 1
        cols = oja[oirp[r] - 1:oirp[r + 1] - 1]
        want = int(np.nonzero(cols == r + 1)[0][0]) + 1 if (cols == r + 1).any() else 0
        assert idiag[p] == want
    # padding of the positions past nr: ja = nr, val = 0
    last = len(hk) - 2
    wl = widths[last]
    for p in range(f.nr, (last + 1) * hksz):
        lane = p - last * hksz
        sl = hk[last] + np.arange(wl) * hksz + lane
        assert np.all(ja[sl] == f.nr) and np.all(val[sl] == 0)
    if name == "powerlaw_20000":
        # a whole-matrix sort removes nearly all of the padding of the
        # rows-in-order layout; a window sort keeps each window's first hack
        # as wide as its longest row (5 windows of 4096 rows here: 1.85x),
        # which the width check above already pins
        if window <= 0:
            assert a.a.isz <= 1.2 * f.nz
        h = orc.hll_from_coo(f, hksz)
        if window == hksz:
            # sorting inside one hack leaves every hack's width unchanged
            assert a.a.isz == h.hkoffs[-1]
        else:
            assert a.a.isz < h.hkoffs[-1]
    a.free()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("window", [0, 4096])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-2.0, 0.5)])
def test_sorted_spmv_fp64(gpu, orc, name, window, alpha, beta):
    _say("case sorted spmv fp64", name, window, alpha, beta)
    coo = CASES[name](orc)
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    x = orc.uniform_pm1(1, f.nc)
    y0 = orc.uniform_pm1(2, f.nr)
    srt = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64, sort_window=window))
    ref = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64))
    nu, ns = srt.a.units()
    _say(f"  {nu} units, {ns} cut hacks")
    if name == "powerlaw_20000":
        assert ns > 0
    ys, yr = P.vect_from(y0), P.vect_from(y0)
    xv = P.vect_from(x)
    P.psb_spmm(alpha, srt, xv, beta, ys)
    P.psb_spmm(alpha, ref, xv, beta, yr)
    ys, yr = ys.get_vect(), yr.get_vect()
    yo = orc.csr_spmv(irp, ja, val, x, alpha, beta, y0.copy())
    assert np.all(np.abs(ys - yo) <= _row_bound(f, x, alpha, beta, y0, np.finfo(float).eps))
    keep = np.ones(f.nr, bool)
    keep[_cut_rows(srt)] = False
    assert np.array_equal(ys[keep], yr[keep])      # same slots, same order
    srt.free()
    ref.free()


def test_sorted_spmv_fp32_and_single_copy(gpu, orc):
    _say("case sorted spmv fp32")
    coo = CASES["powerlaw_20000"](orc)
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    x = orc.uniform_pm1(3, f.nc)
    srt = _mat(coo, P.psb_s_hip_hlg_sparse_mat(64, sort_window=0))
    ref = _mat(coo, P.psb_s_hip_hlg_sparse_mat(64))
    x32 = x.astype(np.float32)
    ys = P.vect_from(np.zeros(f.nr, np.float32), mold=P.psb_s_hip_vect())
    yr = P.vect_from(np.zeros(f.nr, np.float32), mold=P.psb_s_hip_vect())
    xv = P.vect_from(x32, mold=P.psb_s_hip_vect())
    P.psb_spmm(1.0, srt, xv, 0.0, ys)
    P.psb_spmm(1.0, ref, xv, 0.0, yr)
    ys, yr = ys.get_vect(), yr.get_vect()
    yo = orc.csr_spmv(irp, ja, val.astype(np.float32).astype(np.float64), x32.astype(np.float64))
    bound = 2 * _row_bound(f, x, 1.0, 0.0, np.zeros(f.nr), np.finfo(np.float32).eps)
    assert np.all(np.abs(ys - yo) <= bound)
    keep = np.ones(f.nr, bool)
    keep[_cut_rows(srt)] = False
    assert np.array_equal(ys[keep], yr[keep])
    # the FP32 copy of a double sorted matrix (psb_hip_hll_spmv_single)
    from psblas3_amd import _native as N
    d = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64, sort_window=0))
    N.call("psb_hip_hll_make_single", d.a.dev)
    y1 = P.vect_from(np.zeros(f.nr, np.float32), mold=P.psb_s_hip_vect())
    N.call("psb_hip_hll_spmv_single", d.a.dev, 1.0, xv.v.dev, 0.0, y1.v.dev)
    y1.v.set_dev()
    assert np.array_equal(y1.get_vect(), ys)   # same slots, same FP32 values
    # the fused solvers take rows in order only
    prec = P.psb_dprec_type("DIAG").build(d, P.psb_cdall(nl=f.nr))
    with pytest.raises(P.PsbError):
        P.psb_krylov_fused("CG", d, prec, P.vect_from(np.ones(f.nr)),
                           P.vect_from(np.zeros(f.nr)), 1e-6)
    for m in (srt, ref, d):
        m.free()


@pytest.mark.parametrize("k", [3, 8])
def test_sorted_spmm(gpu, orc, k):
    _say(f"case sorted spmm k={k}")
    coo = CASES["powerlaw_3000"](orc)
    f = orc.fix_coo(coo)
    h = orc.hll_from_coo(f, 64)
    a = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64, sort_window=0))
    X = np.stack([orc.uniform_pm1(s, coo.nc) for s in range(k)], axis=1)
    Y0 = np.stack([orc.uniform_pm1(20 + s, coo.nr) for s in range(k)], axis=1)
    yv = P.vect_from(Y0)
    P.psb_spmm(2.0, a, P.vect_from(X), -1.0, yv)
    Y = yv.get_vect()
    Yref = orc.hll_spmm(h, X, 2.0, -1.0, Y0)
    for col in range(k):
        bound = _row_bound(f, X[:, col], 2.0, -1.0, Y0[:, col], np.finfo(float).eps)
        assert np.all(np.abs(Y[:, col] - Yref[:, col]) <= bound)
    a.free()


def test_sorted_diag_scal_and_solve(gpu, orc):
    _say("case sorted get_diag / scal / drop-in BiCGSTAB")
    coo = orc.gen_convdiff(12, conv=3.0)
    # make the rows uneven: drop some off-diagonal entries
    keep = (coo.ia == coo.ja) | ((coo.ia * 7 + coo.ja) % 5 != 0)
    coo = orc.COO(coo.nr, coo.nc, coo.ia[keep], coo.ja[keep], coo.val[keep])
    f = orc.fix_coo(coo)
    n = f.nr
    srt = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64, sort_window=256))
    ref = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64))
    assert not np.array_equal(srt.a.perm, np.arange(1, n + 1))
    d1, d2 = P.psb_d_hip_vect().all(n), P.psb_d_hip_vect().all(n)
    srt.a.get_diag(d1)
    ref.a.get_diag(d2)
    assert np.array_equal(d1.get_vect(), d2.get_vect())
    dl = orc.uniform_pm1(9, n) + 2.0
    for side in ("L", "R"):
        srt.a.scal(dl, side)
        ref.a.scal(dl, side)
        a1, a2 = srt.a.to_csr(), ref.a.to_csr()
        assert all(np.array_equal(u, v) for u, v in zip(a1, a2))
    bh = orc.csr_spmv(*ref.a.to_csr(), np.ones(n))
    its = []
    for m in (srt, ref):
        desc = P.psb_cdall(nl=n)
        prec = P.psb_dprec_type("DIAG").build(m, desc)
        x = P.vect_from(np.zeros(n))
        it, err, _ = P.psb_krylov("BICGSTAB", m, prec, P.vect_from(bh), x, 1e-9, desc)
        assert err <= 1e-9 and np.allclose(x.get_vect(), 1.0, atol=1e-6)
        its.append(it)
    # no hack is cut (rows of at most 7 entries): the same products, the same run
    assert its[0] == its[1]
    srt.free()
    ref.free()


@pytest.mark.parametrize("slices", [2, 3, 5])
@pytest.mark.parametrize("single", [False, True])
def test_sorted_column_slices(gpu, orc, slices, single):
    """Row-sorted HLL with column slices (forced; automatic only for x over
    12 MiB with scattered columns): each slice is a row-sorted matrix of its
    own over a column range, the product adds the slices' sums; every row
    meets the per-row bound against the oracle, for alpha / beta and
    several columns."""
    from psblas3_amd import _native as N
    _say("case sorted slices", slices, "fp32" if single else "fp64")
    coo = orc.gen_powerlaw(20000, cap=1000, seed=5)
    f = orc.fix_coo(coo)
    old = N.fn("psb_hip_hll_set_slices")(slices)
    try:
        mold = P.psb_s_hip_hlg_sparse_mat(64, sort_window=0) if single \
            else P.psb_d_hip_hlg_sparse_mat(64, sort_window=0)
        a = _mat(coo, mold)
    finally:
        N.fn("psb_hip_hll_set_slices")(old)
    assert N.fn("psb_hip_hll_slices")(a.a.dev) == slices
    dt = np.float32 if single else np.float64
    eps = np.finfo(dt).eps * (2.0 if single else 1.0)
    vm = P.psb_s_hip_vect() if single else P.psb_d_hip_vect()
    x = orc.uniform_pm1(8, f.nc)
    y0 = orc.uniform_pm1(9, f.nr)
    irp, ja, val = orc.csr_from_coo(f)
    for alpha, beta in ((1.0, 0.0), (-1.5, 0.25)):
        xv = P.vect_from(x.astype(dt), mold=vm)
        yv = P.vect_from(y0.astype(dt), mold=vm)
        P.psb_spmm(alpha, a, xv, beta, yv)
        yref = orc.csr_spmv(irp, ja, val, x, alpha, beta, y0.copy())
        bound = _row_bound(f, x, alpha, beta, y0, eps)
        assert np.all(np.abs(yv.get_vect().astype(np.float64) - yref) <= bound)
    if not single:
        # three columns at once: each column as the single product
        X = np.stack([orc.uniform_pm1(30 + c, f.nc) for c in range(3)], axis=1)
        xk = P.psb_d_hip_vect().bld(X)
        yk = P.psb_d_hip_vect().all(f.nr, 3)
        P.psb_spmm(1.0, a, xk, 0.0, yk)
        Y = yk.get_vect()
        for c in range(3):
            ys = P.vect_from(np.zeros(f.nr))
            P.psb_spmm(1.0, a, P.vect_from(np.ascontiguousarray(X[:, c])), 0.0, ys)
            assert np.array_equal(Y[:, c], ys.get_vect())
    # the stored (unsliced) arrays are unchanged: read back, they are the oracle's CSR
    irp2, ja2, val2 = a.a.to_csr()
    assert np.array_equal(irp2, irp) and np.array_equal(ja2, ja)
    a.free()
