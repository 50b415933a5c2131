"""16-bit column offsets of rows-in-order HLL matrices (psb_hip_hll_index16):
the products read a 16-bit residual per slot plus a base per hack slot
column instead of ja.  The same slots in the same order, so every product
equals the ja path bit for bit; matrices whose slot columns span more than
65535 columns keep ja only.  The products are also checked against the
oracle within the per-row bound."""
import numpy as np
import pytest

import psblas3_amd as P
from psblas3_amd import _native as N

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _mat(coo, mold):
    return P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val, mold=mold)


CASES = {
    "convdiff_L40": (lambda orc: orc.gen_convdiff(40), True),
    "convdiff_L23": (lambda orc: orc.gen_convdiff(23), True),     # hacks straddle planes
    "poisson_L31": (lambda orc: orc.gen_poisson3d(31), True),
    "aniso_L77": (lambda orc: orc.gen_aniso2d(77, 0.01), True),
    # columns within 20000 of every row: offsets always fit
    "powerlaw_20000": (lambda orc: orc.gen_powerlaw(20000, cap=1000, seed=5), True),
    # random columns over 100000: slot columns span more than 65535
    "powerlaw_100000": (lambda orc: orc.gen_powerlaw(100000, cap=64, seed=6), False),
}


def _fits(orc, f, hksz):
    """The encoding rule on the oracle's HLL arrays: every hack slot column's
    offsets (column - row) span at most 65535."""
    h = orc.hll_from_coo(f, hksz)
    w = np.diff(h.hkoffs.astype(np.int64)) // hksz
    cols = h.ja.astype(np.int64).reshape(-1, hksz) - 1
    rows = np.repeat(np.arange(len(w)), w)[:, None] * hksz + np.arange(hksz)[None, :]
    d = cols - rows
    return bool(((d.max(1) - d.min(1)) <= 65535).all())


def _both(m, xv, vm, alpha, beta, y0):
    out = []
    for on in (0, 1):
        old = N.fn("psb_hip_hll_set_index16")(on)
        try:
            yv = P.vect_from(y0.copy(), mold=vm)
            P.psb_spmm(alpha, m, xv, beta, yv)
            out.append(yv.get_vect().copy())
        finally:
            N.fn("psb_hip_hll_set_index16")(old)
    return out


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("hksz", [64, 32, 7])
@pytest.mark.parametrize("single", [False, True])
def test_index16_products_bitexact(gpu, orc, name, hksz, single):
    _say("case index16", name, hksz, "fp32" if single else "fp64")
    gen, fits = CASES[name]
    coo = gen(orc)
    f = orc.fix_coo(coo)
    assert _fits(orc, f, hksz) == fits
    mold = P.psb_s_hip_hlg_sparse_mat(hksz) if single else P.psb_d_hip_hlg_sparse_mat(hksz)
    m = _mat(coo, mold)
    assert N.fn("psb_hip_hll_index16")(m.a.dev) == int(fits)
    dt = np.float32 if single else np.float64
    vm = P.psb_s_hip_vect() if single else P.psb_d_hip_vect()
    x = orc.uniform_pm1(3, coo.nc).astype(dt)
    y0 = orc.uniform_pm1(4, coo.nr).astype(dt)
    xv = P.vect_from(x, mold=vm)
    irp, ja, val = orc.csr_from_coo(f)
    for alpha, beta in ((1.0, 0.0), (-2.0, 0.5)):
        off, on = _both(m, xv, vm, alpha, beta, y0)
        assert np.array_equal(off.view(np.uint8), on.view(np.uint8))
        if not single:
            yref = orc.csr_spmv(irp, ja, val, x, alpha, beta, y0.copy())
            absax = orc.csr_spmv(irp, ja, np.abs(val), np.abs(x))
            nnz_row = np.diff(irp)
            bound = 4.0 * (nnz_row + 1) * np.finfo(float).eps * (abs(alpha) * absax +
                                                                  abs(beta) * np.abs(y0))
            assert np.all(np.abs(on - yref) <= bound + 1e-300)
    m.free()


@pytest.mark.parametrize("k", [3, 8])
def test_index16_spmm_bitexact(gpu, orc, k):
    _say("case index16 spmm", k)
    coo = orc.gen_convdiff(30)
    m = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64))
    assert N.fn("psb_hip_hll_index16")(m.a.dev) == 1
    n = coo.nr
    X = np.stack([orc.uniform_pm1(10 + c, n) for c in range(k)], axis=1)
    xv = P.psb_d_hip_vect().all(n, k)
    xv.bld(X)
    yv = P.psb_d_hip_vect().all(n, k)
    outs = []
    for on in (0, 1):
        old = N.fn("psb_hip_hll_set_index16")(on)
        try:
            P.psb_spmm(1.0, m, xv, 0.0, yv)
            outs.append(yv.get_vect().copy())
        finally:
            N.fn("psb_hip_hll_set_index16")(old)
    assert np.array_equal(outs[0].view(np.uint8), outs[1].view(np.uint8))
    # and each column equals the single-vector product
    for c in range(k):
        xs = P.vect_from(np.ascontiguousarray(X[:, c]))
        ys = P.vect_from(np.zeros(n))
        P.psb_spmm(1.0, m, xs, 0.0, ys)
        assert np.array_equal(ys.get_vect(), outs[1][:, c])
    m.free()


def test_index16_dropped_for_sorted_rows(gpu, orc):
    """A row-sorted matrix has no 16-bit offsets (its products run work
    units over ja)."""
    coo = orc.gen_convdiff(12)
    m = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64, sort_window=0))
    assert N.fn("psb_hip_hll_index16")(m.a.dev) == 0
    m.free()


@pytest.mark.parametrize("tile", [1, 2, 4])
def test_plane_order_bitexact(gpu, orc, tile):
    """psb_hip_hll_set_plane only reorders the blocks of the launch: the
    single and multi-RHS products equal the default order's bit for bit."""
    _say("case plane order", tile)
    L = 30
    coo = orc.gen_convdiff(L)
    m = _mat(coo, P.psb_d_hip_hlg_sparse_mat(64))
    n = coo.nr
    x = orc.uniform_pm1(5, n)
    X = np.stack([orc.uniform_pm1(20 + c, n) for c in range(5)], axis=1)
    outs = []
    for period, tb in ((0, 0), (L * L, tile)):
        N.call("psb_hip_hll_set_plane", m.a.dev, period, tb)
        yv = P.vect_from(np.zeros(n))
        P.psb_spmm(1.0, m, P.vect_from(x), 0.0, yv)
        xk = P.psb_d_hip_vect().bld(X)
        yk = P.psb_d_hip_vect().all(n, 5)
        P.psb_spmm(1.0, m, xk, 0.0, yk)
        outs.append((yv.get_vect(), yk.get_vect()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    m.free()


def _ell_fits(orc, f):
    """The ELL encoding rule on the oracle's ELL arrays (ldv = nr): every slot
    column of every group of 64 rows spans at most 65535 columns."""
    e = orc.ell_from_coo(f)
    nr, nzm = f.nr, e.nzm
    ja = e.ja.astype(np.int64).reshape(nzm, -1)[:, :nr]      # column-major, pitch = nr
    d = ja - 1 - np.arange(nr)[None, :]
    for g in range(0, nr, 64):
        blk = d[:, g:g + 64]
        if ((blk.max(1) - blk.min(1)) > 65535).any():
            return False
    return True


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("single", [False, True])
def test_index16_ell_bitexact(gpu, orc, name, single):
    _say("case index16 ell", name, "fp32" if single else "fp64")
    gen, fits = CASES[name]
    coo = gen(orc)
    f = orc.fix_coo(coo)
    assert _ell_fits(orc, f) == fits
    mold = P.psb_s_hip_elg_sparse_mat() if single else P.psb_d_hip_elg_sparse_mat()
    m = _mat(coo, mold)
    assert N.fn("psb_hip_ell_index16")(m.a.dev) == int(fits)
    dt = np.float32 if single else np.float64
    vm = P.psb_s_hip_vect() if single else P.psb_d_hip_vect()
    x = orc.uniform_pm1(3, coo.nc).astype(dt)
    y0 = orc.uniform_pm1(4, coo.nr).astype(dt)
    xv = P.vect_from(x, mold=vm)
    for alpha, beta in ((1.0, 0.0), (-2.0, 0.5)):
        off, on = _both(m, xv, vm, alpha, beta, y0)
        assert np.array_equal(off.view(np.uint8), on.view(np.uint8))
    m.free()


@pytest.mark.parametrize("fmt", ["HLG", "ELG"])
@pytest.mark.parametrize("method", ["CG", "BICGSTAB"])
def test_index16_fused_solvers_bitexact(gpu, orc, fmt, method):
    """The fused solvers' products read the same offsets: iteration counts,
    histories and solutions equal those over ja bit for bit."""
    _say("case index16 fused", fmt, method)
    coo = orc.gen_poisson3d(20) if method == "CG" else orc.gen_convdiff(20)
    mold = P.psb_d_hip_hlg_sparse_mat(64) if fmt == "HLG" else P.psb_d_hip_elg_sparse_mat()
    a = _mat(coo, mold)
    prec = P.psb_dprec_type("DIAG").build(a, P.psb_cdall(nl=coo.nr))
    b = orc.uniform_pm1(6, coo.nr)
    res = []
    for on in (0, 1):
        old = N.fn("psb_hip_hll_set_index16")(on)
        try:
            x = P.vect_from(np.zeros(coo.nr))
            it, err, hist = P.psb_krylov_fused(method, a, prec, P.vect_from(b), x, 1e-9,
                                               history=True)
            res.append((it, err, hist, x.get_vect()))
        finally:
            N.fn("psb_hip_hll_set_index16")(old)
    assert res[0][0] == res[1][0] and res[0][1] == res[1][1] and res[0][2] == res[1][2]
    assert np.array_equal(res[0][3], res[1][3])
    a.free()


@pytest.mark.parametrize("single", [False, True])
def test_band_chunk_order_bitexact(gpu, orc, single):
    """The automatic XCD chunk (about an eighth of the bandwidth: L = 128,
    band 16384, chunks of 4 pair blocks / 8 single blocks) only reorders the
    blocks: products equal those in 16-block chunks bit for bit."""
    _say("case band chunk", "fp32" if single else "fp64")
    coo = orc.gen_convdiff(128)
    mold = P.psb_s_hip_hlg_sparse_mat(64) if single else P.psb_d_hip_hlg_sparse_mat(64)
    m = _mat(coo, mold)
    dt = np.float32 if single else np.float64
    vm = P.psb_s_hip_vect() if single else P.psb_d_hip_vect()
    x = orc.uniform_pm1(7, coo.nc).astype(dt)
    outs = []
    for chunk in (16, -1):
        old = N.fn("psb_hip_hll_set_xcd_chunk")(chunk)
        try:
            yv = P.vect_from(np.zeros(coo.nr, dt), mold=vm)
            P.psb_spmm(1.0, m, P.vect_from(x, mold=vm), 0.0, yv)
            res = [yv.get_vect()]
            if not single:
                X = np.stack([orc.uniform_pm1(40 + c, coo.nc) for c in range(4)], axis=1)
                yk = P.psb_d_hip_vect().all(coo.nr, 4)
                P.psb_spmm(1.0, m, P.psb_d_hip_vect().bld(X), 0.0, yk)
                res.append(yk.get_vect())
            outs.append(res)
        finally:
            N.fn("psb_hip_hll_set_xcd_chunk")(old)
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    m.free()


@pytest.mark.parametrize("name", ["convdiff_L23", "aniso_L77", "powerlaw_20000", "powerlaw_100000"])
@pytest.mark.parametrize("single", [False, True])
def test_index16_transposed_bitexact(gpu, orc, name, single):
    """The transposed product (psb_spmm trans 'T') reads 16-bit source offsets
    when the transpose's slot columns allow them: the same bits as over the
    32-bit source rows, and the oracle's column-oriented loop bit for bit."""
    _say("case index16 transposed", name, "fp32" if single else "fp64")
    gen, _ = CASES[name]
    coo = gen(orc)
    f = orc.fix_coo(coo)
    mold = P.psb_s_hip_hlg_sparse_mat(64) if single else P.psb_d_hip_hlg_sparse_mat(64)
    m = _mat(coo, mold)
    dt = np.float32 if single else np.float64
    vm = P.psb_s_hip_vect() if single else P.psb_d_hip_vect()
    x = orc.uniform_pm1(11, coo.nr).astype(dt)
    y0 = orc.uniform_pm1(12, coo.nc).astype(dt)
    outs = []
    for on in (0, 1):
        old = N.fn("psb_hip_hll_set_index16")(on)
        try:
            yv = P.vect_from(y0.copy(), mold=vm)
            P.psb_spmm(-1.5, m, P.vect_from(x, mold=vm), 0.5, yv, trans="T")
            outs.append(yv.get_vect())
        finally:
            N.fn("psb_hip_hll_set_index16")(old)
        m.a._drop_t()      # rebuild the transpose under the next setting
    assert np.array_equal(outs[0].view(np.uint8), outs[1].view(np.uint8))
    if not single:
        irp, ja, val = orc.csr_from_coo(f)
        yref = orc.csr_spmv_t(irp, ja, val, x, -1.5, 0.5, y0.copy(), nc=f.nc)
        assert np.array_equal(outs[1].view(np.uint8), yref.view(np.uint8))
    m.free()
