"""GPU parity of the vector path: elementwise kernels bit-exact against the
oracle's restatement of the CPU routines; reductions within their error
bounds; gather/scatter/insert exact."""
import ctypes as C
import math

import numpy as np
import pytest

import psblas3_amd as P
from psblas3_amd import _native as N

pytestmark = pytest.mark.gpu


def _v(a, single=False):
    return P.vect_from(a, mold=P.psb_s_hip_vect() if single else P.psb_d_hip_vect())


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 1 << 20])
@pytest.mark.parametrize("ab", [(0.0, 0.0), (1.0, 0.0), (-1.0, 1.0), (0.0, 3.0), (2.5, -0.75),
                                (0.0, 1.0)])
def test_axpby_bitexact(gpu, orc, n, ab):
    print("case axpby", n, ab, flush=True)
    x = orc.uniform_pm1(1, n)
    y = orc.uniform_pm1(2, n)
    xv, yv = _v(x), _v(y)
    P.psb_geaxpby(ab[0], xv, ab[1], yv)
    assert _bits_equal(yv.get_vect(), orc.axpby(ab[0], x, ab[1], y))


def test_vecoperation_known_answers(gpu):
    print("case vecoperation", flush=True)
    n = 5000
    x, y = _v(np.ones(n)), _v(np.full(n, -2.0))
    P.psb_geaxpby(1.0, x, 1.0, y)
    assert np.all(y.get_vect() == -1.0)
    x, y = _v(np.ones(n)), _v(np.full(n, 2.0))
    P.psb_geaxpby(-1.0, x, 1.0, y)
    assert np.all(y.get_vect() == 1.0)
    x, y = _v(np.full(n, 2.0)), _v(np.full(n, -2.0))
    P.psb_geaxpby(0.5, x, 1.0, y)
    assert np.all(y.get_vect() == -1.0)
    x, y, z = _v(np.full(n, 2.0)), _v(np.full(n, -0.5)), _v(np.zeros(n))
    P.psb_gemlt(1.0, x, y, 0.0, z)
    assert np.all(z.get_vect() == -1.0)
    x, y = _v(np.full(n, 2.0)), _v(np.full(n, 0.5))
    assert P.psb_gedot(x, y) == n
    assert P.psb_geamax(_v(np.full(n, -0.5))) == 0.5
    assert P.psb_geasum(_v(np.full(n, -0.5))) == n * 0.5
    z = _v(np.zeros(n))
    P.psb_geabs(_v(np.full(n, -1.0)), z)
    assert np.all(z.get_vect() == 1.0)


def test_upd_xyz_xyzw_mlt_bitexact(gpu, orc):
    print("case upd_xyz/xyzw/mlt", flush=True)
    n = 100003
    x, y, z, w = (orc.uniform_pm1(s, n) for s in range(4))
    for coefs in ((2.0, 3.0, 0.5, -1.0), (0.0, 0.0, 0.5, 2.0), (1.0, 0.0, 0.0, 1.0)):
        xv, yv, zv = _v(x), _v(y), _v(z)
        P.psb_upd_xyz(*coefs, xv, yv, zv)
        ey, ez = orc.upd_xyz(*coefs, x, y, z)
        assert _bits_equal(yv.get_vect(), ey) and _bits_equal(zv.get_vect(), ez)
    xv, yv, zv, wv = _v(x), _v(y), _v(z), _v(w)
    P.psb_xyzw(1.0, 2.0, 3.0, 4.0, 5.0, 6.0, xv, yv, zv, wv)
    ey, ez, ew = orc.xyzw(1.0, 2.0, 3.0, 4.0, 5.0, 6.0, x, y, z, w)
    assert _bits_equal(yv.get_vect(), ey) and _bits_equal(zv.get_vect(), ez)
    assert _bits_equal(wv.get_vect(), ew)
    for alpha, beta in ((1.0, 0.0), (1.0, 1.0), (-1.0, 0.5), (0.3, 0.0), (0.3, 2.0), (0.0, 2.0)):
        xv, yv, zv = _v(x), _v(y), _v(z)
        P.psb_gemlt(alpha, xv, yv, beta, zv)
        assert _bits_equal(zv.get_vect(), orc.mlt(alpha, x, y, beta, z)), (alpha, beta)
    # beta == 0 must not read z (NaN stays out)
    zv = _v(np.full(n, np.nan))
    P.psb_gemlt(1.0, _v(x), _v(y), 0.0, zv)
    assert np.all(np.isfinite(zv.get_vect()))


@pytest.mark.parametrize("n", [1, 100, 4096, 1 << 22])
def test_reductions(gpu, orc, n):
    print("case reductions", n, flush=True)
    x = orc.uniform_pm1(5, n)
    y = orc.uniform_pm1(6, n)
    xv, yv = _v(x), _v(y)
    eps = np.finfo(float).eps
    d = P.psb_gedot(xv, yv)
    exact = math.fsum((x * y).tolist())
    assert abs(d - exact) <= 2 * math.log2(n + 1) * eps * np.sum(np.abs(x * y)) + 1e-300
    nr = P.psb_genrm2(xv)
    ex = math.sqrt(math.fsum((x * x).tolist()))
    assert abs(nr - ex) <= 2 * math.log2(n + 1) * eps * ex
    assert abs(nr - orc.nrm2(x)) <= 4 * math.log2(n + 1) * eps * ex
    assert P.psb_geamax(xv) == np.abs(x).max()
    s = P.psb_geasum(xv)
    assert abs(s - math.fsum(np.abs(x).tolist())) <= 2 * math.log2(n + 1) * eps * s
    # deterministic: the same call twice gives the same bits
    assert P.psb_gedot(xv, yv) == d


def test_reductions_fp32_accumulate_in_fp64(gpu, orc):
    print("case fp32 reductions", flush=True)
    n = 1 << 20
    x = orc.uniform_pm1(7, n).astype(np.float32)
    xv = _v(x, single=True)
    ex = math.sqrt(math.fsum((x.astype(np.float64) ** 2).tolist()))
    assert abs(P.psb_genrm2(xv) - ex) <= 1e-12 * ex


def test_nrm2_no_overflow(gpu):
    print("case nrm2 overflow", flush=True)
    assert P.psb_genrm2(_v(np.full(10, 1e200))) == pytest.approx(1e200 * math.sqrt(10), rel=1e-14)
    assert P.psb_genrm2(_v(np.full(10, 1e-200))) == pytest.approx(1e-200 * math.sqrt(10), rel=1e-14)
    assert P.psb_genrm2(_v(np.zeros(10))) == 0.0


def test_gather_scatter_insert(gpu, orc):
    print("case gath/scat/geins", flush=True)
    n = 1000
    x = orc.uniform_pm1(9, n)
    xv = P.psb_d_hip_vect().bld(x)
    idx = np.array([5, 1, 999, 1000, 17], np.int32)
    iv = P.psb_i_hip_vect().bld(idx)
    out = P.psb_d_hip_vect().all(len(idx))
    N.call("psb_hip_mv_igath", xv.dev, 0, len(idx), 1, N.fn("psb_hip_mv_data")(iv.dev), 1,
           N.fn("psb_hip_mv_data")(out.dev), 1)
    out.set_dev()
    assert np.array_equal(out.get_vect(), x[idx - 1])
    vals = np.arange(1.0, 6.0)
    vv = P.psb_d_hip_vect().bld(vals)
    N.call("psb_hip_mv_iscat", xv.dev, 0, len(idx), 1, N.fn("psb_hip_mv_data")(iv.dev), 1,
           N.fn("psb_hip_mv_data")(vv.dev), 1, 2.0)
    xv.set_dev()
    e = x.copy()
    e[idx - 1] = 2.0 * x[idx - 1] + vals
    assert np.array_equal(xv.get_vect(), e)
    # geins with dupl = add and a repeated index
    idx2 = P.psb_i_hip_vect().bld(np.array([3, 3, 7], np.int32))
    v2 = P.psb_d_hip_vect().bld(np.array([1.0, 2.0, 4.0]))
    z = P.psb_d_hip_vect().all(10)
    N.call("psb_hip_mv_geins", 3, idx2.dev, v2.dev, N.INS_ADD, 1, z.dev)
    z.set_dev()
    assert z.get_vect().tolist() == [0, 0, 3.0, 0, 0, 0, 4.0, 0, 0, 0]


def test_setscal_first_last(gpu):
    print("case setscal", flush=True)
    v = P.psb_d_hip_vect().all(20)
    v.set_scal(3.0, 5, 9)
    v.set_dev()
    got = v.get_vect()
    assert got[4:9].tolist() == [3.0] * 5 and got[:4].sum() == 0 and got[9:].sum() == 0
    d = P.psb_d_hip_vect().bld(np.array([2.0, 0.0, -4.0, 8.0]))
    N.call("psb_hip_mv_diag_inv", 3, d.dev)
    d.set_dev()
    assert d.get_vect().tolist() == [0.5, 1.0, -0.25, 1.0]   # zero -> 1; beyond n -> 1


def test_vector_plugin_methods(gpu, orc):
    """The type-bound procedures of the device vector (dot_v, axpby_v,
    upd_xyz, xyzw, mlt_v, mlt_v_2, scal, nrm2/amax/asum, absval, ins_v/ins_a)
    against the oracle's CPU routines."""
    print("case vector plug-in methods", flush=True)
    n = 3000
    x, y, z, w = (orc.uniform_pm1(s, n) for s in (11, 12, 13, 14))
    D = lambda a: P.psb_d_hip_vect().bld(a)
    xv, yv, zv, wv = D(x), D(y), D(z), D(w)
    assert abs(xv.dot_v(n, yv) - orc.dot(x, y)) <= 2 * n * np.finfo(float).eps * np.abs(x) @ np.abs(y)
    yv.axpby_v(n, 1.5, xv, -0.5)
    ye = orc.axpby(1.5, x, -0.5, y.copy())
    assert _bits_equal(yv.get_vect(), ye)
    yv2, zv2 = D(y), D(z)
    zv2.upd_xyz(n, 2.0, 0.5, -1.0, 3.0, xv, yv2)
    ye2, ze2 = orc.upd_xyz(2.0, 0.5, -1.0, 3.0, x, y.copy(), z.copy())
    assert _bits_equal(yv2.get_vect(), ye2) and _bits_equal(zv2.get_vect(), ze2)
    yv3, zv3 = D(y), D(z)
    wv.xyzw(n, 1.0, 2.0, 3.0, -1.0, 0.25, 4.0, xv, yv3, zv3)
    e = orc.xyzw(1.0, 2.0, 3.0, -1.0, 0.25, 4.0, x, y.copy(), z.copy(), w.copy())
    assert all(_bits_equal(g.get_vect(), r) for g, r in zip((yv3, zv3, wv), e))
    m = D(y)
    m.mlt_v(xv)
    assert _bits_equal(m.get_vect(), x * y)
    m2 = D(z)
    m2.mlt_v_2(2.0, xv, D(y), 0.5)
    assert _bits_equal(m2.get_vect(), orc.mlt(2.0, x, y, 0.5, z.copy()))
    s = D(x)
    s.scal(-3.0)
    assert _bits_equal(s.get_vect(), -3.0 * x)
    assert xv.amax() == np.max(np.abs(x))
    assert xv.asum() == pytest.approx(np.sum(np.abs(x)), rel=1e-13)
    assert xv.nrm2() == pytest.approx(np.linalg.norm(x), rel=1e-13)
    a = D(x)
    a.absval()
    assert _bits_equal(a.get_vect(), np.abs(x))
    b = P.psb_d_hip_vect().all(n)
    xv.absval(b)
    assert _bits_equal(b.get_vect(), np.abs(x))
    # insertion: overwrite, then add with a repeated index
    t = P.psb_d_hip_vect().all(10)
    t.ins_a(3, np.array([2, 9, 2], np.int32), np.array([1.0, 2.0, 5.0]), 1)
    assert t.get_vect().tolist() == [0, 6.0, 0, 0, 0, 0, 0, 0, 2.0, 0]
    t.ins_v(2, P.psb_i_hip_vect().bld(np.array([9, 1], np.int32)), D(np.array([-1.0, 7.0])), 0)
    assert t.get_vect().tolist() == [7.0, 6.0, 0, 0, 0, 0, 0, 0, -1.0, 0]
    with pytest.raises(P.PsbError):
        t.ins_a(1, np.array([11], np.int32), np.array([1.0]), 0)
    # asb grows the vector and keeps its values; cnv moves them to another type
    t.asb(14)
    assert t.get_nrows() == 14
    assert t.get_vect().tolist() == [7.0, 6.0, 0, 0, 0, 0, 0, 0, -1.0, 0, 0, 0, 0, 0]
    t.asb(5)
    assert t.get_nrows() == 14
    vt = P.psb_d_vect_type(D(x))
    vt.cnv(P.psb_s_hip_vect())
    assert vt.v._dtype == np.float32 and np.array_equal(vt.get_vect(), x.astype(np.float32))
    vt.cnv()
    assert vt.v._dtype == np.float64 and np.array_equal(vt.get_vect(), x.astype(np.float32))
    vt.free()


def test_halo_buffer_methods(gpu, orc):
    """gthzv_x / gthzbuf / sctb / sctb_x / sctb_buf: the halo-exchange halves
    (cuda/psb_d_cuda_vect_mod.F90:222-555), exact against numpy indexing."""
    print("case halo buffers", flush=True)
    n = 5000
    x = orc.uniform_pm1(21, n)
    xv = P.psb_d_hip_vect().bld(x)
    idx = np.array([17, 1, 5000, 250, 4999, 3, 1200, 77], np.int32)
    iv = P.psb_i_hip_vect().bld(idx)
    assert np.array_equal(xv.gthzv_x(3, 4, iv), x[idx[2:6] - 1])
    yh = np.full(6, -1.0)
    xv.gthzv_x(1, 5, iv, yh)
    assert np.array_equal(yh[:5], x[idx[:5] - 1]) and yh[5] == -1.0
    xv.new_buffer(len(idx))
    xv.gthzbuf(1, 3, iv)
    xv.gthzbuf(4, 5, iv)
    assert np.array_equal(xv.combuf, x[idx - 1])
    # sctb_buf: y(idx(i:i+n-1)) = beta*y + combuf(i:i+n-1)
    yv = P.psb_d_hip_vect().bld(np.ones(n))
    yv.new_buffer(len(idx))
    yv.combuf[:] = np.arange(1.0, 9.0)
    yv.sctb_buf(2, 6, iv, 0.5)
    ye = np.ones(n)
    ye[idx[1:7] - 1] = 0.5 + np.arange(2.0, 8.0)
    assert np.array_equal(yv.get_vect(), ye)
    # sctb_x from a host array, beta = 0
    yv.sctb_x(1, 2, iv, np.array([9.0, -9.0]), 0.0)
    ye[idx[:2] - 1] = [9.0, -9.0]
    assert np.array_equal(yv.get_vect(), ye)
    # sctb on the host copy (the reference's path), repeated index accumulates
    yv.sctb(3, np.array([4, 4, 10], np.int32), np.array([1.0, 2.0, 3.0]), 1.0)
    ye[3] += 3.0
    ye[9] += 3.0
    assert yv.is_host() and np.array_equal(yv.get_vect(), ye)
    with pytest.raises(P.PsbError):
        xv.gthzv_x(6, 4, iv)                    # window beyond idx
    xv.free()
    yv.free()


def test_stream_kernels(gpu):
    """The bandwidth probes compute what they claim: the read-only stream's
    per-block sums add up to the array sum, the SpMV-shaped stream writes
    y[i] = sum_k a[k*m + i] (16-byte elements)."""
    print("case stream kernels", flush=True)
    rng = np.random.default_rng(31)
    m, ratio = 4096, 12                          # 16-byte elements
    a = rng.integers(-8, 8, size=(ratio * m * 2)).astype(np.float64)   # exact sums
    nbytes = a.nbytes
    da, dy, dp = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for p, nb in ((da, nbytes), (dy, m * 16), (dp, 1 << 20)):
        N.call("psb_hip_alloc_buffer", C.byref(p), nb)
    try:
        N.call("psb_hip_write_buffer", da, a.ctypes.data, nbytes)
        N.call("psb_hip_write_buffer", dp, np.zeros(1 << 17).ctypes.data, 1 << 20)
        N.call("psb_hip_stream_read", da, nbytes, dp)
        N.call("psb_hip_device_sync")
        part = np.empty(1 << 17)
        N.call("psb_hip_read_buffer", part.ctypes.data, dp, 1 << 20)
        assert part.sum() == a.sum()
        N.call("psb_hip_stream_mix", da, ratio, m * 16, dy)
        N.call("psb_hip_device_sync")
        y = np.empty(2 * m)
        N.call("psb_hip_read_buffer", y.ctypes.data, dy, m * 16)
        cols = a[:ratio * m * 2].reshape(ratio, m, 2)      # column k, element i, two doubles
        assert np.array_equal(y.reshape(m, 2), cols.sum(axis=0))
    finally:
        for p in (da, dy, dp):
            N.call("psb_hip_free_buffer", p)
