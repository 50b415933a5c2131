"""The rest of the matrix plug-in's methods on the device copy: scals / scal
(psb_d_cuda_{hlg,elg,csrg}_scals / _scal, which the reference runs as the
CPU psb_d_{hll,ell,csr}_scal[s] then to_gpu), make_nonunit, the host-array
products csmv / csmm, cp_from_fmt / mv_from_coo / mv_from_fmt and the
dual-memory state methods.  Scaling is one multiply per value, so the device
arrays must equal the CPU rule applied to the arrays read back, bit for bit;
which values each format scales (padding or not) follows the CPU code."""
import numpy as np
import pytest

import psblas3_amd as P
from psblas3_amd import _native as N

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def _molds(single):
    if single:
        return {"hlg": lambda: P.psb_s_hip_hlg_sparse_mat(32), "elg": P.psb_s_hip_elg_sparse_mat,
                "csrg": P.psb_s_hip_csrg_sparse_mat}
    return {"hlg": lambda: P.psb_d_hip_hlg_sparse_mat(32), "elg": P.psb_d_hip_elg_sparse_mat,
            "csrg": P.psb_d_hip_csrg_sparse_mat}


def _matrix(orc):
    """Convection-diffusion L=9 plus a few empty and long rows, so that
    hacks and ELL rows carry padding."""
    coo = orc.gen_convdiff(9, conv=2.0)
    keep = (coo.ia % 37) != 0                     # some rows emptied
    ia, ja, val = coo.ia[keep], coo.ja[keep], coo.val[keep]
    extra_i = np.full(20, 100, np.int32)
    extra_j = np.arange(200, 220, dtype=np.int32)
    ia = np.concatenate([ia, extra_i]); ja = np.concatenate([ja, extra_j])
    val = np.concatenate([val, np.linspace(-1.0, 1.0, 20)])
    return P.psb_d_coo_sparse_mat(coo.nr, coo.nc, ia, ja, val).fix()


def _arrays(m, fmt):
    """(values, rows, cols, entry-mask) of the whole device value array: row
    and column of every slot and whether it holds an entry (not padding)."""
    if fmt == "csrg":
        irp, ja, val = m.to_host()
        rows = np.repeat(np.arange(m.nr), np.diff(irp))
        return val, rows, ja - 1, np.ones(len(val), bool)
    if fmt == "hlg":
        hk, irn, ja, val, _ = m.to_host()
        hs = m.hksz
        rows = np.empty(len(val), np.int64)
        ent = np.zeros(len(val), bool)
        for h in range(len(hk) - 1):
            w = (hk[h + 1] - hk[h]) // hs
            for j in range(w):
                for lane in range(hs):
                    p = hk[h] + j * hs + lane
                    r = h * hs + lane
                    rows[p] = min(r, m.nr - 1)
                    ent[p] = r < m.nr and j < irn[r]
        return val, rows, ja - 1, ent
    irn, ja, val, _ = m.to_host()
    nzm = len(val) // m.nr
    rows = np.tile(np.arange(m.nr), nzm)
    ent = np.concatenate([np.arange(m.nr) * 0 + j < irn for j in range(nzm)])
    return val, rows, ja - 1, ent


@pytest.mark.parametrize("fmt", ["hlg", "elg", "csrg"])
@pytest.mark.parametrize("single", [False, True], ids=["d", "s"])
def test_scals_scal(gpu, orc, fmt, single):
    _say(f"case scals/scal {fmt} single={single}")
    coo = _matrix(orc)
    dt = np.float32 if single else np.float64
    m = _molds(single)[fmt]().cp_from_coo(coo)
    val0, rows, cols, ent = _arrays(m, fmt)
    s = dt(1.75)
    m.scals(1.75)
    val1 = _arrays(m, fmt)[0]
    # every stored value, padding included (hll / ell scals the whole array)
    assert np.array_equal(val1, val0 * s)
    dl = orc.uniform_pm1(31, m.nr).astype(dt)
    m.scal(dl, side="L")
    val2 = _arrays(m, fmt)[0]
    exp = val1.copy()
    sel = ent | (fmt == "elg")                    # ELL's left scaling takes the whole row
    exp[sel] = val1[sel] * dl[rows[sel]]
    assert np.array_equal(val2, exp)
    dr = orc.uniform_pm1(32, m.nc).astype(dt)
    dv = P.psb_s_hip_vect().bld(dr) if single else P.psb_d_hip_vect().bld(dr)
    m.scal(dv, side="R")
    val3 = _arrays(m, fmt)[0]
    exp = val2.copy()
    exp[ent] = val2[ent] * dr[cols[ent]]
    assert np.array_equal(val3, exp)
    with pytest.raises(P.PsbError):
        m.scal(dl[:-1], side="L")
    dv.free()
    m.free()


def test_scaling_follows_single_copy(gpu, orc):
    """A double HLL matrix's FP32 copy (psb_hip_hll_make_single) follows
    scals / scal: its product equals that of a copy made after scaling."""
    _say("case scals with FP32 copy")
    oc = orc.gen_convdiff(10, conv=1.0)
    coo = P.psb_d_coo_sparse_mat(oc.nr, oc.nc, oc.ia, oc.ja, oc.val).fix()
    a = P.psb_d_hip_hlg_sparse_mat(64).cp_from_coo(coo)
    b = P.psb_d_hip_hlg_sparse_mat(64).cp_from_coo(coo)
    N.call("psb_hip_hll_make_single", a.dev)
    d = orc.uniform_pm1(33, coo.nr)
    a.scals(0.5)
    a.scal(d, side="L")
    b.scals(0.5)
    b.scal(d, side="L")
    N.call("psb_hip_hll_make_single", b.dev)
    x = P.psb_s_hip_vect().bld(orc.uniform_pm1(34, coo.nr).astype(np.float32))
    ya = P.psb_s_hip_vect().bld(np.zeros(coo.nr, np.float32))
    yb = P.psb_s_hip_vect().bld(np.zeros(coo.nr, np.float32))
    N.call("psb_hip_hll_spmv_single", a.dev, 1.0, x.dev, 0.0, ya.dev)
    N.call("psb_hip_hll_spmv_single", b.dev, 1.0, x.dev, 0.0, yb.dev)
    ya.set_dev(); yb.set_dev()
    assert np.array_equal(ya.get_vect(), yb.get_vect())
    for v in (x, ya, yb):
        v.free()
    a.free()
    b.free()


def test_make_nonunit_then_scals(gpu, orc):
    """A unit triangle stores its unit diagonal before it is scaled (the
    reference's scals calls make_nonunit first); the result is a non-unit
    triangle whose solve matches the oracle's."""
    _say("case make_nonunit + scals + solve")
    from helpers import tri_csr
    irp, ja, val = tri_csr(orc, 80, True, True, seed=51)
    desc = P.psb_cdall(nl=80)
    a = P.psb_spall(desc, len(val))
    a.set_triangle(); a.set_lower(); a.set_unit()
    P.psb_spins(len(val), np.repeat(np.arange(1, 81, dtype=np.int32), np.diff(irp)), ja, val,
                a, desc)
    P.psb_spasb(a, desc, mold=P.psb_d_hip_csrg_sparse_mat())
    a.a.scals(2.0)
    assert not a.a.is_unit() and a.a.is_triangle() and a.a.is_lower()
    irp2, ja2, val2 = a.a.to_csr()
    assert len(val2) == len(val) + 80
    x = orc.uniform_pm1(35, 80)
    y = P.vect_from(np.zeros(80))
    a.a.inner_vect_sv(1.0, P.vect_from(x).v, 0.0, y.v)
    ref = orc.csr_cssv(irp2, ja2, val2, x, lower=True, unit=False)
    assert np.array_equal(y.get_vect(), ref)
    assert np.all(val2[irp2[1:] - 2] == 2.0)         # the stored diagonal, scaled
    a.free()


@pytest.mark.parametrize("fmt", ["hlg", "elg", "csrg"])
def test_csmv_csmm_host_arrays(gpu, orc, fmt):
    _say(f"case csmv / csmm {fmt}")
    coo = _matrix(orc)
    m = _molds(False)[fmt]().cp_from_coo(coo)
    x = orc.uniform_pm1(36, coo.nc)
    y0 = orc.uniform_pm1(37, coo.nr)
    y = m.csmv(1.5, x, -0.5, y0)
    yv = P.psb_d_hip_vect().bld(y0)
    m.vect_mv(1.5, P.psb_d_hip_vect().bld(x), -0.5, yv)
    assert np.array_equal(y, yv.get_vect())
    xk = np.asfortranarray(np.stack([orc.uniform_pm1(38 + q, coo.nc) for q in range(3)], 1))
    yk = m.csmm(1.0, xk, 0.0, np.zeros((coo.nr, 3)))
    ykv = P.psb_d_hip_vect().bld(np.zeros((coo.nr, 3)))
    m.vect_mv(1.0, P.psb_d_hip_vect().bld(xk), 0.0, ykv)
    assert np.array_equal(np.asarray(yk), np.asarray(ykv.get_vect()))
    # each column within the per-row bound of the CSR loop
    irp, ja, val = m.to_csr()
    for q in range(3):
        ref = orc.csr_spmv(irp, ja, val, xk[:, q].copy())
        bound = orc.csr_spmv(irp, ja, np.abs(val), np.abs(xk[:, q]))
        assert np.all(np.abs(np.asarray(yk)[:, q] - ref) <= 4 * 25 * 2.2e-16 * bound)
    # trans 'T': the device transposed product, equal to the CPU loop's bits
    xt = orc.uniform_pm1(41, coo.nr)
    yt0 = orc.uniform_pm1(42, coo.nc)
    yt = m.csmv(2.0, xt, 0.5, yt0, trans="T")
    irp, ja, val = m.to_csr()
    assert np.array_equal(yt, orc.csr_spmv_t(irp, ja, val, xt, 2.0, 0.5, yt0, nc=coo.nc))
    with pytest.raises(P.PsbError):
        m.csmv(1.0, x, 0.0, y0, trans="X")
    m.free()


def test_format_conversions_and_state(gpu, orc):
    _say("case cp_from_fmt / mv_from_coo / mv_from_fmt / state")
    coo = _matrix(orc)
    irp, ja, val = orc.csr_from_coo(orc.fix_coo(orc.COO(coo.nr, coo.nc, coo.ia, coo.ja,
                                                        coo.val)))
    c = P.psb_d_hip_csrg_sparse_mat().cp_from_coo(coo)
    h = P.psb_d_hip_hlg_sparse_mat(64).cp_from_fmt(c)
    e = P.psb_d_hip_elg_sparse_mat().mv_from_fmt(h)
    assert h.dev is None
    for m in (c, e):
        i2, j2, v2 = m.to_csr()
        assert np.array_equal(i2, irp) and np.array_equal(j2, ja) and np.array_equal(v2, val)
        assert m.is_dev() and not m.is_host() and not m.is_sync()
        m.sync()
        m.to_gpu()
    coo2 = P.psb_d_coo_sparse_mat(coo.nr, coo.nc, coo.ia.copy(), coo.ja.copy(), coo.val.copy())
    coo2.fix()
    c2 = P.psb_d_hip_csrg_sparse_mat().mv_from_coo(coo2)
    assert coo2.get_nzeros() == 0
    assert np.array_equal(c2.to_csr()[2], val)
    for m in (c, e, c2):
        m.free()
