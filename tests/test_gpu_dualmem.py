"""Dual-memory semantics of the matrix plug-ins (cuda/psb_d_cuda_hlg_mat_mod.F90:
40-84 host / sync / dev state; cuda/impl/psb_d_cuda_hlg_vect_mv.F90:77-108
the host product when a vector is not a device vector), for HLG (rows in
order and row-sorted), ELG and CSRG, FP64 and FP32:

- a matrix built on the host is "host"; a device product syncs it first
  ("sync") and equals, bit for bit, the product of the same matrix built on
  the device ("dev");
- a device matrix with host vectors computes on its host copy (synced from
  the device first): equal bit for bit to the oracle's CSR product (the CPU
  arithmetic), and within the per-row bound of the device product;
- the transposed product with host vectors equals the oracle's CPU loop;
- state transitions: scaling on the device leaves "dev", sync reads the
  device back, values changed on the host copy reach the device on sync.
"""
import numpy as np
import pytest

import psblas3_amd as P

pytestmark = pytest.mark.gpu

MOLDS = {
    "HLG": lambda s: (P.psb_s_hip_hlg_sparse_mat() if s else P.psb_d_hip_hlg_sparse_mat()),
    "HLG32sorted": lambda s: (P.psb_s_hip_hlg_sparse_mat(32, sort_window=0) if s
                              else P.psb_d_hip_hlg_sparse_mat(32, sort_window=0)),
    "ELG": lambda s: (P.psb_s_hip_elg_sparse_mat() if s else P.psb_d_hip_elg_sparse_mat()),
    "CSRG": lambda s: (P.psb_s_hip_csrg_sparse_mat() if s else P.psb_d_hip_csrg_sparse_mat()),
}


def _say(*a):
    print(*a, flush=True)


def _coo(orc):
    return orc.fix_coo(orc.gen_powerlaw(3000, cap=300, seed=3))


def _pc(f):
    c = P.psb_d_coo_sparse_mat(f.nr, f.nc, f.ia, f.ja, f.val)
    c.sorted = True
    return c


def _wrap(a_, n):
    m = P.psb_dspmat_type()
    m.m = m.n = n
    m.a = a_
    return m


@pytest.mark.parametrize("fmt", list(MOLDS))
@pytest.mark.parametrize("single", [False, True], ids=["d", "s"])
def test_build_on_host_then_device_product(gpu, orc, fmt, single):
    _say("case dual-memory host build", fmt, "single" if single else "double")
    f = _coo(orc)
    dt = np.float32 if single else np.float64
    vm = P.psb_s_hip_vect() if single else P.psb_d_hip_vect()
    h = _wrap(MOLDS[fmt](single).cp_from_coo(_pc(f), on_host=True), f.nr)
    d = _wrap(MOLDS[fmt](single).cp_from_coo(_pc(f)), f.nr)
    assert h.a.is_host() and not h.a.is_dev() and d.a.is_dev()
    # the host copy built on the host equals the device arrays read back
    hc = h.a.host_copy()
    dc = d.a.host_copy()          # reads the device back: "sync"
    assert d.a.is_sync()
    for u, v in zip(hc, dc):
        assert np.array_equal(np.asarray(u), np.asarray(v))
    x = P.vect_from(orc.uniform_pm1(1, f.nc).astype(dt), mold=vm)
    y1 = P.vect_from(np.zeros(f.nr, dt), mold=vm)
    y2 = P.vect_from(np.zeros(f.nr, dt), mold=vm)
    P.psb_spmm(1.0, h, x, 0.0, y1)          # syncs the host-built matrix to the device
    assert h.a.is_sync()
    P.psb_spmm(1.0, d, x, 0.0, y2)
    assert np.array_equal(y1.get_vect(), y2.get_vect())
    h.free()
    d.free()


@pytest.mark.parametrize("fmt", list(MOLDS))
@pytest.mark.parametrize("ab", [(1.0, 0.0), (-1.5, 0.5)])
def test_device_matrix_host_vectors(gpu, orc, fmt, ab):
    alpha, beta = ab
    _say("case dual-memory host vectors", fmt, alpha, beta)
    f = _coo(orc)
    irp, ja, val = orc.csr_from_coo(f)
    x = orc.uniform_pm1(2, f.nc)
    y0 = orc.uniform_pm1(3, f.nr)
    a = _wrap(MOLDS[fmt](False).cp_from_coo(_pc(f)), f.nr)
    assert a.a.is_dev()
    xh = P.psb_d_vect_type(P.psb_d_base_vect().bld(x))
    yh = P.psb_d_vect_type(P.psb_d_base_vect().bld(y0))
    P.psb_spmm(alpha, a, xh, beta, yh)
    assert a.a.is_sync()            # synced to compute on the host copy
    yref = orc.csr_spmv(irp, ja, val, x, alpha, beta, y0.copy())
    got = yh.get_vect()
    # the CPU arithmetic: equal to the oracle's product wherever it is nonzero
    nz = yref != 0
    assert np.array_equal(got[nz].view(np.int64), yref[nz].view(np.int64))
    assert np.all(got[~nz] == 0)
    # the device product on device vectors stays within the per-row bound
    yd = P.vect_from(y0.copy())
    P.psb_spmm(alpha, a, P.vect_from(x), beta, yd)
    absax = orc.csr_spmv(irp, ja, np.abs(val), np.abs(x))
    bound = 4.0 * (np.diff(irp) + 1) * np.finfo(float).eps * (abs(alpha) * absax +
                                                               abs(beta) * np.abs(y0))
    assert np.all(np.abs(yd.get_vect() - got) <= 2 * bound + 1e-300)
    # transposed with host vectors: the CPU loop's order, bit for bit
    yt = P.psb_d_vect_type(P.psb_d_base_vect().bld(orc.uniform_pm1(4, f.nc)))
    y0t = yt.get_vect()
    xt = P.psb_d_vect_type(P.psb_d_base_vect().bld(orc.uniform_pm1(5, f.nr)))
    P.psb_spmm(alpha, a, xt, beta, yt, trans="T")
    ytref = orc.csr_spmv_t(irp, ja, val, xt.get_vect(), alpha, beta, y0t.copy(), nc=f.nc)
    assert np.array_equal(yt.get_vect(), ytref)
    a.free()


def test_state_transitions(gpu, orc):
    _say("case dual-memory state transitions")
    f = _coo(orc)
    irp, ja, val = orc.csr_from_coo(f)
    x = orc.uniform_pm1(6, f.nc)
    a = _wrap(P.psb_d_hip_hlg_sparse_mat().cp_from_coo(_pc(f), on_host=True), f.nr)
    assert a.a.is_host()
    a.a.sync()
    assert a.a.is_sync()
    a.a.scals(2.0)                      # scaled on the device copy
    assert a.a.is_dev()
    hk, irn, hja, hval, idiag = a.a.host_copy()   # read back: sync
    assert a.a.is_sync()
    # the host copy now holds the scaled values
    h0 = orc.hll_from_coo(f, 64)
    assert np.array_equal(hval, 2.0 * h0.val)
    # change the host copy, mark it current, the next device product sees it
    hval *= 0.5
    a.a.set_host()
    y = P.vect_from(np.zeros(f.nr))
    P.psb_spmm(1.0, a, P.vect_from(x), 0.0, y)
    assert a.a.is_sync()
    ref = _wrap(P.psb_d_hip_hlg_sparse_mat().cp_from_coo(_pc(f)), f.nr)
    yr = P.vect_from(np.zeros(f.nr))
    P.psb_spmm(1.0, ref, P.vect_from(x), 0.0, yr)
    assert np.array_equal(y.get_vect(), yr.get_vect())
    # a device product on a "dev" matrix does not touch the host copy
    a.a.set_dev()
    P.psb_spmm(1.0, a, P.vect_from(x), 0.0, y)
    assert a.a.is_dev()
    a.free()
    ref.free()
