"""Parity at the benchmark sizes: the workloads bench.py measures, checked
against the oracle (the reference's own GPU benchmark compares the GPU
product with the CPU product on its benchmark matrix,
test/cudakern/dpdegenmv.F90:815-821,848-852).

- 3D convection-diffusion L = 200 (8.0M rows, 55.76M entries): HLL64
  conversion bit for bit; CSR / ELL / HLL64 products in FP64 and FP32 within
  the per-row bound (31,250 blocks of XCD-chunked grids, 117M-slot offsets).
- power-law n = 4M (Zipf 2, cap 4096): CSR on its automatic column slices and
  on the whole matrix (rows of <= 512 entries bit for bit with beta = 0), HLL
  with rows in order and row-sorted, within the per-row bound.
- SPD Poisson L = 256 (16.8M rows): FP64 CG and both mixed-precision forms to
  1e-10, each final residual recomputed by the oracle's CSR product.
- the 256-system anisotropic batch: every system's final FP64 residual
  recomputed on the host.
- the 64-matrix serving set: a sample of requests served in the mix
  compared bit for bit with the same requests served alone.
"""
import numpy as np
import pytest

import psblas3_amd as P
from psblas3_amd import _native as N

# assembling and checking at these sizes takes tens of seconds on the host
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def _say(*a):
    print(*a, flush=True)


def _dev_matrix(coo, mold):
    """Device matrix from an oracle-fixed COO (already sorted and merged)."""
    pc = P.psb_d_coo_sparse_mat(coo.nr, coo.nc, coo.ia, coo.ja, coo.val)
    pc.sorted = True
    m = P.psb_dspmat_type()
    m.m, m.n = coo.nr, coo.nc
    m.a = mold.cp_from_coo(pc)
    return m


def _absprod(orc, csr, x):
    irp, ja, val = csr
    return orc.csr_spmv(irp, ja, np.abs(val), np.abs(x))


def _check_bound(y, yref, absax, nnz_row, eps, alpha=1.0, factor=1.0):
    bound = factor * 4.0 * (nnz_row + 1) * eps * abs(alpha) * absax + 1e-300
    bad = np.abs(y - yref) > bound
    assert not bad.any(), (int(bad.sum()), float(np.max(np.abs(y - yref) - bound)))


@pytest.fixture(scope="module")
def stencil200(orc):
    _say("[fullsize] assembling the L=200 stencil")
    f = orc.fix_coo(orc.gen_convdiff(200))
    csr = orc.csr_from_coo(f)
    return f, csr


def test_fullsize_stencil_L200(gpu, orc, stencil200):
    f, csr = stencil200
    irp = csr[0]
    nnz_row = np.diff(irp).astype(np.float64)
    x = orc.uniform_pm1(1, f.nc)
    yref = orc.csr_spmv(*csr, x)
    absax = _absprod(orc, csr, x)
    xv = P.vect_from(x)
    yv = P.vect_from(np.zeros(f.nr))
    _say("[fullsize] L=200 HLL64 conversion")
    hll = _dev_matrix(f, P.psb_d_hip_hlg_sparse_mat(64))
    hk, irn, ja, val, idiag = hll.a.to_host()
    h = orc.hll_from_coo(f, 64)
    assert np.array_equal(hk, h.hkoffs) and np.array_equal(irn, h.irn)
    assert np.array_equal(ja, h.ja) and np.array_equal(idiag, h.idiag)
    assert np.array_equal(val.view(np.int64), h.val.view(np.int64))
    # closed form: hack widths are 7 except where a hack holds only boundary
    # rows, so the slots are 64 * sum(widths) and at least one per entry
    assert hk[-1] == 64 * int(sum(irn[k:k + 64].max() for k in range(0, f.nr, 64)))
    assert f.nz <= hk[-1] <= 7 * (f.nr + 63) // 64 * 64
    del hk, irn, ja, val, idiag, h
    for name, mat in (("HLL64", hll), ("ELL", None), ("CSR", None)):
        if mat is None:
            mold = P.psb_d_hip_elg_sparse_mat() if name == "ELL" else P.psb_d_hip_csrg_sparse_mat()
            mat = _dev_matrix(f, mold)
        _say(f"[fullsize] L=200 {name} FP64 product")
        P.psb_spmm(1.0, mat, xv, 0.0, yv)
        _check_bound(yv.get_vect(), yref, absax, nnz_row, np.finfo(np.float64).eps)
        mat.free()
    # FP32: values and x rounded to single, the oracle in single precision
    x32 = x.astype(np.float32)
    y32ref = orc.csr_spmv(csr[0], csr[1], csr[2].astype(np.float32), x32).astype(np.float64)
    xs = P.vect_from(x32, mold=P.psb_s_hip_vect())
    ys = P.vect_from(np.zeros(f.nr, np.float32), mold=P.psb_s_hip_vect())
    for name, mold in (("HLL64", P.psb_s_hip_hlg_sparse_mat(64)),
                       ("ELL", P.psb_s_hip_elg_sparse_mat()), ("CSR", P.psb_s_hip_csrg_sparse_mat())):
        _say(f"[fullsize] L=200 {name} FP32 product")
        mat = _dev_matrix(f, mold)
        P.psb_spmm(1.0, mat, xs, 0.0, ys)
        _check_bound(ys.get_vect().astype(np.float64), y32ref, absax, nnz_row,
                     np.finfo(np.float32).eps, factor=2.0)
        mat.free()


def test_fullsize_spmm_L200(gpu, orc, stencil200):
    """Multi-RHS products at the benchmark size, k = 8 (two rows per lane) and
    k = 12 (one row per lane): every column equals the single-vector product
    of that column bit for bit (both sum each row in slot order from zero),
    and the single-vector product is held to the oracle's per-row bound in
    test_fullsize_stencil_L200."""
    f, _ = stencil200
    hll = _dev_matrix(f, P.psb_d_hip_hlg_sparse_mat(64))
    y1 = P.vect_from(np.zeros(f.nr))
    for k in (8, 12):
        _say(f"[fullsize] L=200 HLL64 SpMM k={k}")
        X = np.stack([orc.uniform_pm1(100 + s, f.nc) for s in range(k)], axis=1)
        yv = P.vect_from(np.zeros((f.nr, k)))
        P.psb_spmm(1.0, hll, P.vect_from(X), 0.0, yv)
        Y = yv.get_vect()
        for col in range(k):
            P.psb_spmm(1.0, hll, P.vect_from(X[:, col].copy()), 0.0, y1)
            assert np.array_equal(Y[:, col], y1.get_vect()), (k, col)
        del X, Y, yv
    hll.free()


@pytest.fixture(scope="module")
def powerlaw4m(orc):
    _say("[fullsize] assembling the n=4M power-law matrix")
    f = orc.fix_coo(orc.gen_powerlaw(4_000_000, cap=4095, seed=1))
    csr = orc.csr_from_coo(f)
    return f, csr


def test_fullsize_powerlaw_n4M(gpu, orc, powerlaw4m):
    f, csr = powerlaw4m
    irp = csr[0]
    nnz_row = np.diff(irp).astype(np.float64)
    x = orc.uniform_pm1(1, f.nc)
    y0 = orc.uniform_pm1(2, f.nr)
    yref = orc.csr_spmv(*csr, x)
    yref_b = orc.csr_spmv(*csr, x, 1.5, -0.5, y0.copy())
    absax = _absprod(orc, csr, x)
    xv = P.vect_from(x)
    _say("[fullsize] power-law CSR, automatic column slices")
    m = _dev_matrix(f, P.psb_d_hip_csrg_sparse_mat())
    assert N.fn("psb_hip_csr_slices")(m.a.dev) >= 2
    yv = P.vect_from(np.zeros(f.nr))
    P.psb_spmm(1.0, m, xv, 0.0, yv)
    _check_bound(yv.get_vect(), yref, absax, nnz_row, np.finfo(np.float64).eps)
    yv = P.vect_from(y0)
    P.psb_spmm(1.5, m, xv, -0.5, yv)
    bound_b = 4.0 * (nnz_row + 1) * np.finfo(np.float64).eps * (1.5 * absax + 0.5 * np.abs(y0))
    assert np.all(np.abs(yv.get_vect() - yref_b) <= bound_b)
    _say("[fullsize] power-law CSR, whole matrix (row-block kernel)")
    old = N.fn("psb_hip_csr_set_variant")(3)
    try:
        yv = P.vect_from(np.zeros(f.nr))
        P.psb_spmm(1.0, m, xv, 0.0, yv)
    finally:
        N.fn("psb_hip_csr_set_variant")(old)
    y = yv.get_vect()
    short = np.diff(irp) <= 512
    assert np.array_equal(y[short].view(np.int64), yref[short].view(np.int64))
    _check_bound(y, yref, absax, nnz_row, np.finfo(np.float64).eps)
    m.free()
    for name, mold in (("HLL64 rows in order", P.psb_d_hip_hlg_sparse_mat(64)),
                       ("HLL64 sorted", P.psb_d_hip_hlg_sparse_mat(64, sort_window=0)),
                       ("HLL64 sorted 64K", P.psb_d_hip_hlg_sparse_mat(64, sort_window=65536))):
        _say(f"[fullsize] power-law {name}")
        m = _dev_matrix(f, mold)
        if "sorted" in name:
            assert m.a.isz <= 1.2 * f.nz
            # x is 32 MB with scattered columns: the automatic column slices
            assert N.fn("psb_hip_hll_slices")(m.a.dev) >= 2
        yv = P.vect_from(np.zeros(f.nr))
        P.psb_spmm(1.0, m, xv, 0.0, yv)
        _check_bound(yv.get_vect(), yref, absax, nnz_row, np.finfo(np.float64).eps)
        m.free()


def test_fullsize_cg_poisson_L256(gpu, orc):
    _say("[fullsize] assembling the SPD Poisson L=256")
    L = 256
    f = orc.fix_coo(orc.gen_poisson3d(L))
    csr = orc.csr_from_coo(f)
    n = f.nr
    xt = orc.uniform_pm1(2, n)
    bh = orc.csr_spmv(*csr, xt)
    bn = np.linalg.norm(bh)
    a = _dev_matrix(f, P.psb_d_hip_hlg_sparse_mat(64))
    desc = P.psb_cdall(nl=n)
    prec = P.psb_dprec_type("DIAG").build(a, desc)
    b = P.vect_from(bh)
    errs = {}
    for kind in ("fp64", "restart", "reliable"):
        _say(f"[fullsize] L=256 CG {kind}")
        x = P.vect_from(np.zeros(n))
        if kind == "fp64":
            it, err, _ = P.psb_krylov_fused("CG", a, prec, b, x, 1e-10, desc, itmax=5000)
            info = f"{it} its"
        else:
            oi, ii, err = P.psb_krylov_mixed("CG", a, prec, b, x, 1e-10, desc, outer_itmax=50,
                                             inner_eps=1e-4, inner_itmax=20000, variant=kind)
            info = f"{oi} outer, {ii} inner its"
        xs = x.get_vect()
        rr = np.linalg.norm(bh - orc.csr_spmv(*csr, xs)) / bn
        errs[kind] = np.linalg.norm(xs - xt) / np.linalg.norm(xt)
        _say(f"  {info}: reported {err:.3e}, oracle residual {rr:.3e}, "
             f"|x - xt|/|xt| {errs[kind]:.3e}")
        assert err <= 1e-10
        # the recomputed residual: the reported one, up to the rounding of
        # the two products (a few ulps of ||b|| relative)
        assert rr <= 1e-10 * 1.05
    for kind in ("restart", "reliable"):
        assert errs[kind] <= 2.0 * errs["fp64"]
    a.free()


def test_fullsize_batch_256(gpu, orc):
    """The benchmark's 256 anisotropic systems (side log-uniform in [32, 512],
    eps log-uniform in [1e-3, 1], seed 3) to 1e-8 in one engine."""
    nsys, seed = 256, 3
    u = orc.uniform_pm1(seed, nsys, stream=0)
    ue = orc.uniform_pm1(seed, nsys, stream=1)
    Ls = np.rint(32.0 * 16.0 ** ((u + 1.0) / 2.0)).astype(int)
    es = 10.0 ** (-3.0 + 3.0 * (ue + 1.0) / 2.0)
    _say(f"[fullsize] batch of {nsys} systems, {int((Ls ** 2).sum())} rows")
    systems, total = [], 0
    for k in range(nsys):
        f = orc.fix_coo(orc.gen_aniso2d(int(Ls[k]), float(es[k])))
        csr = orc.csr_from_coo(f)
        a = _dev_matrix(f, P.psb_d_hip_hlg_sparse_mat(64))
        prec = P.psb_dprec_type("DIAG").build(a, P.psb_cdall(nl=f.nr))
        bh = orc.csr_spmv(*csr, orc.uniform_pm1(seed, f.nr, stream=1000 + k))
        systems.append((a, prec, csr, bh))
        n = f.nr
        total += 3 * ((8 * n + 255) // 256 * 256) + 5 * ((4 * n + 255) // 256 * 256)
    eng = P.psb_hip_batch("CG", max_systems=nsys, arena_bytes=total + (1 << 20), check_every=8)
    jobs = []
    for a, prec, csr, bh in systems:
        mid = eng.register(a, prec)
        jobs.append((mid, P.vect_from(bh), P.vect_from(np.zeros(len(bh)))))
    out = eng.solve(jobs, eps=1e-8, itmax=20000)
    eng.free()
    worst = 0.0
    for (a, prec, csr, bh), (_, _, xv), o in zip(systems, jobs, out):
        rr = np.linalg.norm(bh - orc.csr_spmv(*csr, xv.get_vect())) / np.linalg.norm(bh)
        worst = max(worst, rr)
        assert o["status"] == 1 and rr <= 1e-8 * 1.01, (o, rr)
        assert abs(o["relres"] - rr) <= 1e-3 * rr + 1e-15
        a.free()
    _say(f"  all {nsys} converged, worst recomputed residual {worst:.3e}")


def test_fullsize_serving_sample(gpu, orc):
    """The 64 resident matrices of the serving benchmark (L uniform in
    [24, 160], convection uniform in [0, 10], seed 4): 48 requests at 1000
    req/s served in the mix; 4 of them served again alone must give the same
    bits, and their residuals are recomputed by the oracle."""
    nmat, seed = 64, 4
    u = orc.uniform_pm1(seed, nmat, stream=0)
    w = orc.uniform_pm1(seed, nmat, stream=1)
    Ls = np.minimum(24 + ((u + 1.0) / 2.0 * 137).astype(int), 160)
    convs = (w + 1.0) / 2.0 * 10.0
    _say(f"[fullsize] serving set: {nmat} matrices, {int((Ls ** 3).sum())} rows")
    mats = []
    fn = N.fn("psb_hip_gen_pde3d")
    a3 = np.ones(3)
    for L, cv in zip(Ls, convs):
        L = int(L)
        b3 = np.full(3, cv / np.sqrt(3.0))
        nz = fn(L, a3.ctypes.data, b3.ctypes.data, 0.0, 1, 0, None, None, None, None)
        ia = np.empty(nz, np.int32); ja = np.empty(nz, np.int32); val = np.empty(nz)
        fn(L, a3.ctypes.data, b3.ctypes.data, 0.0, 1, 0, ia.ctypes.data, ja.ctypes.data,
           val.ctypes.data, None)
        coo = P.psb_d_coo_sparse_mat(L ** 3, L ** 3, ia, ja, val).fix()
        m = P.psb_dspmat_type()
        m.m = m.n = L ** 3
        m.a = P.psb_d_hip_hlg_sparse_mat(64).cp_from_coo(coo)
        prec = P.psb_dprec_type("DIAG").build(m, P.psb_cdall(nl=L ** 3))
        mats.append((m, prec, (L, float(cv))))
    reqs = P.poisson_arrivals(1000.0, 48, nmat, seed=seed, stream=20)
    srv = P.psb_hip_server("BICGSTAB", max_systems=256, arena_bytes=768 << 20, check_every=4,
                           eps=1e-10)
    for m, prec, _ in mats:
        srv.register(m, prec)
    rec = srv.run(reqs, keep_solutions=True, time_limit=100.0)
    assert all(r["status"] == 1 for r in rec)
    # the sample: the first two requests, the last, and the first request on
    # the largest matrix among those requested
    big = max(range(len(reqs)), key=lambda k: (int(Ls[reqs[k][1]]), -k))
    pick = sorted({0, 1, len(reqs) - 1, big})
    for k in pick:
        t, mi, stream = reqs[k]
        alone = srv.run([(0.0, mi, stream)], keep_solutions=True)[0]
        assert alone["status"] == 1
        x_mix = rec[k]["x"].get_vect()
        x_alone = alone["x"].get_vect()
        assert alone["iterations"] == rec[k]["iterations"]
        assert np.array_equal(x_mix.view(np.int64), x_alone.view(np.int64))
        # residual of that request, recomputed by the oracle
        L, cv = mats[mi][2]
        f = orc.fix_coo(orc.gen_convdiff(L, conv=cv / np.sqrt(3.0)))
        csr = orc.csr_from_coo(f)
        bh = orc.uniform_pm1(seed, f.nr, stream=stream)
        rr = np.linalg.norm(bh - orc.csr_spmv(*csr, x_mix)) / np.linalg.norm(bh)
        _say(f"  request {k} (L={L}): {rec[k]['iterations']} its, alone = mix bitwise, "
             f"oracle residual {rr:.3e}")
        assert rr <= 1e-10 * 1.05
        alone["x"].free()
    srv.free()
    for m, _, _ in mats:
        m.free()
