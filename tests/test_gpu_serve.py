"""Serving loop (continuous batching) on the GPU.

Each request's solution must be the same alone and in the mix: bit-identical
(the engine's per-system arithmetic does not depend on the other systems),
which is stricter than the solve tolerance.  Residuals are recomputed on the
host from the oracle CSR product; admission is limited by the arena size.
"""
import numpy as np
import pytest

import psblas3_amd as P

pytestmark = pytest.mark.gpu

MATS = [(10, 0.0), (14, 2.0), (18, 5.0), (12, 10.0)]


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def _mats(orc):
    out = []
    for L, c in MATS:
        coo = orc.gen_convdiff(L, conv=c / np.sqrt(3.0))
        a = P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val,
                              mold=P.psb_d_hip_hlg_sparse_mat(64))
        desc = P.psb_cdall(nl=coo.nr)
        prec = P.psb_dprec_type("DIAG").build(a, desc)
        out.append((a, prec, desc, orc.csr_from_coo(orc.fix_coo(coo))))
    return out


def _rhs(orc, n, stream):
    return orc.uniform_pm1(4, n, stream)


def test_serve_alone_vs_mix(gpu, orc):
    _say("case serve: 8 overlapping requests vs each alone (BiCGSTAB mixed, 1e-10)")
    mats = _mats(orc)
    reqs = [(0.0005 * k, k % 4, 500 + k) for k in range(8)]
    srv = P.psb_hip_server("BICGSTAB", max_systems=16, arena_bytes=64 << 20, eps=1e-10)
    ids = [srv.register(a, p) for a, p, _, _ in mats]
    assert ids == [0, 1, 2, 3]
    mix = srv.run(reqs, keep_solutions=True)
    srv.free()
    overlap = sum(1 for i in range(8) for j in range(8)
                  if i < j and mix[j]["admitted"] < mix[i]["done"] and mix[i]["admitted"] < mix[j]["done"])
    _say(f"  {overlap} overlapping request pairs in the mix")
    assert overlap > 0
    for k, (t, mi, stream) in enumerate(reqs):
        srv = P.psb_hip_server("BICGSTAB", max_systems=4, arena_bytes=16 << 20, eps=1e-10)
        for a, p, _, _ in mats:
            srv.register(a, p)
        (alone,) = srv.run([(0.0, mi, stream)], keep_solutions=True)
        srv.free()
        xm, xa = mix[k]["x"].get_vect(), alone["x"].get_vect()
        csr = mats[mi][3]
        b = _rhs(orc, len(xm), stream)
        rr = np.linalg.norm(b - orc.csr_spmv(*csr, xm)) / np.linalg.norm(b)
        _say(f"  request {k} (matrix {mi}): {mix[k]['iterations']} its, relres {rr:.2e}")
        assert mix[k]["status"] == 1 and rr <= 1.01e-10
        assert mix[k]["iterations"] == alone["iterations"]
        assert np.array_equal(xm, xa)


def test_serve_admission_by_memory(gpu, orc):
    _say("case serve: arena room for two solves, six requests at once")
    mats = _mats(orc)
    srv = P.psb_hip_server("BICGSTAB", max_systems=16, arena_bytes=1 << 20, eps=1e-10)
    for a, p, _, _ in mats:
        srv.register(a, p)
    big = srv.eng.footprint(2)               # L = 18: 5832 rows
    srv.free()
    srv = P.psb_hip_server("BICGSTAB", max_systems=16, arena_bytes=2 * big + 512, eps=1e-10)
    for a, p, _, _ in mats:
        srv.register(a, p)
    reqs = [(0.0, 2, 700 + k) for k in range(6)]
    rec = srv.run(reqs)
    srv.free()
    assert all(r["status"] == 1 for r in rec)
    events = sorted([(r["admitted"], 1) for r in rec] + [(r["done"], -1) for r in rec])
    live = peak = 0
    for _, d in events:
        live += d
        peak = max(peak, live)
    _say(f"  peak concurrent solves {peak}")
    assert peak <= 2


def test_serve_fp64_baseline(gpu, orc):
    _say("case serve: FP64 one-at-a-time baseline on the same stream")
    mats = _mats(orc)
    reqs = [(0.0002 * k, k % 4, 900 + k) for k in range(4)]
    rec = P.serve_one_at_a_time([(a, p, d) for a, p, d, _ in mats], reqs, eps=1e-10)
    for r in rec:
        assert r["status"] == 1 and r["relres"] <= 1e-10
    s = P.latency_summary(rec)
    assert s["finished"] == 4 and s["p99_ms"] >= s["p50_ms"] > 0


def test_restart_refinement_survives_nonfinite_breakdown(gpu):
    """A request whose FP32 inner BiCGSTAB overflows near a breakdown: matrix
    25 of the 64-matrix serving set (3-D convection-diffusion, L = 159,
    convection magnitude from host uniform stream 1 of seed 4, about 6.34)
    with the right-hand side of device stream 9100026.  With restarted
    refinement its third inner cycle went non-finite and, when only exact
    zeros counted as breakdowns, ran to the iteration limit with a NaN
    residual.  It must converge to 1e-10; its FP64 residual is checked on
    the device by the engine and recomputed here through the product."""
    from psblas3_amd import _native as N
    _say("case serve: restarted refinement through a non-finite FP32 breakdown (L=159)")
    nmat, seed, k = 64, 4, 25
    u = np.empty(nmat)
    N.fn("psb_hip_fill_uniform_host")(seed, 0, nmat, u.ctypes.data)
    w = np.empty(nmat)
    N.fn("psb_hip_fill_uniform_host")(seed, 1, nmat, w.ctypes.data)
    L = int(min(24 + int((u[k] + 1.0) / 2.0 * 137), 160))
    conv = (w[k] + 1.0) / 2.0 * 10.0
    assert L == 159 and abs(conv - 6.341) < 1e-3
    fn = N.fn("psb_hip_gen_pde3d")
    a3, b3 = np.ones(3), np.full(3, conv / np.sqrt(3.0))
    nz = fn(L, a3.ctypes.data, b3.ctypes.data, 0.0, 1, 0, None, None, None, None)
    ia = np.empty(nz, np.int32); ja = np.empty(nz, np.int32); val = np.empty(nz)
    fn(L, a3.ctypes.data, b3.ctypes.data, 0.0, 1, 0, ia.ctypes.data, ja.ctypes.data,
       val.ctypes.data, None)
    coo = P.psb_d_coo_sparse_mat(L ** 3, L ** 3, ia, ja, val)
    coo.fix()
    m = P.psb_dspmat_type()
    m.m = m.n = L ** 3
    m.a = P.psb_d_hip_hlg_sparse_mat(64).cp_from_coo(coo)
    desc = P.psb_cdall(nl=L ** 3)
    prec = P.psb_dprec_type("DIAG").build(m, desc)
    srv = P.psb_hip_server("BICGSTAB", max_systems=4, arena_bytes=768 << 20, check_every=4,
                           eps=1e-10, refine="restart")
    j = srv.register(m, prec)
    rec = srv.run([(0.0, j, 9100026)], keep_solutions=True, time_limit=120.0)[0]
    _say(f"  status {rec['status']} iterations {rec['iterations']} outer {rec['outer']} "
         f"relres {rec['relres']:.3e}")
    assert rec["status"] == 1 and rec["relres"] <= 1e-10
    # the residual again, through the device product
    b = P.psb_d_hip_vect().all(m.m)
    N.call("psb_hip_mv_fill_uniform", b.dev, seed, 9100026)
    b.set_dev()
    bh = b.get_vect()
    ax = P.vect_from(np.zeros(m.m))
    P.psb_spmm(1.0, m, P.psb_d_vect_type(rec["x"]), 0.0, ax)
    r = bh - ax.get_vect()
    assert np.isfinite(r).all()
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(bh)
    srv.free()
    m.free()
