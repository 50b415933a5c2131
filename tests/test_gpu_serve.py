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
