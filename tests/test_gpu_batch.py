"""Batched mixed-precision engine (psb_hip_batch_*) on the GPU.

Checked against the oracle: every system's final FP64 residual recomputed on
the host from the oracle CSR product, iteration counts against the oracle's
FP64 CG, bit-identical results alone and in a mix, with and without the
graph, admission by arena footprint, the iteration limit and a zero rhs.
"""
import numpy as np
import pytest

import psblas3_amd as P

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def _system(orc, kind, L, p, hksz=64):
    if kind == "aniso":
        coo = orc.gen_aniso2d(L, p)
    else:
        coo = orc.gen_convdiff(L, conv=p / np.sqrt(3.0))
    a = P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val,
                          mold=P.psb_d_hip_hlg_sparse_mat(hksz))
    desc = P.psb_cdall(nl=coo.nr)
    prec = P.psb_dprec_type("DIAG").build(a, desc)
    csr = orc.csr_from_coo(orc.fix_coo(coo))
    return a, prec, csr


def _rhs(orc, csr, seed):
    xt = orc.uniform_pm1(seed, len(csr[0]) - 1)
    return orc.csr_spmv(*csr, xt)


def _relres(orc, csr, b, x):
    r = b - orc.csr_spmv(*csr, x)
    return np.linalg.norm(r) / np.linalg.norm(b)


def _run(orc, eng, systems, eps, seeds, itmax=20000):
    jobs, mids = [], []
    for (a, prec, csr), s in zip(systems, seeds):
        mid = eng.register(a, prec)
        mids.append(mid)
        b = _rhs(orc, csr, s)
        jobs.append((mid, P.vect_from(b), P.vect_from(np.zeros(len(b)))))
    out = eng.solve(jobs, eps=eps, itmax=itmax)
    return [j[2].get_vect() for j in jobs], out, [j[1].get_vect() for j in jobs]


CG_CASES = [(12, 1.0), (20, 0.1), (33, 0.01), (47, 0.003), (64, 0.001)]


def test_batch_cg_aniso(gpu, orc):
    _say("case batch CG aniso2d 5 systems eps 1e-8")
    systems = [_system(orc, "aniso", L, e) for L, e in CG_CASES]
    eng = P.psb_hip_batch("CG", max_systems=16, arena_bytes=64 << 20, check_every=8)
    xs, out, bs = _run(orc, eng, systems, 1e-8, seeds=range(5))
    eng.free()
    for (L, e), sys_, x, b, o in zip(CG_CASES, systems, xs, bs, out):
        csr = sys_[2]
        rr = _relres(orc, csr, b, x)
        _, ito, _, _ = orc.krylov("CG", csr, b, eps=1e-8, itmax=20000)
        _say(f"  L={L} eps={e}: {o['iterations']} FP32 its, {o['outer']} replacements, "
             f"relres {rr:.2e} (oracle FP64 CG {ito} its)")
        assert o["status"] == 1
        assert rr <= 1.0e-8 * 1.01
        assert abs(o["relres"] - rr) <= 1e-3 * rr + 1e-15
        # FP32 CG with FP64 residual replacement follows the FP64 recurrence
        # closely: within 30% (+5) above and 10% (+2) below the oracle's count
        assert 0.9 * ito - 2 <= o["iterations"] <= 1.3 * ito + 5


def test_batch_alone_vs_mix_bitwise(gpu, orc):
    _say("case batch alone vs mix (CG)")
    target = _system(orc, "aniso", 40, 0.01)
    others = [_system(orc, "aniso", L, e) for L, e in [(25, 0.5), (57, 0.002), (16, 0.05)]]
    b = _rhs(orc, target[2], 11)
    # alone
    eng = P.psb_hip_batch("CG", max_systems=8, arena_bytes=32 << 20)
    mid = eng.register(target[0], target[1])
    x1 = P.vect_from(np.zeros(len(b)))
    (o1,) = eng.solve([(mid, P.vect_from(b), x1)], eps=1e-8)
    eng.free()
    # in a mix: the others first, the target admitted three steps later
    eng = P.psb_hip_batch("CG", max_systems=8, arena_bytes=32 << 20)
    ids = [eng.register(a, p) for a, p, _ in others]
    tid = eng.register(target[0], target[1])
    outs = {}
    for k, (mid_k, sys_k) in enumerate(zip(ids, others)):
        bk = _rhs(orc, sys_k[2], 20 + k)
        sid = eng.submit(mid_k, P.vect_from(bk), eps=1e-8)
        outs[sid] = ("other", P.vect_from(np.zeros(len(bk))))
    for _ in range(3):
        eng.step()
    x2 = P.vect_from(np.zeros(len(b)))
    sid_t = eng.submit(tid, P.vect_from(b), eps=1e-8)
    outs[sid_t] = ("target", x2)
    res = {}
    while outs:
        eng.step()
        done, _ = eng.poll()
        for d in done:
            kind, xv = outs.pop(d["id"])
            eng.result(d["id"], xv)
            res.setdefault(kind, []).append(d)
    eng.free()
    o2 = res["target"][0]
    _say(f"  alone {o1['iterations']} its / mix {o2['iterations']} its")
    assert o1["iterations"] == o2["iterations"] and o1["outer"] == o2["outer"]
    assert np.array_equal(x1.get_vect(), x2.get_vect())
    assert o1["relres"] == o2["relres"]


def test_batch_graph_equals_eager(gpu, orc):
    _say("case batch graph vs eager launches")
    systems = [_system(orc, "aniso", L, e) for L, e in [(18, 0.3), (29, 0.01)]]
    res = []
    for g in (True, False):
        eng = P.psb_hip_batch("CG", max_systems=4, arena_bytes=16 << 20, use_graph=g)
        xs, out, _ = _run(orc, eng, systems, 1e-8, seeds=[3, 4])
        eng.free()
        res.append((xs, [o["iterations"] for o in out]))
    assert res[0][1] == res[1][1]
    for a, b in zip(res[0][0], res[1][0]):
        assert np.array_equal(a, b)


BI_CASES = [(10, 0.0), (16, 3.0), (23, 10.0)]


@pytest.mark.parametrize("refine", ["reliable", "restart"])
def test_batch_bicgstab_convdiff(gpu, orc, refine):
    _say(f"case batch BiCGSTAB ({refine}) convection-diffusion eps 1e-10")
    systems = [_system(orc, "convdiff", L, c) for L, c in BI_CASES]
    eng = P.psb_hip_batch("BICGSTAB", max_systems=8, arena_bytes=64 << 20, check_every=4,
                          refine=refine)
    xs, out, bs = _run(orc, eng, systems, 1e-10, seeds=[5, 6, 7])
    eng.free()
    for (L, c), sys_, x, b, o in zip(BI_CASES, systems, xs, bs, out):
        rr = _relres(orc, sys_[2], b, x)
        _, ito, _, _ = orc.krylov("BICGSTAB", sys_[2], b, eps=1e-10, itmax=5000)
        _say(f"  L={L} conv={c}: {o['iterations']} FP32 its, {o['outer']} transitions, "
             f"relres {rr:.2e} (oracle FP64 BiCGSTAB {ito} its)")
        assert o["status"] == 1
        assert rr <= 1.0e-10 * 1.01
        # FP32 BiCGSTAB needs more iterations than FP64 (rounding in the
        # inner recurrence, restarts at each FP64 transition): at most 1.5x the
        # oracle's FP64 count (+10), and not fewer than 0.7x
        assert 0.7 * ito <= o["iterations"] <= 1.5 * ito + 10


def test_batch_admission_by_footprint(gpu, orc):
    _say("case batch admission by arena footprint")
    systems = [_system(orc, "aniso", L, 0.1) for L in (30, 30, 30)]
    probe = P.psb_hip_batch("CG", max_systems=4, arena_bytes=1 << 20)
    mid = probe.register(systems[0][0], systems[0][1])
    fp = probe.footprint(mid)
    probe.free()
    n = 900
    expect = 3 * (((8 * n) + 255) // 256 * 256) + 5 * (((4 * n) + 255) // 256 * 256)
    assert fp == expect
    eng = P.psb_hip_batch("CG", max_systems=4, arena_bytes=2 * fp + 100)
    mids = [eng.register(a, p) for a, p, _ in systems]
    bs = [P.vect_from(_rhs(orc, s[2], 30 + k)) for k, s in enumerate(systems)]
    s0 = eng.submit(mids[0], bs[0], eps=1e-8)
    s1 = eng.submit(mids[1], bs[1], eps=1e-8)
    assert s0 is not None and s1 is not None
    assert eng.submit(mids[2], bs[2], eps=1e-8) is None      # no arena room
    live = {s0, s1}
    while live:
        eng.step()
        done, _ = eng.poll()
        for d in done:
            eng.result(d["id"])
            live.discard(d["id"])
    s2 = eng.submit(mids[2], bs[2], eps=1e-8)
    assert s2 is not None
    while True:
        eng.step()
        done, nlive = eng.poll(wait_last=True)
        if done:
            assert done[0]["status"] == 1
            eng.result(done[0]["id"])
            break
    eng.free()


def test_batch_itmax_and_zero_rhs(gpu, orc):
    _say("case batch iteration limit and zero rhs")
    a, prec, csr = _system(orc, "aniso", 40, 0.001)
    eng = P.psb_hip_batch("CG", max_systems=4, arena_bytes=8 << 20)
    mid = eng.register(a, prec)
    n = len(csr[0]) - 1
    x_lim, x_zero = P.vect_from(np.zeros(n)), P.vect_from(np.ones(n))
    out = eng.solve([(mid, P.vect_from(_rhs(orc, csr, 9)), x_lim),
                     (mid, P.vect_from(np.zeros(n)), x_zero)], eps=1e-8, itmax=5)
    eng.free()
    assert out[0]["status"] == 2 and out[0]["relres"] > 1e-8
    assert out[1]["status"] == 1 and out[1]["iterations"] == 0
    assert np.array_equal(x_zero.get_vect(), np.zeros(n))
