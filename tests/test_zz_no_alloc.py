"""psb_hip_krylov_fused_noalloc: the fused CG / BiCGSTAB with every device
buffer carved from the library's static scratch array.  Called through the
C ABI and checked against psb_hip_krylov_fused on the same inputs: the same
kernels in the same order, so iteration counts, errors, histories and
solutions agree bit for bit.  Two calls in a row, the second on a smaller
system, so the second runs over the first one's leftovers in the scratch
array; then the size limit."""
import ctypes as C

import numpy as np
import pytest

import psblas3_amd as P
from psblas3_amd import _native as N

pytestmark = pytest.mark.gpu

FMT = {"HLG": 0, "ELG": 1, "CSRG": 2}


def _say(*a):
    print(*a, flush=True)


def _last_error():
    msg = N.fn("psb_hip_last_error")()
    return msg.decode() if msg else ""


def _system(orc, kind, L, mold, single=False):
    coo = orc.gen_convdiff(L) if kind == "convdiff" else orc.gen_poisson3d(L)
    a = P.matrix_from_coo(coo.nr, coo.nc, coo.ia, coo.ja, coo.val, mold=mold)
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    bh = orc.csr_spmv(irp, ja, val, orc.uniform_pm1(7, coo.nr))
    vm = P.psb_s_hip_vect() if single else P.psb_d_hip_vect()
    dt = np.float32 if single else np.float64
    prec = P.psb_dprec_type("DIAG").build(a, P.psb_cdall(nl=coo.nr))
    return a, prec, bh.astype(dt), vm, dt


def _solve(entry, method, a, prec, bh, vm, dt, eps, itmax):
    b = P.vect_from(bh, mold=type(vm)())
    x = P.vect_from(np.zeros(len(bh), dt), mold=type(vm)())
    it, err = C.c_int(0), C.c_double(0.0)
    hist = np.zeros(itmax)
    rc = N.fn(entry)(method, FMT[a.a.get_fmt()], a.a.dev, prec.dv.dev, b.v.dev, x.v.dev,
                     eps, itmax, 2, 8, C.byref(it), C.byref(err), hist.ctypes.data)
    assert rc == 0, (entry, rc, _last_error())
    x.v.set_dev()
    return it.value, err.value, hist[:it.value].copy(), x.get_vect()


CASES = [
    # (method, matrix, L, mold, single): the second and later cases are smaller
    (1, "convdiff", 40, lambda: P.psb_d_hip_hlg_sparse_mat(64), False),
    (0, "poisson", 24, lambda: P.psb_d_hip_hlg_sparse_mat(64), False),
    (1, "convdiff", 16, lambda: P.psb_d_hip_csrg_sparse_mat(), False),
    (0, "poisson", 12, lambda: P.psb_s_hip_elg_sparse_mat(), True),
]


def test_noalloc_equals_fused(gpu, orc):
    for method, kind, L, mold, single in CASES:
        _say(f"case noalloc method {method} {kind} L={L} single={single}")
        a, prec, bh, vm, dt = _system(orc, kind, L, mold(), single)
        eps = 1e-6 if single else 1e-9
        ref = _solve("psb_hip_krylov_fused", method, a, prec, bh, vm, dt, eps, 1000)
        got = _solve("psb_hip_krylov_fused_noalloc", method, a, prec, bh, vm, dt, eps, 1000)
        assert got[0] == ref[0] and got[0] > 0
        assert got[1] == ref[1] and got[1] <= eps
        assert np.array_equal(got[2], ref[2])
        xb = got[3].view(np.int32 if single else np.int64)
        assert np.array_equal(xb, ref[3].view(np.int32 if single else np.int64))
        # a second call right after, same system: same bits again
        again = _solve("psb_hip_krylov_fused_noalloc", method, a, prec, bh, vm, dt, eps, 1000)
        assert again[0] == ref[0] and np.array_equal(again[3], ref[3])
        a.free()


def test_noalloc_rejects_oversize(gpu, orc):
    _say("case noalloc oversize")
    a, prec, bh, vm, dt = _system(orc, "poisson", 8, P.psb_d_hip_hlg_sparse_mat(64))
    b = P.vect_from(bh)
    x = P.vect_from(np.zeros(len(bh)))
    it, err = C.c_int(0), C.c_double(0.0)
    # 10M iterations of history alone need 80 MB, over the 64 MiB array
    rc = N.fn("psb_hip_krylov_fused_noalloc")(0, 0, a.a.dev, prec.dv.dev, b.v.dev, x.v.dev,
                                              1e-9, 10_000_000, 2, 8, C.byref(it), C.byref(err),
                                              None)
    assert rc == N.INVALID
    assert "scratch" in _last_error() and "maximum" in _last_error()
    # the matrix and vectors are untouched and the entry point still works
    it2, err2, _, xs = _solve("psb_hip_krylov_fused_noalloc", 0, a, prec, bh, vm, dt, 1e-9, 1000)
    ref = _solve("psb_hip_krylov_fused", 0, a, prec, bh, vm, dt, 1e-9, 1000)
    assert it2 == ref[0] and np.array_equal(xs, ref[3])
    with pytest.raises(N.PsbError):
        P.psb_krylov_fused("CG", a, prec, b, x, 1e-9, itmax=10_000_000, noalloc=True)
    a.free()
