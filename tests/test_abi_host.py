"""CPU-only checks of libpsb_hip.so: it loads, exports every entry point
include/psb_hip.h declares, and its host-side assembly routines (COO fix,
hack offsets, generators) agree bit for bit with the oracle.  No device
work is done here."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from psblas3_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "psb_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psb_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = N.load()
    names = _header_functions()
    assert len(names) > 80
    out = subprocess.check_output(["nm", "-D", "--defined-only", N.LIB_PATH]).decode()
    exported = set(re.findall(r" T (psb_hip_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)


def test_binding_table_matches_header():
    assert sorted(N.exported_names()) == _header_functions()


def _fix_native(nr, nc, ia, ja, val, dupl=1):
    ia, ja, val = ia.copy(), ja.copy(), val.copy()
    nzo = C.c_int64()
    N.call("psb_hip_fix_coo_d", nr, nc, len(val), ia.ctypes.data, ja.ctypes.data,
           val.ctypes.data, dupl, C.byref(nzo))
    k = nzo.value
    return ia[:k], ja[:k], val[:k]


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("dupl", [0, 1])
def test_fix_coo_matches_oracle(orc, seed, dupl):
    rng = np.random.default_rng(seed)
    nr, nc, nz = 300, 80, 5000
    ia = rng.integers(1, nr + 1, nz).astype(np.int32)
    ja = rng.integers(1, nc + 1, nz).astype(np.int32)
    val = rng.standard_normal(nz)
    # a few very long rows exercise the stable_sort branch
    ia[:400] = 7
    a = _fix_native(nr, nc, ia, ja, val, dupl)
    o = orc.fix_coo(orc.COO(nr, nc, ia, ja, val), dupl=dupl)
    assert np.array_equal(a[0], o.ia) and np.array_equal(a[1], o.ja)
    assert np.array_equal(a[2].view(np.int64), o.val.view(np.int64))


def test_fix_coo_rejects_bad_rows():
    ia = np.array([1, 9], np.int32)
    ja = np.array([1, 1], np.int32)
    with pytest.raises(N.PsbError):
        _fix_native(4, 4, ia, ja, np.ones(2))


def test_fix_coo_dupl_err():
    ia = np.array([2, 2, 1], np.int32)
    ja = np.array([1, 1, 1], np.int32)
    with pytest.raises(N.PsbError):
        _fix_native(4, 4, ia, ja, np.ones(3), dupl=2)


@pytest.mark.parametrize("hksz", [1, 5, 32, 64])
def test_hll_offsets_match_oracle(orc, hksz):
    o = orc.fix_coo(orc.gen_powerlaw(1000, cap=100, seed=2))
    nr = o.nr
    noffs = (nr + hksz - 1) // hksz
    irn = np.zeros(nr, np.int32)
    idisp = np.zeros(nr + 1, np.int32)
    hk = np.zeros(noffs + 1, np.int32)
    isz = C.c_int64()
    N.call("psb_hip_hll_offsets", nr, o.nz, o.ia.ctypes.data, hksz, irn.ctypes.data,
           idisp.ctypes.data, hk.ctypes.data, C.byref(isz))
    h = orc.hll_from_coo(o, hksz)
    assert np.array_equal(hk, h.hkoffs) and np.array_equal(irn, h.irn)
    assert isz.value == h.hkoffs[-1]
    assert np.array_equal(idisp, np.concatenate([[0], np.cumsum(h.irn)]))


def _gen_native_pde3d(L, a, b, c, hscale, gflag):
    fn = N.fn("psb_hip_gen_pde3d")
    av = np.array(a, np.float64)
    bv = np.array(b, np.float64)
    nz = fn(L, av.ctypes.data, bv.ctypes.data, c, hscale, gflag, None, None, None, None)
    ia = np.empty(nz, np.int32); ja = np.empty(nz, np.int32); val = np.empty(nz)
    rhs = np.zeros(L ** 3)
    fn(L, av.ctypes.data, bv.ctypes.data, c, hscale, gflag, ia.ctypes.data, ja.ctypes.data,
       val.ctypes.data, rhs.ctypes.data)
    return ia, ja, val, rhs


@pytest.mark.parametrize("L", [1, 2, 5, 17])
def test_pde3d_generator_matches_oracle(orc, L):
    for a, b, c, hs, g in (((1, 1, 1), (0.577, 0.577, 0.577), 0.0, 1, 0),
                           ((1 / 80,) * 3, (0, 0, 0), 0.0, 1, 1),
                           ((1, 1, 1), (0, 0, 0), 0.0, 0, 0)):
        ia, ja, val, rhs = _gen_native_pde3d(L, a, b, c, hs, g)
        o, orhs = orc.gen_pde3d(L, a, b, c, bool(hs), bool(g))
        assert np.array_equal(ia, o.ia) and np.array_equal(ja, o.ja)
        assert np.array_equal(val.view(np.int64), o.val.view(np.int64))
        assert np.array_equal(rhs.view(np.int64), orhs.view(np.int64))


def test_aniso_and_powerlaw_generators_match_oracle(orc):
    fn = N.fn("psb_hip_gen_aniso2d")
    L, eps = 13, 0.003
    nz = fn(L, eps, None, None, None)
    ia = np.empty(nz, np.int32); ja = np.empty(nz, np.int32); val = np.empty(nz)
    fn(L, eps, ia.ctypes.data, ja.ctypes.data, val.ctypes.data)
    o = orc.gen_aniso2d(L, eps)
    assert np.array_equal(ia, o.ia) and np.array_equal(ja, o.ja) and np.array_equal(val, o.val)
    fp = N.fn("psb_hip_gen_powerlaw")
    n = 70000   # above the threaded-generation threshold
    nz = fp(n, 4095, 1, None, None, None)
    ia = np.empty(nz, np.int32); ja = np.empty(nz, np.int32); val = np.empty(nz)
    fp(n, 4095, 1, ia.ctypes.data, ja.ctypes.data, val.ctypes.data)
    o = orc.gen_powerlaw(n, 4095, 1)
    assert np.array_equal(ia, o.ia) and np.array_equal(ja, o.ja) and np.array_equal(val, o.val)


def test_uniform_fill_matches_oracle(orc):
    x = np.empty(100000)
    N.fn("psb_hip_fill_uniform_host")(5, 0, len(x), x.ctypes.data)
    assert np.array_equal(x, orc.uniform_pm1(5, len(x)))
    assert x.min() >= -1.0 and x.max() < 1.0 and abs(x.mean()) < 0.01


@pytest.mark.parametrize("window,hksz", [(0, 64), (64, 64), (100, 32), (1000, 64), (-1, 7)])
def test_hll_sort_rows_host(window, hksz):
    """psb_hip_hll_sort_rows against a numpy restatement: inside each window
    of ceil(window/hksz)*hksz rows (all rows for window <= 0), positions take
    rows by descending count, ties by ascending row; perm is a bijection of
    1..nr; hack offsets are hksz times each hack's widest count."""
    rng = np.random.default_rng(3)
    nr = 1000
    irn = np.minimum(rng.zipf(2.0, nr), 300).astype(np.int32)
    irn[::17] = 0
    noffs = (nr + hksz - 1) // hksz
    perm = np.zeros(nr, np.int32)
    ipos = np.zeros(nr, np.int32)
    hk = np.zeros(noffs + 1, np.int32)
    isz = C.c_int64()
    N.call("psb_hip_hll_sort_rows", nr, hksz, window, irn.ctypes.data, perm.ctypes.data,
           ipos.ctypes.data, hk.ctypes.data, C.byref(isz))
    win = nr if window <= 0 else -(-window // hksz) * hksz
    exp = []
    for w0 in range(0, nr, win):
        idx = np.arange(w0, min(nr, w0 + win))
        exp.extend(idx[np.lexsort((idx, -irn[idx]))])
    exp = np.array(exp) + 1
    assert np.array_equal(perm, exp)
    assert np.array_equal(np.sort(perm), np.arange(1, nr + 1))
    assert np.array_equal(ipos, irn[perm - 1])
    widths = [ipos[h * hksz:(h + 1) * hksz].max() for h in range(noffs)]
    assert np.array_equal(np.diff(hk), hksz * np.array(widths))
    assert isz.value == hk[-1]
    # sorting never pads more than rows in order
    assert isz.value <= sum(hksz * irn[h * hksz:(h + 1) * hksz].max() for h in range(noffs))


def test_noalloc_solver_needs_init():
    """The allocation-free fused solver does not initialise the library: in a
    fresh process without psb_hip_init it returns PSB_HIP_INVALID before any
    device work."""
    code = (
        "import ctypes as C\n"
        "from psblas3_amd import _native as N\n"
        "it, err = C.c_int(0), C.c_double(0.0)\n"
        "rc = N.fn('psb_hip_krylov_fused_noalloc')(0, 0, None, None, None, None, 1e-9, 10, 2, 8,"
        " C.byref(it), C.byref(err), None)\n"
        "print(rc, N.fn('psb_hip_last_error')().decode())\n")
    out = subprocess.check_output([sys.executable, "-c", code], cwd=ROOT,
                                  env=dict(os.environ, HIP_VISIBLE_DEVICES="")).decode()
    rc, msg = out.split(" ", 1)
    assert int(rc) == N.INVALID and "psb_hip_init" in msg
