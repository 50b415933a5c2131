"""Triangular solves with a CSR matrix on the device (psb_spsm, the CSRG
type's inner_vect_sv): equal to the oracle's restatement of psb_d_csr_cssv
bit for bit for lower / upper, unit / non-unit, op 'N' / 'T' / 'C', FP64 and
FP32 and every alpha / beta rule; the torture known answers; the errors of
the reference."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import psblas3_amd as P
from psblas3_amd import _native as N
from helpers import tri_csr
from test_gpu_ilu import _band6, _mixed_levels
from test_oracle_cssv import _ussv_cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def _tri_part(orc, coo, lower, unit):
    """tril / triu of the matrix with its diagonal, CSR 1-based; for a unit
    triangle the strict part scaled by the inverse diagonal, D^-1 L (the unit
    factor of the Gauss-Seidel splitting), so that the solution stays bounded
    (the stencil's strict part alone grows by a factor of about 3 per level,
    past the FP32 range at L = 40)."""
    f = orc.fix_coo(coo)
    irp, ja, val = orc.csr_from_coo(f)
    A = sp.csr_matrix((val, ja - 1, irp - 1), shape=(coo.nr, coo.nc))
    if unit:
        A = sp.diags(1.0 / A.diagonal()) @ A
    T = sp.tril(A, k=-1 if unit else 0) if lower else sp.triu(A, k=1 if unit else 0)
    T = sp.csr_matrix(T)
    T.sort_indices()
    return (T.indptr.astype(np.int32) + 1, T.indices.astype(np.int32) + 1,
            T.data.astype(np.float64))


def _tri_mat(irp, ja, val, lower, unit, single=False):
    """psb_spall / psb_spins / psb_spasb of the triangle with its flags set
    before assembly, as the torture tests do."""
    n = len(irp) - 1
    desc = P.psb_cdall(nl=n)
    a = P.psb_spall(desc, len(val))
    a.set_triangle()
    a.set_lower(lower)
    a.set_unit(unit)
    ia = np.repeat(np.arange(1, n + 1, dtype=np.int32), np.diff(irp))
    P.psb_spins(len(val), ia, ja, val, a, desc)
    mold = P.psb_s_hip_csrg_sparse_mat() if single else P.psb_d_hip_csrg_sparse_mat()
    return P.psb_spasb(a, desc, mold=mold), desc


def _vec(x, single=False):
    return P.vect_from(x, mold=P.psb_s_hip_vect() if single else None)


def _bits(a):
    a = np.asarray(a)
    return a.view(np.int32 if a.dtype == np.float32 else np.int64)


CASES = {"convdiff12": lambda o: o.gen_convdiff(12, conv=3.0),
         "poisson40": lambda o: o.gen_poisson3d(40),
         "mixed_levels": _mixed_levels,
         "band6": _band6}


@pytest.fixture(params=[0, 1], ids=["workgroup-runs", "launch-per-level"])
def trsv_variant(request):
    old = N.fn("psb_hip_ilu_set_variant")(request.param)
    yield request.param
    N.fn("psb_hip_ilu_set_variant")(old)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("lower,unit", [(True, False), (False, False), (True, True),
                                        (False, True)], ids=["L", "U", "L-unit", "U-unit"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_spsv_bitexact(gpu, orc, name, lower, unit, trans, trsv_variant):
    _say(f"case csr spsv {name} lower={lower} unit={unit} trans={trans} variant {trsv_variant}")
    coo = CASES[name](orc)
    irp, ja, val = _tri_part(orc, coo, lower, unit)
    a, desc = _tri_mat(irp, ja, val, lower, unit)
    x = orc.uniform_pm1(12, coo.nr)
    y = P.vect_from(np.zeros(coo.nr))
    P.psb_spsm(1.0, a, P.vect_from(x), 0.0, y, desc, trans=trans)
    ref = orc.csr_cssv(irp, ja, val, x, lower=lower, unit=unit, trans=trans)
    assert np.array_equal(_bits(y.get_vect()), _bits(ref))
    lv, sg = a.a.sv_levels()
    _say(f"  levels {lv} workgroup runs {sg}")
    if trsv_variant == 1:
        assert sg == 0
    elif name == "poisson40":
        assert lv == 118 and sg >= 1          # 3 (L - 1) + 1 wavefronts
    a.free()


@pytest.mark.parametrize("name", ["convdiff12", "poisson40"])
@pytest.mark.parametrize("lower,unit,trans", [(True, False, "N"), (False, True, "N"),
                                              (True, False, "T"), (False, False, "C")])
def test_spsv_single_bitexact(gpu, orc, name, lower, unit, trans):
    _say(f"case csr spsv FP32 {name} lower={lower} unit={unit} trans={trans}")
    coo = CASES[name](orc)
    irp, ja, val = _tri_part(orc, coo, lower, unit)
    a, desc = _tri_mat(irp, ja, val, lower, unit, single=True)
    x = orc.uniform_pm1(13, coo.nr).astype(np.float32)
    y = _vec(np.zeros(coo.nr, np.float32), single=True)
    P.psb_spsm(1.0, a, _vec(x, single=True), 0.0, y, desc, trans=trans)
    ref = orc.csr_cssv(irp, ja, val.astype(np.float32), x, lower=lower, unit=unit, trans=trans)
    out = y.get_vect()
    assert out.dtype == np.float32
    assert np.array_equal(_bits(out), _bits(ref))
    a.free()


@pytest.mark.parametrize("alpha,beta", [(0.0, 0.5), (0.0, 0.0), (-1.0, 0.0), (2.5, 0.0),
                                        (-3.0, 0.5), (1.0, 1.0), (1.0, -2.0)])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_spsv_alpha_beta(gpu, orc, alpha, beta, trans):
    _say(f"case csr spsv alpha={alpha} beta={beta} trans={trans}")
    irp, ja, val = tri_csr(orc, 300, True, False, seed=21)
    a, desc = _tri_mat(irp, ja, val, True, False)
    x = orc.uniform_pm1(14, 300)
    y0 = orc.uniform_pm1(15, 300)
    y = P.vect_from(y0.copy())
    P.psb_spsm(alpha, a, P.vect_from(x), beta, y, desc, trans=trans)
    ref = orc.csr_cssv(irp, ja, val, x, lower=True, unit=False, trans=trans, alpha=alpha,
                       beta=beta, y=y0)
    assert np.array_equal(_bits(y.get_vect()), _bits(ref))
    a.free()


@pytest.mark.parametrize("scale", ["L", "R"])
@pytest.mark.parametrize("beta", [0.0, 0.5])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_spsm_diag_scaling(gpu, orc, scale, beta, trans):
    """psb_spsm with a diagonal: 'R' solves with d .* x; 'L' scales the
    solution (y = y .* d when beta = 0, else t .* d then t + beta * y), the
    steps of psb_d_base_vect_cssv, each one rounding like the CPU's."""
    _say(f"case psb_spsm scale={scale} beta={beta} trans={trans}")
    n = 400
    irp, ja, val = tri_csr(orc, n, False, False, seed=61)
    a, desc = _tri_mat(irp, ja, val, False, False)
    x = orc.uniform_pm1(62, n)
    d = orc.uniform_pm1(63, n) + 2.0
    y0 = orc.uniform_pm1(64, n)
    y = P.vect_from(y0.copy())
    P.psb_spsm(2.0, a, P.vect_from(x), beta, y, desc, trans=trans, scale=scale,
               diag=P.vect_from(d))
    if scale == "R":
        ref = orc.csr_cssv(irp, ja, val, d * x, lower=False, trans=trans, alpha=2.0, beta=beta,
                           y=y0)
    else:
        t = orc.csr_cssv(irp, ja, val, x, lower=False, trans=trans, alpha=2.0) * d
        ref = t if beta == 0.0 else orc.axpby(1.0, t, beta, y0.copy())
    assert np.array_equal(_bits(y.get_vect()), _bits(ref))
    a.free()


@pytest.mark.parametrize("case", _ussv_cases(), ids=lambda c: c["name"])
def test_torture_ussv_device(gpu, orc, case):
    single = case["kind"] == "s"
    dt = np.float32 if single else np.float64
    coo = orc.fix_coo(orc.COO(case["m"], case["k"], np.array(case["ia"], np.int32),
                              np.array(case["ja"], np.int32), np.array(case["va"])))
    irp, ja, val = orc.csr_from_coo(coo)
    a, desc = _tri_mat(irp, ja, val, case["lower"], case["unit"], single=single)
    y = _vec(np.array(case["bcy"], dt), single)
    P.psb_spsm(case["alpha"], a, _vec(np.array(case["x"], dt), single), case["beta"], y, desc,
               trans=case["trans"])
    assert y.get_vect().tolist() == case["cy"]
    a.free()


def test_spsv_columns_and_aliasing(gpu, orc):
    """Several columns are solved one after another; x and y may be the same
    vector (the right-hand side is read before the solution is written)."""
    _say("case csr spsv multi-column and aliased x/y")
    coo = orc.gen_convdiff(10, conv=2.0)
    irp, ja, val = _tri_part(orc, coo, True, False)
    a, desc = _tri_mat(irp, ja, val, True, False)
    n = coo.nr
    xs = np.stack([orc.uniform_pm1(16 + q, n) for q in range(3)], axis=1)
    y = P.vect_from(np.zeros((n, 3)))
    P.psb_spsm(1.0, a, P.vect_from(np.asfortranarray(xs)), 0.0, y, desc)
    out = np.asarray(y.get_vect())
    for q in range(3):
        ref = orc.csr_cssv(irp, ja, val, xs[:, q], lower=True, unit=False)
        assert np.array_equal(_bits(out[:, q]), _bits(ref))
    z = P.vect_from(xs[:, 0].copy())
    P.psb_spsm(1.0, a, z, 0.0, z, desc, trans="T")
    ref = orc.csr_cssv(irp, ja, val, xs[:, 0], lower=True, unit=False, trans="T")
    assert np.array_equal(_bits(z.get_vect()), _bits(ref))
    a.free()


def test_spsv_rewrite_and_reuse(gpu, orc):
    """The prepared solve is rebuilt when the values are rewritten
    (psb_hip_csr_write) and when another triangle or op is asked for."""
    _say("case csr spsv rewrite / key change")
    irp, ja, val = tri_csr(orc, 500, True, False, seed=31)
    a, desc = _tri_mat(irp, ja, val, True, False)
    x = orc.uniform_pm1(20, 500)
    y = P.vect_from(np.zeros(500))
    P.psb_spsm(1.0, a, P.vect_from(x), 0.0, y, desc)
    assert np.array_equal(y.get_vect(), orc.csr_cssv(irp, ja, val, x))
    val2 = val * 1.5
    N.call("psb_hip_csr_write", a.a.dev, irp.ctypes.data, ja.ctypes.data,
           np.ascontiguousarray(val2).ctypes.data)
    assert a.a.sv_levels() == (0, 0)
    P.psb_spsm(1.0, a, P.vect_from(x), 0.0, y, desc)
    assert np.array_equal(y.get_vect(), orc.csr_cssv(irp, ja, val2, x))
    P.psb_spsm(1.0, a, P.vect_from(x), 0.0, y, desc, trans="T")
    assert np.array_equal(y.get_vect(), orc.csr_cssv(irp, ja, val2, x, trans="T"))
    P.psb_spsm(1.0, a, P.vect_from(x), 0.0, y, desc)
    assert np.array_equal(y.get_vect(), orc.csr_cssv(irp, ja, val2, x))
    a.free()


def test_spsv_errors(gpu, orc):
    _say("case csr spsv errors")
    irp, ja, val = tri_csr(orc, 50, True, False, seed=41)
    x = P.vect_from(orc.uniform_pm1(22, 50))
    y = P.vect_from(np.zeros(50))
    # not flagged triangular: psb_err_invalid_mat_state_ in the reference
    desc = P.psb_cdall(nl=50)
    a = P.psb_spall(desc, len(val))
    P.psb_spins(len(val), np.repeat(np.arange(1, 51, dtype=np.int32), np.diff(irp)), ja, val,
                a, desc)
    P.psb_spasb(a, desc, mold=P.psb_d_hip_csrg_sparse_mat())
    with pytest.raises(P.PsbError):
        P.psb_spsm(1.0, a, x, 0.0, y, desc)
    a.free()
    # flagged upper, entries below the diagonal
    b, desc = _tri_mat(irp, ja, val, False, True)
    with pytest.raises(P.PsbError):
        P.psb_spsm(1.0, b, x, 0.0, y, desc)
    with pytest.raises(P.PsbError):
        P.psb_spsm(1.0, b, x, 0.0, y, desc, trans="T")
    assert b.a.sv_levels() == (0, 0)
    b.free()
    # a diagonal with a scale other than 'L' / 'R' (info 31), a short
    # diagonal, and the HLG format
    c, desc = _tri_mat(irp, ja, val, True, False)
    with pytest.raises(P.PsbError):
        P.psb_spsm(1.0, c, x, 0.0, y, desc, scale="U", diag=P.vect_from(np.ones(50)))
    with pytest.raises(P.PsbError):
        P.psb_spsm(1.0, c, x, 0.0, y, desc, scale="L", diag=P.vect_from(np.ones(49)))
    c.free()
    # HLG / ELG solve through a device CSR copy of their entries: the CPU's bits
    for mold in (P.psb_d_hip_hlg_sparse_mat(32), P.psb_d_hip_elg_sparse_mat()):
        h = P.matrix_from_coo(50, 50, np.repeat(np.arange(1, 51, dtype=np.int32), np.diff(irp)),
                              ja, val, mold=mold)
        with pytest.raises(P.PsbError):            # not flagged triangular
            P.psb_spsm(1.0, h, x, 0.0, y, desc)
        h.set_triangle()
        for op in ("N", "T"):
            P.psb_spsm(1.0, h, x, 0.0, y, desc, trans=op)
            ref = orc.csr_cssv(irp, ja, val, x.get_vect(), lower=True, unit=False, trans=op)
            assert np.array_equal(y.get_vect(), ref), op
        h.free()
    # a non-unit row with no entries
    irp0 = np.array([1, 2, 2, 3], np.int32)
    ja0 = np.array([1, 3], np.int32)
    d, desc = _tri_mat(irp0, ja0, np.array([2.0, 4.0]), True, False)
    with pytest.raises(P.PsbError):
        P.psb_spsm(1.0, d, P.vect_from(np.ones(3)), 0.0, P.vect_from(np.zeros(3)), desc)
    d.free()
    # the raw entry point rejects mismatched vectors before any work
    e, desc = _tri_mat(irp, ja, val, True, False)
    short = P.vect_from(np.zeros(10))
    rc = N.fn("psb_hip_csr_spsv")(e.a.dev, 1, 0, 0, 1.0, x.v.dev, 0.0, short.v.dev)
    assert rc != 0
    rc = N.fn("psb_hip_csr_spsv")(e.a.dev, 1, 0, 3, 1.0, x.v.dev, 0.0, y.v.dev)
    assert rc != 0
    e.free()
