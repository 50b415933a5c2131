"""Pins the oracle's COO fix and CSR/ELL/HLL conversions.

Checks: a case derived by hand from the reference rules
(ext/impl/psi_d_convert_hll_from_coo.f90, psi_d_xtr_ell_from_coo.f90),
an independent numpy restatement of the sort/merge, closed forms of the
7-point stencil, and a dense rebuild of every format.
"""
import numpy as np
import pytest


def _hand_case(orc):
    # rows (1-based): r1 (1,1)=1,(1,3)=2; r2 empty; r3 (3,2)=3;
    # r4 (4,4)=4,(4,1)=5,(4,5)=6 (unsorted); r5 (5,5)=7
    ia = np.array([1, 1, 3, 4, 4, 4, 5], np.int32)
    ja = np.array([1, 3, 2, 4, 1, 5, 5], np.int32)
    val = np.array([1, 2, 3, 4, 5, 6, 7], np.float64)
    return orc.fix_coo(orc.COO(5, 5, ia, ja, val))


def test_hll_hand_case(orc):
    h = orc.hll_from_coo(_hand_case(orc), hksz=2)
    assert h.hkoffs.tolist() == [0, 4, 10, 12]
    assert h.irn.tolist() == [2, 0, 1, 3, 1]
    # empty row 2 padded with the hack's first column (1); row 3 with its own
    # last column (2); the unused row of the last hack with ja = nr = 5
    assert h.ja.tolist() == [1, 1, 3, 1, 2, 1, 2, 4, 2, 5, 5, 5]
    assert h.val.tolist() == [1, 0, 2, 0, 3, 5, 0, 4, 0, 6, 7, 0]
    assert h.idiag.tolist() == [1, 0, 0, 2, 1]


def test_ell_hand_case(orc):
    e = orc.ell_from_coo(_hand_case(orc))
    assert (e.ldv, e.nzm) == (5, 3)
    # an empty row is padded with the first column of the whole matrix
    assert e.ja.tolist() == [1, 1, 2, 1, 5, 3, 1, 2, 4, 5, 3, 1, 2, 5, 5]
    assert e.val.tolist() == [1, 0, 3, 5, 7, 2, 0, 0, 4, 0, 0, 0, 0, 6, 0]
    assert e.idiag.tolist() == [1, 0, 0, 2, 1]


def _numpy_fix(nr, ia, ja, val, dupl=1):
    """Independent restatement: stable lexsort then a left-to-right merge."""
    order = np.lexsort((np.arange(len(ia)), ja, ia))
    out_i, out_j, out_v = [], [], []
    for k in order:
        if out_i and out_i[-1] == ia[k] and out_j[-1] == ja[k]:
            out_v[-1] = out_v[-1] + val[k] if dupl == 1 else val[k]
        else:
            out_i.append(ia[k]); out_j.append(ja[k]); out_v.append(val[k])
    return np.array(out_i), np.array(out_j), np.array(out_v)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("dupl", [0, 1])
def test_fix_coo_vs_numpy(orc, seed, dupl):
    rng = np.random.default_rng(seed)
    nr, nz = 50, 600
    ia = rng.integers(1, nr + 1, nz).astype(np.int32)
    ja = rng.integers(1, 20, nz).astype(np.int32)       # many duplicates
    val = rng.standard_normal(nz)
    f = orc.fix_coo(orc.COO(nr, 20, ia, ja, val), dupl=dupl)
    ei, ej, ev = _numpy_fix(nr, ia, ja, val, dupl)
    assert np.array_equal(f.ia, ei) and np.array_equal(f.ja, ej)
    assert np.array_equal(f.val.view(np.int64), ev.view(np.int64))  # bit-exact


def test_fix_coo_dupl_err(orc):
    c = orc.COO(2, 2, np.array([1, 1], np.int32), np.array([2, 2], np.int32), np.ones(2))
    with pytest.raises(ValueError):
        orc.fix_coo(c, dupl=2)


def _dense_from_slots(nr, nc, rows, ja, val):
    d = np.zeros((nr, nc))
    np.add.at(d, (rows, ja - 1), val)
    return d


def _dense_from_coo(c):
    d = np.zeros((c.nr, c.nc))
    np.add.at(d, (c.ia - 1, c.ja - 1), c.val)
    return d


def _hll_rows(h):
    rows = np.empty(len(h.ja), np.int64)
    for hk in range(len(h.hkoffs) - 1):
        a, b = h.hkoffs[hk], h.hkoffs[hk + 1]
        lane = np.arange(b - a) % h.hksz
        rows[a:b] = hk * h.hksz + lane
    return rows


@pytest.mark.parametrize("hksz", [2, 7, 32, 64])
def test_hll_dense_rebuild(orc, hksz):
    c = orc.fix_coo(orc.gen_powerlaw(300, cap=40, seed=5))
    h = orc.hll_from_coo(c, hksz)
    rows = _hll_rows(h)
    keep = rows < c.nr
    assert np.all(h.val[~keep] == 0) and np.all(h.ja[~keep] == c.nr)
    assert np.array_equal(_dense_from_slots(c.nr, c.nc, rows[keep], h.ja[keep], h.val[keep]),
                          _dense_from_coo(c))


def test_ell_dense_rebuild(orc):
    c = orc.fix_coo(orc.gen_powerlaw(300, cap=40, seed=6))
    e = orc.ell_from_coo(c)
    rows = np.tile(np.arange(e.ldv), e.nzm)
    keep = rows < c.nr
    assert np.array_equal(_dense_from_slots(c.nr, c.nc, rows[keep], e.ja[keep], e.val[keep]),
                          _dense_from_coo(c))


@pytest.mark.parametrize("L,hksz", [(7, 32), (40, 32), (40, 64)])
def test_stencil_closed_forms(orc, L, hksz):
    c = orc.fix_coo(orc.gen_convdiff(L))
    n = L ** 3
    assert c.nz == 7 * n - 6 * L * L
    ix, iy, iz = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    nb = ((ix == 0) | (ix == L - 1)).astype(int) + ((iy == 0) | (iy == L - 1)) + \
        ((iz == 0) | (iz == L - 1))
    # row index is ix fastest: transpose to (iz, iy, ix)
    irn_expect = (7 - nb).transpose(2, 1, 0).ravel()
    h = orc.hll_from_coo(c, hksz)
    assert np.array_equal(h.irn, irn_expect)
    for hk in range(len(h.hkoffs) - 1):
        r0, r1 = hk * hksz, min(n, (hk + 1) * hksz)
        assert h.hkoffs[hk + 1] - h.hkoffs[hk] == hksz * irn_expect[r0:r1].max()
    # CSR column order per interior row: z-1, y-1, x-1, diag, x+1, y+1, z+1
    irp, ja, _ = orc.csr_from_coo(c)
    i = (L // 2) * L * L + (L // 2) * L + L // 2
    cols = ja[irp[i] - 1:irp[i + 1] - 1] - 1
    assert (cols - i).tolist() == [-L * L, -L, -1, 0, 1, L, L * L]
    assert np.all(h.idiag > 0)


def test_convdiff_coefficients(orc):
    """Diagonal 6/h^2, neighbours -1/h^2 -/+ (1/sqrt3)/(2h), h = 1/(L+1)."""
    L = 5
    c = orc.fix_coo(orc.gen_convdiff(L))
    h = 1.0 / (L + 1)
    b = 1.0 / np.sqrt(3.0)
    d = _dense_from_coo(c)
    i = 2 * L * L + 2 * L + 2
    assert d[i, i] == pytest.approx(6 / h ** 2, rel=1e-15)
    for off in (1, L, L * L):
        assert d[i, i - off] == pytest.approx(-1 / h ** 2 - b / (2 * h), rel=1e-15)
        assert d[i, i + off] == pytest.approx(-1 / h ** 2 + b / (2 * h), rel=1e-15)


def test_powerlaw_row_lengths(orc):
    n = 200000
    lens = np.array([orc.lib().orc_powerlaw_rowlen(n, 4095, 1, i) for i in range(1, 20001)])
    assert lens.min() >= 2
    assert 13.0 < lens.mean() < 19.5          # mean 16 in expectation (heavy tail)
    c = orc.gen_powerlaw(2000, cap=4095, seed=1)
    f = orc.fix_coo(c)
    assert f.nz == c.nz                        # no duplicate columns generated
    assert np.all(orc.hll_from_coo(f, 32).idiag > 0)   # diagonal present in every row
