"""MatrixMarket I/O (psb_mmio_mod): hand-written files, scipy's independent
reader as the second opinion, and exact round trips."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import psblas3_amd as P
from psblas3_amd import _native as N


def _write(path, text):
    path.write_text(text)
    return path


def _dense(coo):
    a = np.zeros((coo.nr, coo.nc))
    np.add.at(a, (coo.ia - 1, coo.ja - 1), coo.val)
    return a


def _sorted_triplets(coo):
    return list(zip(coo.ia.tolist(), coo.ja.tolist(), coo.val.tolist()))


GENERAL = """%%MatrixMarket matrix coordinate real general
% a comment
%
4 5 6
3 2 -1.5
1 1 2.0
4 5 1e-3
1 3 4.25
3 2 0.5
2 4 -7
"""


def test_read_general_real(tmp_path):
    f = _write(tmp_path / "g.mtx", GENERAL)
    coo = P.mm_mat_read(f)
    assert (coo.nr, coo.nc) == (4, 5)
    # rows ascending, columns ascending, the duplicate (3,2) added (psb_d_fix_coo)
    assert _sorted_triplets(coo) == [(1, 1, 2.0), (1, 3, 4.25), (2, 4, -7.0), (3, 2, -1.0),
                                     (4, 5, 1e-3)]
    ref = scipy.io.mmread(str(f)).toarray()
    assert np.array_equal(_dense(coo), ref)


def test_read_pattern_and_symmetric(tmp_path):
    pat = _write(tmp_path / "p.mtx", """%%MatrixMarket matrix coordinate pattern general
3 3 3
1 2
3 1
2 2
""")
    coo = P.mm_mat_read(pat)
    assert _sorted_triplets(coo) == [(1, 2, 1.0), (2, 2, 1.0), (3, 1, 1.0)]
    sym = _write(tmp_path / "s.mtx", """%%MatrixMarket matrix coordinate real symmetric
4 4 5
1 1 4.0
2 1 -1.0
3 2 -1.5
3 3 4.0
4 1 0.25
""")
    coo = P.mm_mat_read(sym)
    assert coo.get_nzeros() == 8
    assert np.array_equal(_dense(coo), scipy.io.mmread(str(sym)).toarray())
    assert np.array_equal(_dense(coo), _dense(coo).T)
    sp_ = _write(tmp_path / "sp.mtx", """%%MatrixMarket matrix coordinate pattern symmetric
3 3 3
2 1
3 3
3 2
""")
    coo = P.mm_mat_read(sp_)
    assert np.array_equal(_dense(coo), scipy.io.mmread(str(sp_)).toarray())


def test_read_unsupported_and_errors(tmp_path):
    ints = _write(tmp_path / "i.mtx", "%%MatrixMarket matrix coordinate integer general\n2 2 1\n1 1 3\n")
    with pytest.raises(N.PsbError) as e:
        P.mm_mat_read(ints)
    assert e.value.info == 904           # the reference reads real and pattern only
    arr = _write(tmp_path / "a.mtx", "%%MatrixMarket matrix array real general\n2 1\n1.0\n2.0\n")
    with pytest.raises(N.PsbError) as e:
        P.mm_mat_read(arr)
    assert e.value.info == 909
    short = _write(tmp_path / "short.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1.0\n2 2 2.0\n")
    with pytest.raises(N.PsbError) as e:
        P.mm_mat_read(short)
    assert e.value.info == 905
    bad = _write(tmp_path / "bad.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(N.PsbError):
        P.mm_mat_read(bad)               # row index out of range
    with pytest.raises(N.PsbError) as e:
        P.mm_mat_read(tmp_path / "missing.mtx")
    assert e.value.info == 901


def test_roundtrip_exact(tmp_path, orc):
    coo = orc.fix_coo(orc.gen_powerlaw(2000, cap=100, seed=6))
    a = P.psb_d_coo_sparse_mat(coo.nr, coo.nc, coo.ia, coo.ja, coo.val)
    f = tmp_path / "rt.mtx"
    P.mm_mat_write(a, "power-law n=2000", f)
    assert f.read_text().startswith("%%MatrixMarket matrix coordinate real general\n% power-law")
    back = P.mm_mat_read(f)
    assert np.array_equal(back.ia, coo.ia) and np.array_equal(back.ja, coo.ja)
    assert np.array_equal(back.val.view(np.int64), coo.val.view(np.int64))
    ref = scipy.io.mmread(str(f)).tocoo()
    assert np.array_equal(_dense(back), ref.toarray())


def test_scipy_written_file(tmp_path):
    rng = np.random.default_rng(5)
    m = sp.random(40, 30, density=0.1, random_state=rng, format="coo")
    f = tmp_path / "scipy.mtx"
    scipy.io.mmwrite(str(f), m, precision=17)
    coo = P.mm_mat_read(f)
    assert np.array_equal(_dense(coo), m.toarray())


def test_array_roundtrip(tmp_path, orc):
    b = orc.uniform_pm1(3, 777)
    f = tmp_path / "b.mtx"
    P.mm_array_write(b, "rhs", f)
    back = P.mm_array_read(f)
    assert np.array_equal(back.view(np.int64), b.view(np.int64))
    assert np.array_equal(scipy.io.mmread(str(f)).ravel(), b)
    sym = _write(tmp_path / "sym.mtx", "%%MatrixMarket matrix array real symmetric\n2 2\n1\n2\n3\n")
    with pytest.raises(N.PsbError):
        P.mm_array_read(sym)


@pytest.mark.gpu
def test_mtx_to_device_spmv(gpu, orc, tmp_path):
    print("[gpu] case MatrixMarket file -> HLG -> SpMV", flush=True)
    coo = orc.fix_coo(orc.gen_convdiff(12))
    f = tmp_path / "cd.mtx"
    P.mm_mat_write(P.psb_d_coo_sparse_mat(coo.nr, coo.nc, coo.ia, coo.ja, coo.val), "cd", f)
    a = P.psb_dspmat_type()
    c = P.mm_mat_read(f)
    a.m, a.n = c.nr, c.nc
    a.a = P.psb_d_hip_hlg_sparse_mat(64).cp_from_coo(c)
    x = orc.uniform_pm1(1, coo.nc)
    y = P.vect_from(np.zeros(coo.nr))
    P.psb_spmm(1.0, a, P.vect_from(x), 0.0, y)
    irp, ja, val = orc.csr_from_coo(coo)
    ref = orc.csr_spmv(irp, ja, val, x)
    absax = orc.csr_spmv(irp, ja, np.abs(val), np.abs(x))
    assert np.all(np.abs(y.get_vect() - ref) <= 4 * 8 * np.finfo(float).eps * absax)
    a.a.free()
