"""DIAG preconditioner build on the device (psb_d_diag_precbld,
prec/impl/psb_d_diagprec_impl.f90:229-293) for every matrix plug-in, against
the oracle's restatement: d(i) = 1/a_ii, a stored zero diagonal and a row
with no stored diagonal both giving 1.  Bit for bit (one IEEE division)."""
import numpy as np
import pytest

import psblas3_amd as P

pytestmark = pytest.mark.gpu

MOLDS = {"HLG": lambda: P.psb_d_hip_hlg_sparse_mat(), "HLG32": lambda: P.psb_d_hip_hlg_sparse_mat(32),
         "ELG": lambda: P.psb_d_hip_elg_sparse_mat(), "CSRG": lambda: P.psb_d_hip_csrg_sparse_mat()}
SMOLDS = {"HLG": lambda: P.psb_s_hip_hlg_sparse_mat(), "ELG": lambda: P.psb_s_hip_elg_sparse_mat(),
          "CSRG": lambda: P.psb_s_hip_csrg_sparse_mat()}


def _matrix(orc):
    """convection-diffusion L=7 with row 3's diagonal stored as 0 and row 9's
    diagonal entry removed (0-based rows)."""
    coo = orc.gen_convdiff(7, conv=2.0)
    ia, ja, val = coo.ia.copy(), coo.ja.copy(), coo.val.copy()
    val[(ia == 4) & (ja == 4)] = 0.0
    keep = ~((ia == 10) & (ja == 10))
    return coo.nr, ia[keep], ja[keep], val[keep]


@pytest.mark.parametrize("fmt", list(MOLDS))
def test_diag_build_fp64(gpu, orc, fmt):
    print("case DIAG build fp64", fmt, flush=True)
    n, ia, ja, val = _matrix(orc)
    a = P.matrix_from_coo(n, n, ia, ja, val, mold=MOLDS[fmt]())
    prec = P.psb_dprec_type("DIAG").build(a, P.psb_cdall(nl=n))
    d = prec.dv.get_vect()[:n]
    csr = orc.csr_from_coo(orc.fix_coo(orc.COO(n, n, ia, ja, val)))
    de = orc.diag_build(*csr)
    assert d[3] == 1.0 and d[9] == 1.0
    assert np.array_equal(d.view(np.uint64), de.view(np.uint64))
    a.free()


@pytest.mark.parametrize("fmt", list(SMOLDS))
def test_diag_build_fp32(gpu, orc, fmt):
    print("case DIAG build fp32", fmt, flush=True)
    n, ia, ja, val = _matrix(orc)
    a = P.matrix_from_coo(n, n, ia, ja, val, mold=SMOLDS[fmt]())
    prec = P.psb_dprec_type("DIAG").build(a, P.psb_cdall(nl=n))
    d = prec.dv.get_vect()[:n]
    assert d.dtype == np.float32
    # closed form in single precision: 1/a_ii, zero or absent -> 1
    dense = np.zeros(n, np.float32)
    for i, j, v in zip(ia, ja, val):
        if i == j:
            dense[i - 1] = np.float32(v)
    de = np.where(dense == 0, np.float32(1), np.float32(1) / np.where(dense == 0, 1, dense))
    assert np.array_equal(d.view(np.uint32), de.astype(np.float32).view(np.uint32))
    a.free()
