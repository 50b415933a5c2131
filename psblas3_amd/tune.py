"""Autotuner for the HLL product: hack size x kernel variant.

For each candidate hack size the matrix is converted (psb_hip_hll_offsets +
device fill) and psb_hip_hll_tune times the four kernel variants (one row or
two adjacent rows per lane, launch-order or XCD-contiguous blocks) with HIP
events; the fastest (hack size, variant) wins and keeps its variant in the
matrix.  The reference fixes the hack size at conversion
(psi_get_hksz, ext/psi_i_ext_util_mod.f90:39-66; the CUDA type uses the warp
size, cuda/impl/psb_d_cuda_cp_hlg_from_coo.F90:60) and has one kernel.
"""
from . import _native as N
from .mat import psb_d_hip_hlg_sparse_mat, psb_s_hip_hlg_sparse_mat
from .vect import psb_d_hip_vect, psb_s_hip_vect


def psb_hip_hll_autotune(coo, hksz_candidates=(32, 64), reps=20, single=False, seed=1):
    """Returns (HLG matrix with the best hack size and variant, report).  coo:
    psb_d_coo_sparse_mat (fixed or not).  x ~ U[-1,1) (seed, stream 0)."""
    cls = psb_s_hip_hlg_sparse_mat if single else psb_d_hip_hlg_sparse_mat
    vcls = psb_s_hip_vect if single else psb_d_hip_vect
    if not coo.sorted:
        coo = coo.fix()
    x = vcls().all(coo.nc)
    y = vcls().all(coo.nr)
    if single:
        import numpy as np
        xd = psb_d_hip_vect().all(coo.nc)
        N.call("psb_hip_mv_fill_uniform", xd.dev, seed, 0)
        xd.set_dev()
        x.bld(xd.get_vect().astype(np.float32))
        xd.free()
    else:
        N.call("psb_hip_mv_fill_uniform", x.dev, seed, 0)
        x.set_dev()
    best, report = None, {"table": {}}
    for hk in hksz_candidates:
        m = cls(hksz=hk).cp_from_coo(coo)
        var, times = m.tune(x, y, reps)
        for v, t in times.items():
            report["table"][f"hksz{hk}_v{v}"] = round(t, 3)
        if best is None or times[var] < best[2]:
            if best is not None:
                best[0].free()
            best = (m, (hk, var), times[var])
        else:
            m.free()
    x.free()
    y.free()
    report.update(hksz=best[1][0], variant=best[1][1], us=round(best[2], 3))
    return best[0], report
