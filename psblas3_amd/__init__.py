"""psblas3_amd: PSBLAS's device SpMV and Krylov vector path, MI355X-native.

The device layer is libpsb_hip.so (hand-written HIP for gfx950, C ABI in
include/psb_hip.h).  This package is the host-side mirror of the PSBLAS
plug-in interface: psb_[sd]_hip_{hlg,elg,csrg}_sparse_mat and
psb_[sd]_hip_vect types behind psb_spmm / psb_geaxpby / psb_gedot /
psb_genrm2 / psb_krylov.
"""
from ._native import PsbError  # noqa: F401
from .env import (DeviceTimer, psb_exit, psb_hip_DeviceName, psb_hip_DeviceSync,  # noqa: F401
                  psb_hip_MemoryPeakBandwidth, psb_hip_MultiProcessors, psb_init, psb_wtime)
from .mat import (psb_d_coo_sparse_mat, psb_d_hip_csrg_sparse_mat,  # noqa: F401
                  psb_d_hip_elg_sparse_mat, psb_d_hip_hlg_sparse_mat, psb_dspmat_type,
                  psb_s_hip_csrg_sparse_mat, psb_s_hip_elg_sparse_mat, psb_s_hip_hlg_sparse_mat)
from .vect import (psb_d_base_vect, psb_d_hip_vect, psb_d_vect_type,  # noqa: F401
                   psb_i_hip_vect, psb_s_base_vect, psb_s_hip_vect)
from .psblas import (matrix_from_coo, psb_cdall, psb_cdasb, psb_cscnv, psb_desc_type,  # noqa: F401
                     psb_geabs, psb_geall, psb_geamax, psb_geasb, psb_geasum, psb_geaxpby,
                     psb_geaxpbyz, psb_gedot, psb_gefree, psb_gemlt, psb_genrm2, psb_gescal,
                     psb_spall, psb_spasb, psb_spfree, psb_spins, psb_spmm, psb_spsm, psb_upd_xyz,
                     psb_xyzw, vect_from)
from .prec import psb_dprec_type, psb_precinit  # noqa: F401
from .krylov import psb_krylov, psb_krylov_fused, psb_krylov_mixed, psb_spnrmi  # noqa: F401
from .batch import psb_hip_batch  # noqa: F401
from .serve import (latency_summary, poisson_arrivals, psb_hip_server,  # noqa: F401
                    serve_one_at_a_time)
from .tune import psb_hip_hll_autotune  # noqa: F401
from .mmio import mm_array_read, mm_array_write, mm_mat_read, mm_mat_write  # noqa: F401
