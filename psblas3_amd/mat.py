"""Device sparse matrices: psb_[ds]_hip_{hlg,elg,csrg}_sparse_mat, the plug-ins
that take the place of psb_[ds]_cuda_{hlg,elg,csrg}_sparse_mat
(cuda/psb_d_cuda_hlg_mat_mod.F90:44-84, psb_d_cuda_elg_mat_mod.F90:44-60,
psb_d_cuda_csrg_mat_mod.F90:43-81), and the psb_dspmat_type wrapper whose
``a`` component holds one of them (base/modules/serial/psb_d_mat_mod.F90).

Assembly follows psb_spall / psb_spins / psb_spasb(mold=) (base/tools/
psb_dspasb.f90:47,144): triplets are collected host side, fixed by
psb_hip_fix_coo_d (sort + duplicate merge, psb_d_fix_coo) and handed to the
plug-in's cp_from_coo, which for HLG computes the hack offsets on the host
(psi_compute_hckoff_from_coo, cuda/impl/psb_d_cuda_cp_hlg_from_coo.F90:117-188)
and fills the device arrays with a kernel, as the reference GPU type does.
"""
import ctypes as C

import numpy as np

from . import _native as N

DUPL_OVWRT, DUPL_ADD, DUPL_ERR = 0, 1, 2
IS_HOST, IS_SYNC, IS_DEV = -1, 0, 1
_DUPL = {"ovwrt": DUPL_OVWRT, "add": DUPL_ADD, "err": DUPL_ERR}


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


class psb_d_coo_sparse_mat:
    """Host COO in build state (1-based triplets), psb_d_coo_sparse_mat."""

    def __init__(self, nr, nc, ia=None, ja=None, val=None, dtype=np.float64):
        self.nr, self.nc = int(nr), int(nc)
        self.ia = np.zeros(0, np.int32) if ia is None else np.ascontiguousarray(ia, np.int32)
        self.ja = np.zeros(0, np.int32) if ja is None else np.ascontiguousarray(ja, np.int32)
        self.val = np.zeros(0) if val is None else np.ascontiguousarray(val, np.float64)
        self.sorted = False

    def get_nzeros(self):
        return int(self.val.size)

    def fix(self, dupl=DUPL_ADD):
        """psb_d_fix_coo (row major): in place; raises on dupl=err duplicates."""
        ia = self.ia.copy()
        ja = self.ja.copy()
        val = self.val.copy()
        nzout = C.c_int64(0)
        N.call("psb_hip_fix_coo_d", self.nr, self.nc, int(val.size), _ptr(ia), _ptr(ja),
               _ptr(val), int(dupl), C.byref(nzout))
        k = nzout.value
        self.ia, self.ja, self.val = ia[:k].copy(), ja[:k].copy(), val[:k].copy()
        self.sorted = True
        return self


class _hip_base_mat:
    fmt = None
    _etype = None
    _dtype = None

    def __init__(self):
        self.dev = None
        self.nr = self.nc = self.nnz = 0
        # dual-memory state (cuda/psb_d_cuda_hlg_mat_mod.F90:40-42) and the
        # host copy of the format's arrays (None until built on the host or
        # read back by sync)
        self._state = IS_DEV
        self._hc = None
        # triangle flags of psb_d_base_sparse_mat (set_triangle, set_lower,
        # set_unit: base/modules/serial/psb_base_mat_mod.f90); a triangular
        # solve reads them as the reference's inner_vect_sv does
        self._triangle = False
        self._upper = False
        self._unit = False

    def set_triangle(self, val=True):
        self._triangle = bool(val)

    def set_lower(self, val=True):
        self._upper = not val

    def set_upper(self, val=True):
        self._upper = bool(val)

    def set_unit(self, val=True):
        self._unit = bool(val)

    def is_triangle(self):
        return self._triangle

    def is_lower(self):
        return not self._upper

    def is_upper(self):
        return self._upper

    def is_unit(self):
        return self._unit

    def inner_vect_sv(self, alpha, x, beta, y, trans="N"):
        """Triangular solve y = alpha * op(T)^-1 x + beta * y for the HLG and ELG
        formats (psb_d_cuda_hlg_inner_vect_sv copies to the host and runs the
        CPU HLL solve): here a CSR copy of the stored entries, built on the
        device at the first solve, solves on the device with the CPU's
        results (the CPU HLL and CSR solves take each row's entries in the
        same order)."""
        if not self.is_triangle():
            raise N.PsbError(N.INVALID, f"{self.fmt}_inner_vect_sv", "matrix is not triangular")
        c = getattr(self, "_svcsr", None)
        if c is None:
            csr_type = psb_s_hip_csrg_sparse_mat if self._dtype == np.float32 \
                else psb_d_hip_csrg_sparse_mat
            c = csr_type().cp_from_coo(self.to_coo())
            self._svcsr = c
        c._triangle, c._upper, c._unit = self._triangle, self._upper, self._unit
        c.inner_vect_sv(alpha, x, beta, y, trans=trans)

    def get_nrows(self):
        return self.nr

    def get_ncols(self):
        return self.nc

    def get_nzeros(self):
        return self.nnz

    def get_fmt(self):
        return self.fmt

    def is_dev(self):
        return self.dev is not None

    def mold(self):
        return type(self)()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def _check_vects(self, x, y):
        from .vect import psb_hip_base_vect
        for v in (x, y):
            if not isinstance(v, psb_hip_base_vect):
                raise N.PsbError(N.INVALID, f"{self.fmt}_vect_mv",
                                 "x and y must be device vectors (no host fallback)")
            if v._dtype != self._dtype:
                raise N.PsbError(N.INVALID, f"{self.fmt}_vect_mv", "precision mismatch")
        x.ensure_dev()
        if y.is_host():
            y.sync()

    def get_diag(self, d):
        """get_diag on the device into the vector d (0 where no diagonal)."""
        if self.is_host():
            self.sync()
        d.ensure_dev()
        N.call(self._diag_fn, self.dev, d.dev)
        d.set_dev()

    def vect_mv(self, alpha, x, beta, y):
        """psb_d_cuda_*_vect_mv (cuda/impl/psb_d_cuda_hlg_vect_mv.F90:77-108):
        with device vectors y = alpha*A*x + beta*y on the device (a "host"
        matrix is synced first) and y left "dev"; otherwise on the host copy
        (a "dev" matrix is synced first) with the CPU format's product, and y
        rebuilt from the host result."""
        from .vect import psb_hip_base_vect
        if isinstance(x, psb_hip_base_vect) and isinstance(y, psb_hip_base_vect):
            self._check_vects(x, y)
            if self.is_host():
                self.sync()
            self._dev_mv(alpha, x, beta, y)
            y.set_dev()
        else:
            self._host_mv(alpha, x, beta, y, trans=False)

    def _host_mv(self, alpha, x, beta, y, trans):
        rx = np.asarray(x.get_vect(), dtype=self._dtype)
        ry = np.array(y.get_vect(), dtype=self._dtype)
        nx, ny = (self.nr, self.nc) if trans else (self.nc, self.nr)
        if rx.shape[0] < nx or ry.shape[0] < ny:
            raise N.PsbError(N.INVALID, f"{self.fmt}_vect_mv", "x or y too short")
        if self.is_dev():
            self.sync()
        cols = 1 if rx.ndim == 1 else rx.shape[1]
        for c in range(cols):
            xc = np.ascontiguousarray(rx if rx.ndim == 1 else rx[:, c])
            yc = np.ascontiguousarray(ry if ry.ndim == 1 else ry[:, c])
            if trans:
                self._host_spmv_t(alpha, xc, beta, yc)
            else:
                self._host_spmv(alpha, xc, beta, yc)
            if ry.ndim == 1:
                ry = yc
            else:
                ry[:, c] = yc
        y.bld(ry)

    def _host_spmv_t(self, alpha, x, beta, y):
        """y = alpha*A^T x + beta*y on the host copy, in the CPU loop's order
        (base/serial/impl/psb_d_csr_impl.F90:308-370): y = beta*y, then rows
        ascending, each row's entries in stored order, y(ja) updated one entry
        at a time."""
        irp, ja, val = self._host_csr()
        dt = self._dtype
        a = dt(alpha)
        b = dt(beta)
        if b == 0:
            y[:] = 0
        elif b == -1:
            y[:] = -y
        elif b != 1:
            y[:] = b * y
        if a == 0:
            return
        rows = np.repeat(np.arange(self.nr), np.diff(irp))
        cols = ja.astype(np.int64) - 1
        if a == 1:
            np.add.at(y, cols, val * x[rows])
        elif a == -1:
            np.subtract.at(y, cols, val * x[rows])
        else:
            np.add.at(y, cols, (a * val) * x[rows])

    # ---- state of the dual-memory plug-in (cuda/psb_d_cuda_hlg_mat_mod.F90:
    # 40-84, 303-364): host (-1), sync (0) or dev (1).  Built on the device
    # (cp_from_coo) the matrix is "dev" with no host copy, as the reference's
    # GPU cp_from_coo leaves it; built on the host (cp_from_coo(on_host=True))
    # it is "host" until sync / to_gpu writes the device copy.  sync brings
    # the stale side up to date; device products run on the device copy
    # (syncing a "host" matrix first), products with a host vector on the
    # host copy (syncing a "dev" matrix first); device-side scaling leaves
    # the matrix "dev".
    def is_host(self):
        return self._state == IS_HOST

    def is_dev(self):
        return self._state == IS_DEV

    def is_sync(self):
        return self._state == IS_SYNC

    def set_host(self):
        self._state = IS_HOST

    def set_dev(self):
        self._state = IS_DEV

    def set_sync(self):
        self._state = IS_SYNC

    def sync(self):
        """d_cuda_hlg_sync: host -> to_gpu, dev -> from_gpu, then sync."""
        if self._state == IS_HOST:
            self.to_gpu()
        elif self._state == IS_DEV:
            self.from_gpu()
        self._state = IS_SYNC

    def to_gpu(self):
        """psb_d_cuda_hlg_to_gpu: write the host copy to the device."""
        if self._hc is None:
            raise N.PsbError(N.INVALID, f"{self.fmt}_to_gpu", "no host copy")
        self._write_dev(self._hc)
        self._drop_t()
        self._state = IS_SYNC

    def from_gpu(self):
        """psb_d_cuda_hlg_from_gpu: read the device arrays into the host copy."""
        if self._state != IS_SYNC and self._state != IS_HOST:
            self._hc = self._read_dev()
            self._state = IS_SYNC
        return self._hc

    def host_copy(self):
        """The host copy of the format's arrays (read back first when stale)."""
        if self._state == IS_DEV:
            self.sync()
        return self._hc

    def to_host(self):
        """The format's arrays: the host copy when it is the current one,
        else read from the device."""
        if self._state == IS_HOST:
            return self._hc
        return self._read_dev()

    def to_coo(self):
        """The stored entries as a host COO (row-major sorted)."""
        irp, ja, val = self.to_csr()
        ia = np.repeat(np.arange(1, self.nr + 1, dtype=np.int32), np.diff(irp))
        return psb_d_coo_sparse_mat(self.nr, self.nc, ia, ja, val).fix()

    def cp_from_fmt(self, b):
        """cp_from_fmt: this format built from another storage's entries."""
        return self.cp_from_coo(b.to_coo())

    def mv_from_coo(self, coo):
        """mv_from_coo: build, then leave the source empty."""
        self.cp_from_coo(coo)
        coo.ia = np.zeros(0, np.int32)
        coo.ja = np.zeros(0, np.int32)
        coo.val = np.zeros(0)
        return self

    def mv_from_fmt(self, b):
        self.cp_from_fmt(b)
        b.free()
        return self

    # ---- host-array products (csmv / csmm: device temporaries, as
    # cuda/impl/psb_d_cuda_hlg_csmv.F90:62-81 does)
    def csmv(self, alpha, x, beta, y, trans="N"):
        """y = alpha*op(A)*x + beta*y with host arrays; returns the new y."""
        from .vect import psb_d_hip_vect, psb_s_hip_vect
        t = str(trans).upper()
        if t not in ("N", "T", "C"):
            raise N.PsbError(N.INVALID, f"{self.fmt}_csmv", f"trans {trans!r}")
        mold = psb_s_hip_vect if self._dtype == np.float32 else psb_d_hip_vect
        x = np.asarray(x, self._dtype)
        y = np.asarray(y, self._dtype)
        nx, ny = (self.nc, self.nr) if t == "N" else (self.nr, self.nc)
        if x.shape[0] < nx or y.shape[0] < ny:
            raise N.PsbError(N.INVALID, f"{self.fmt}_csmv", "x or y too short")
        xv, yv = mold().bld(x), mold().bld(y)
        try:
            if t == "N":
                self.vect_mv(alpha, xv, beta, yv)
            else:
                self.vect_mv_t(alpha, xv, beta, yv)
            return yv.get_vect()
        finally:
            xv.free()
            yv.free()

    csmm = csmv   # columns of 2-D x / y are the multivector's columns

    # ---- transposed product (the reference runs it on the host): the
    # transpose is built on the device from the stored entries on first use
    # and kept until the storage changes
    def _drop_t(self):
        t = getattr(self, "_tmat", None)
        if t is not None:
            N.fn("psb_hip_tmat_free")(t)
            self._tmat = None
        c = getattr(self, "_svcsr", None)
        if c is not None:
            c.free()
            self._svcsr = None

    def vect_mv_t(self, alpha, x, beta, y):
        """y = alpha * A^T x + beta * y on the device, equal bit for bit to the
        CPU's trans 'T' / 'C' product (psb_d_csr_csmv, psb_d_hll_csmv); with a
        host vector operand, on the host copy (as the reference always does)."""
        from .vect import psb_hip_base_vect
        if not (isinstance(x, psb_hip_base_vect) and isinstance(y, psb_hip_base_vect)):
            self._host_mv(alpha, x, beta, y, trans=True)
            return
        self._check_vects(x, y)
        if self.is_host():
            self.sync()
        if getattr(self, "_tmat", None) is None:
            irp, ja, val = self.to_csr()
            h = C.c_void_p()
            val = np.ascontiguousarray(val, self._dtype)
            N.call("psb_hip_tmat_alloc", C.byref(h), self.nr, self.nc, irp.ctypes.data,
                   np.ascontiguousarray(ja, np.int32).ctypes.data, val.ctypes.data, self._etype)
            self._tmat = h
        N.call("psb_hip_tmat_spmv", self._tmat, float(alpha), x.dev, float(beta), y.dev)
        y.set_dev()

    # ---- scaling (scals / scal): on the device copy
    def make_nonunit(self):
        """make_nonunit: a unit triangle gets its unit diagonal stored."""
        if not self._unit:
            return
        coo = self.to_coo()
        k = min(self.nr, self.nc)
        d = np.arange(1, k + 1, dtype=np.int32)
        coo = psb_d_coo_sparse_mat(self.nr, self.nc, np.concatenate([coo.ia, d]),
                                   np.concatenate([coo.ja, d]), np.concatenate([coo.val, np.ones(k)]))
        flags = (self._triangle, self._upper)
        self.cp_from_coo(coo.fix())
        self._triangle, self._upper, self._unit = flags[0], flags[1], False

    def scals(self, d):
        """psb_d_cuda_*_scals: every value times the scalar d (on the device
        copy; the matrix is then "dev")."""
        self.make_nonunit()
        self._drop_t()
        if self.is_host():
            self.sync()
        N.call(self._pfx + "_scals", self.dev, float(d))
        self._state = IS_DEV

    def scal(self, d, side="L"):
        """psb_d_cuda_*_scal: rows (side 'L') or columns ('R') times d."""
        from .vect import psb_d_hip_vect, psb_hip_base_vect, psb_s_hip_vect
        left = str(side).upper() == "L"
        need = self.nr if left else self.nc
        self.make_nonunit()
        self._drop_t()
        if self.is_host():
            self.sync()
        self._state = IS_DEV
        if isinstance(d, psb_hip_base_vect):
            if d._dtype != self._dtype or d.get_nrows() < need:
                raise N.PsbError(N.INVALID, f"{self.fmt}_scal", "d precision or size")
            d.ensure_dev()
            N.call(self._pfx + "_scal", self.dev, d.dev, 1 if left else 0)
            return
        d = np.asarray(d, self._dtype)
        if d.shape[0] < need:
            raise N.PsbError(N.INVALID, f"{self.fmt}_scal", "d too short")
        mold = psb_s_hip_vect if self._dtype == np.float32 else psb_d_hip_vect
        dv = mold().bld(d)
        try:
            N.call(self._pfx + "_scal", self.dev, dv.dev, 1 if left else 0)
        finally:
            dv.free()


class psb_d_hip_hlg_sparse_mat(_hip_base_mat):
    """HLL on the device (psb_d_cuda_hlg_sparse_mat).  hksz defaults to the
    wavefront size, 64 on gfx950, as the reference GPU type uses WarpSize
    (cuda/impl/psb_d_cuda_cp_hlg_from_coo.F90:60).

    sort_window: None (the default) stores rows in order, the reference's
    layout.  Otherwise rows are sorted by descending length inside windows of
    sort_window rows (0 or less: the whole matrix) and the hacks are built
    over that order (the reorder array of the SPGPU HLL kernel,
    cuda/spgpu/kernels/hell_spmv_base_template.cuh:227-252): each hack then
    holds rows of similar length, so a matrix with very uneven rows is not
    padded to its longest rows.  irn / idiag / the hack layout are then per
    position and self.perm gives the row (1-based) at each position."""
    fmt = "HLG"
    _etype = N.TYPE_DOUBLE
    _dtype = np.float64
    _diag_fn = "psb_hip_hll_get_diag"
    _pfx = "psb_hip_hll"

    def __init__(self, hksz=64, sort_window=None):
        super().__init__()
        self.hksz = int(hksz)
        self.sort_window = None if sort_window is None else int(sort_window)
        self.irn = self.hkoffs = self.perm = None
        self.isz = 0

    def mold(self):
        return type(self)(hksz=self.hksz, sort_window=self.sort_window)

    def cp_from_coo(self, coo, on_host=False):
        """psb_d_cuda_cp_hlg_from_coo: offsets on the host, fill on the device
        (with the row order of sort_window when set); the matrix is then "dev".
        on_host=True fills the host copy instead (the CPU format's
        psi_d_convert_hll_from_coo) and leaves the matrix "host" until sync."""
        if not coo.sorted:
            coo = psb_d_coo_sparse_mat(coo.nr, coo.nc, coo.ia, coo.ja, coo.val).fix()
        self.free()
        nr, nz = coo.nr, coo.get_nzeros()
        noffs = (nr + self.hksz - 1) // self.hksz
        irn = np.zeros(max(nr, 1), np.int32)
        idisp = np.zeros(nr + 1, np.int32)
        hkoffs = np.zeros(noffs + 1, np.int32)
        isz = C.c_int64(0)
        N.call("psb_hip_hll_offsets", nr, nz, _ptr(coo.ia), self.hksz, irn.ctypes.data,
               idisp.ctypes.data, hkoffs.ctypes.data, C.byref(isz))
        perm = None
        if self.sort_window is not None:
            perm = np.zeros(max(nr, 1), np.int32)
            irn_pos = np.zeros(max(nr, 1), np.int32)
            N.call("psb_hip_hll_sort_rows", nr, self.hksz, self.sort_window, irn.ctypes.data,
                   perm.ctypes.data, irn_pos.ctypes.data, hkoffs.ctypes.data, C.byref(isz))
        h = C.c_void_p()
        N.call("psb_hip_hll_alloc", C.byref(h), self.hksz, nr, nz, int(isz.value), self._etype, 1)
        self.dev = h
        val = np.ascontiguousarray(coo.val, dtype=self._dtype)
        if on_host:
            cnt = irn if perm is None else irn_pos
            hja = np.zeros(max(int(isz.value), 1), np.int32)
            hval = np.zeros(max(int(isz.value), 1), self._dtype)
            hdiag = np.zeros(max(nr, 1), np.int32)
            N.call("psb_hip_host_hll_fill", nr, self.hksz, cnt.ctypes.data, hkoffs.ctypes.data,
                   idisp.ctypes.data, _ptr(coo.ja), _ptr(val),
                   None if perm is None else perm.ctypes.data, self._etype, hja.ctypes.data,
                   hval.ctypes.data, hdiag.ctypes.data)
            self._hc = (hkoffs, cnt[:nr].copy(), hja[:int(isz.value)], hval[:int(isz.value)],
                        hdiag[:nr])
            self._state = IS_HOST
            self.nr, self.nc, self.nnz = nr, coo.nc, nz
            self.irn, self.hkoffs, self.isz = cnt[:nr], hkoffs, int(isz.value)
            self.perm = None if perm is None else perm[:nr]
            return self
        if perm is None:
            N.call("psb_hip_hll_copy_coo", nr, coo.nc, nz, self.hksz, noffs, int(isz.value),
                   irn.ctypes.data, hkoffs.ctypes.data, idisp.ctypes.data, _ptr(coo.ja), _ptr(val),
                   self.dev)
        else:
            N.call("psb_hip_hll_copy_coo_sorted", nr, coo.nc, nz, self.hksz, noffs,
                   int(isz.value), perm.ctypes.data, irn_pos.ctypes.data, hkoffs.ctypes.data,
                   idisp.ctypes.data, _ptr(coo.ja), _ptr(val), self.dev)
            irn = irn_pos
        self.nr, self.nc, self.nnz = nr, coo.nc, nz
        self.irn, self.hkoffs, self.isz = irn[:nr], hkoffs, int(isz.value)
        self.perm = None if perm is None else perm[:nr]
        self._hc, self._state = None, IS_DEV
        return self

    def _write_dev(self, hc):
        hk, irn, ja, val, idiag = hc
        val = np.ascontiguousarray(val, self._dtype)
        N.call("psb_hip_hll_write", self.dev, val.ctypes.data, ja.ctypes.data, hk.ctypes.data,
               irn.ctypes.data, idiag.ctypes.data)
        if self.perm is not None:
            N.call("psb_hip_hll_set_perm", self.dev, self.perm.ctypes.data, hk.ctypes.data)

    def _dev_mv(self, alpha, x, beta, y):
        N.call("psb_hip_hll_spmv", self.dev, float(alpha), x.dev, float(beta), y.dev)

    def _host_spmv(self, alpha, x, beta, y):
        hk, irn, ja, val, _ = self._hc
        N.call("psb_hip_host_hll_spmv", self.nr, self.hksz, hk.ctypes.data, irn.ctypes.data,
               ja.ctypes.data, val.ctypes.data,
               None if self.perm is None else self.perm.ctypes.data, self._etype,
               float(alpha), x.ctypes.data, float(beta), y.ctypes.data)

    def _host_csr(self):
        return self._csr_of(self._hc)

    def _read_dev(self):
        """Read the device arrays back: (hkoffs, irn, ja, val, idiag)."""
        val = np.empty(max(self.isz, 1), self._dtype)
        ja = np.empty(max(self.isz, 1), np.int32)
        hk = np.empty((self.nr + self.hksz - 1) // self.hksz + 1, np.int32)
        irn = np.empty(max(self.nr, 1), np.int32)
        idiag = np.empty(max(self.nr, 1), np.int32)
        N.call("psb_hip_hll_read", self.dev, val.ctypes.data, ja.ctypes.data, hk.ctypes.data,
               irn.ctypes.data, idiag.ctypes.data)
        return hk, irn[:self.nr], ja[:self.isz], val[:self.isz], idiag[:self.nr]

    def to_csr(self):
        """The stored entries as host CSR (irp, ja, val), 1-based, padding dropped:
        row r's entries are the first irn slots of its hack column (of the
        position holding row r when the rows are sorted)."""
        return self._csr_of(self.to_host())

    def _csr_of(self, arrays):
        hk, irn, ja, val, _ = arrays
        nr, hs = self.nr, self.hksz
        if self.perm is None:
            order = np.arange(nr, dtype=np.int64)          # position of each row
        else:
            order = np.empty(nr, np.int64)
            order[self.perm.astype(np.int64) - 1] = np.arange(nr, dtype=np.int64)
        cnt = irn[order] if nr else irn
        posn = np.repeat(order, cnt)
        start = np.concatenate([[0], np.cumsum(cnt, dtype=np.int64)[:-1]])
        j = np.arange(len(posn), dtype=np.int64) - np.repeat(start, cnt)
        pos = hk[posn // hs].astype(np.int64) + j * hs + posn % hs
        irp = np.concatenate([[1], 1 + np.cumsum(cnt, dtype=np.int64)]).astype(np.int32)
        return irp, ja[pos].astype(np.int32), val[pos].copy()

    def units(self):
        """(work units, hacks cut into several units) of a row-sorted matrix."""
        nu, ns = C.c_int(0), C.c_int(0)
        N.call("psb_hip_hll_units", self.dev, C.byref(nu), C.byref(ns))
        return nu.value, ns.value

    def tune(self, x, y, reps=20):
        """Time the HLL kernel variants on this matrix with vectors x, y and keep
        the fastest for it (psb_hip_hll_tune).  Returns (variant, {variant: us})."""
        self._check_vects(x, y)
        best = C.c_int(0)
        us = (C.c_float * 4)()
        N.call("psb_hip_hll_tune", self.dev, x.dev, y.dev, int(reps), C.byref(best), us)
        y.set_dev()
        return best.value, {v: float(t) for v, t in zip((1, 2, 5, 6), us)}

    def sizeof(self):
        es = np.dtype(self._dtype).itemsize
        return self.isz * (es + 4) + 4 * (2 * self.nr + len(self.hkoffs))

    def free(self):
        self._drop_t()
        if self.dev is not None:
            N.fn("psb_hip_hll_free")(self.dev)
            self.dev = None


class psb_s_hip_hlg_sparse_mat(psb_d_hip_hlg_sparse_mat):
    """Single-precision HLL on the device (psb_s_cuda_hlg_sparse_mat)."""
    _etype = N.TYPE_FLOAT
    _dtype = np.float32


class psb_d_hip_elg_sparse_mat(_hip_base_mat):
    """ELL on the device (psb_d_cuda_elg_sparse_mat), pitch ceil32(rows)."""
    fmt = "ELG"
    _etype = N.TYPE_DOUBLE
    _dtype = np.float64
    _diag_fn = "psb_hip_ell_get_diag"
    _pfx = "psb_hip_ell"

    def __init__(self):
        super().__init__()
        self.nzm = 0

    def cp_from_coo(self, coo, on_host=False):
        """psb_d_cuda_cp_elg_from_coo: filled on the device ("dev"), or with
        on_host=True the host copy in the CPU layout (leading dimension nr,
        psi_d_convert_ell_from_coo + psi_d_xtr_ell_from_coo's padding: a row's
        last column, an empty row the matrix's first column) left "host"."""
        if not coo.sorted:
            coo = psb_d_coo_sparse_mat(coo.nr, coo.nc, coo.ia, coo.ja, coo.val).fix()
        self.free()
        nr, nz = coo.nr, coo.get_nzeros()
        irn = np.bincount(coo.ia - 1, minlength=nr).astype(np.int32) if nz else np.zeros(nr, np.int32)
        idisp = np.zeros(nr + 1, np.int32)
        np.cumsum(irn, out=idisp[1:])
        nzm = int(irn.max()) if nr else 0
        slots = ((nr + 31) // 32) * 32 * nzm
        es = np.dtype(self._dtype).itemsize
        if slots * (es + 4) > 200 * 2**30:
            raise N.PsbError(N.OUTOFMEM, "ELG cp_from_coo",
                             f"ELL needs {nr} x {nzm} slots: not representable")
        h = C.c_void_p()
        N.call("psb_hip_ell_alloc", C.byref(h), nr, nzm, nz, self._etype, 1)
        self.dev = h
        val = np.ascontiguousarray(coo.val, dtype=self._dtype)
        self.nr, self.nc, self.nnz, self.nzm = nr, coo.nc, nz, nzm
        self.irn = irn
        if on_host:
            rows = (coo.ia - 1).astype(np.int64)
            j = np.arange(nz, dtype=np.int64) - idisp[rows]
            last = np.where(irn > 0, coo.ja[np.maximum(idisp[1:] - 1, 0)] if nz else 1,
                            coo.ja[0] if nz else 1).astype(np.int32)
            ja2 = np.repeat(last[None, :], nzm, axis=0)            # (nzm, nr): column-major
            va2 = np.zeros((nzm, nr), self._dtype)
            ja2[j, rows] = coo.ja
            va2[j, rows] = val
            diag = np.zeros(nr, np.int32)
            on = coo.ia == coo.ja
            diag[rows[on]] = (j[on] + 1).astype(np.int32)
            self._hc = (irn.copy(), ja2.ravel(), va2.ravel(), diag)
            self._state = IS_HOST
            return self
        N.call("psb_hip_ell_copy_coo", nr, coo.nc, nz, _ptr(irn), _ptr(idisp), _ptr(coo.ja),
               _ptr(val), self.dev)
        self._hc, self._state = None, IS_DEV
        return self

    def _read_dev(self):
        """(irn, ja, val, idiag) with ja/val of shape (nr, nzm) column-major."""
        ja = np.empty(max(self.nr * self.nzm, 1), np.int32)
        val = np.empty(max(self.nr * self.nzm, 1), self._dtype)
        irn = np.empty(max(self.nr, 1), np.int32)
        idiag = np.empty(max(self.nr, 1), np.int32)
        N.call("psb_hip_ell_read", self.dev, val.ctypes.data, ja.ctypes.data, max(self.nr, 1),
               irn.ctypes.data, idiag.ctypes.data)
        n = self.nr * self.nzm
        return irn[:self.nr], ja[:n], val[:n], idiag[:self.nr]

    def to_csr(self):
        """The stored entries as host CSR (irp, ja, val), padding dropped."""
        return self._csr_of(self.to_host())

    def _host_csr(self):
        return self._csr_of(self._hc)

    def _write_dev(self, hc):
        irn, ja, val, idiag = hc
        val = np.ascontiguousarray(val, self._dtype)
        N.call("psb_hip_ell_write", self.dev, val.ctypes.data, ja.ctypes.data, max(self.nr, 1),
               irn.ctypes.data, idiag.ctypes.data)

    def _dev_mv(self, alpha, x, beta, y):
        N.call("psb_hip_ell_spmv", self.dev, float(alpha), x.dev, float(beta), y.dev)

    def _host_spmv(self, alpha, x, beta, y):
        irn, ja, val, _ = self._hc
        N.call("psb_hip_host_ell_spmv", self.nr, max(self.nr, 1), irn.ctypes.data,
               ja.ctypes.data, val.ctypes.data, self._etype, float(alpha), x.ctypes.data,
               float(beta), y.ctypes.data)

    def _csr_of(self, arrays):
        irn, ja, val, _ = arrays
        rows = np.repeat(np.arange(self.nr, dtype=np.int64), irn)
        start = np.concatenate([[0], np.cumsum(irn, dtype=np.int64)[:-1]])
        j = np.arange(len(rows), dtype=np.int64) - np.repeat(start, irn)
        pos = j * self.nr + rows
        irp = np.concatenate([[1], 1 + np.cumsum(irn, dtype=np.int64)]).astype(np.int32)
        return irp, ja[pos].astype(np.int32), val[pos].copy()

    def pitch(self):
        return N.fn("psb_hip_ell_pitch")(self.dev)

    def free(self):
        self._drop_t()
        if self.dev is not None:
            N.fn("psb_hip_ell_free")(self.dev)
            self.dev = None


class psb_s_hip_elg_sparse_mat(psb_d_hip_elg_sparse_mat):
    _etype = N.TYPE_FLOAT
    _dtype = np.float32


class psb_d_hip_csrg_sparse_mat(_hip_base_mat):
    """CSR on the device (psb_d_cuda_csrg_sparse_mat), with hand-written
    vector-CSR and wave-per-row kernels in place of cuSPARSE."""
    fmt = "CSRG"
    _etype = N.TYPE_DOUBLE
    _dtype = np.float64
    _diag_fn = "psb_hip_csr_get_diag"
    _pfx = "psb_hip_csr"

    def cp_from_coo(self, coo, on_host=False):
        """psb_d_cp_csr_from_coo (base/serial/impl/psb_d_csr_impl.F90:3154): the
        device copy written ("dev"), or with on_host=True only the host copy
        ("host" until sync)."""
        if not coo.sorted:
            coo = psb_d_coo_sparse_mat(coo.nr, coo.nc, coo.ia, coo.ja, coo.val).fix()
        self.free()
        nr, nz = coo.nr, coo.get_nzeros()
        cnt = np.bincount(coo.ia - 1, minlength=nr) if nz else np.zeros(nr, np.int64)
        irp = np.ones(nr + 1, np.int32)
        np.cumsum(cnt, out=irp[1:])
        irp[1:] += 1
        h = C.c_void_p()
        N.call("psb_hip_csr_alloc", C.byref(h), nr, coo.nc, nz, self._etype, 1)
        self.dev = h
        val = np.ascontiguousarray(coo.val, dtype=self._dtype)
        self.nr, self.nc, self.nnz = nr, coo.nc, nz
        if on_host:
            self._hc = (irp, np.ascontiguousarray(coo.ja, np.int32).copy(), val.copy())
            self._state = IS_HOST
            return self
        N.call("psb_hip_csr_write", self.dev, irp.ctypes.data, _ptr(coo.ja), _ptr(val))
        self._hc, self._state = None, IS_DEV
        return self

    def _read_dev(self):
        irp = np.empty(self.nr + 1, np.int32)
        ja = np.empty(max(self.nnz, 1), np.int32)
        val = np.empty(max(self.nnz, 1), self._dtype)
        N.call("psb_hip_csr_read", self.dev, irp.ctypes.data, ja.ctypes.data, val.ctypes.data)
        return irp, ja[:self.nnz], val[:self.nnz]

    def to_csr(self):
        return self.to_host()

    def _host_csr(self):
        return self._hc

    def _write_dev(self, hc):
        irp, ja, val = hc
        val = np.ascontiguousarray(val, self._dtype)
        N.call("psb_hip_csr_write", self.dev, irp.ctypes.data, _ptr(ja), _ptr(val))

    def _dev_mv(self, alpha, x, beta, y):
        N.call("psb_hip_csr_spmv", self.dev, float(alpha), x.dev, float(beta), y.dev)

    def _host_spmv(self, alpha, x, beta, y):
        irp, ja, val = self._hc
        N.call("psb_hip_host_csr_spmv", self.nr, irp.ctypes.data, _ptr(ja), _ptr(val),
               self._etype, float(alpha), x.ctypes.data, float(beta), y.ctypes.data)

    def analysis(self):
        lanes, nlong, thr = C.c_int(), C.c_int(), C.c_int()
        N.call("psb_hip_csr_get_analysis", self.dev, C.byref(lanes), C.byref(nlong), C.byref(thr))
        return lanes.value, nlong.value, thr.value

    def inner_vect_sv(self, alpha, x, beta, y, trans="N"):
        """psb_d_cuda_csrg_inner_vect_sv (cuda/impl/psb_d_cuda_csrg_inner_vect_sv.F90):
        y = alpha * op(T)^-1 x + beta * y on the device, with the results of
        psb_d_csr_cssv (base/serial/impl/psb_d_csr_impl.F90:782-1039) for every
        trans ('N', 'T', 'C') and beta; the matrix must be flagged triangular
        (psb_err_invalid_mat_state_ otherwise, as the reference)."""
        if not self.is_triangle():
            raise N.PsbError(N.INVALID, "csrg_inner_vect_sv", "matrix is not triangular")
        t = str(trans).upper()
        if t not in ("N", "T", "C"):
            raise N.PsbError(N.INVALID, "csrg_inner_vect_sv", f"trans {trans!r}")
        self._check_vects(x, y)
        N.call("psb_hip_csr_spsv", self.dev, 1 if self.is_lower() else 0,
               1 if self.is_unit() else 0, "NTC".index(t), float(alpha), x.dev, float(beta),
               y.dev)
        y.set_dev()

    def slices(self):
        """Column slices the product runs on (0: the whole matrix); see
        psb_hip_csr_set_slices."""
        return int(N.fn("psb_hip_csr_slices")(self.dev))

    def sv_levels(self):
        """(levels, workgroup runs) of the prepared triangular solve, (0, 0)
        before the first solve."""
        lv, sg = C.c_int(), C.c_int()
        N.call("psb_hip_csr_spsv_levels", self.dev, C.byref(lv), C.byref(sg))
        return lv.value, sg.value

    def free(self):
        self._drop_t()
        if self.dev is not None:
            N.fn("psb_hip_csr_free")(self.dev)
            self.dev = None


class psb_s_hip_csrg_sparse_mat(psb_d_hip_csrg_sparse_mat):
    _etype = N.TYPE_FLOAT
    _dtype = np.float32


class psb_dspmat_type:
    """psb_dspmat_type: ``a`` holds the storage plug-in.  Build state is a
    host COO; psb_spasb converts it to the mold's format."""

    def __init__(self):
        self.a = None
        self.coo = None
        self.m = self.n = 0
        self._tri = (False, False, False)   # triangle, upper, unit (until assembly)

    # the triangle flags (psb_dspmat_type%set_triangle etc.): kept here in the
    # build state and carried into the storage at assembly / conversion
    def _set_flag(self, k, val):
        t = list(self._tri)
        t[k] = bool(val)
        self._tri = tuple(t)
        if self.a is not None:
            self.a._triangle, self.a._upper, self.a._unit = self._tri

    def set_triangle(self, val=True):
        self._set_flag(0, val)

    def set_upper(self, val=True):
        self._set_flag(1, val)

    def set_lower(self, val=True):
        self._set_flag(1, not val)

    def set_unit(self, val=True):
        self._set_flag(2, val)

    def is_triangle(self):
        return self._tri[0]

    def is_upper(self):
        return self._tri[1]

    def is_lower(self):
        return not self._tri[1]

    def is_unit(self):
        return self._tri[2]

    def get_nrows(self):
        return self.m

    def get_ncols(self):
        return self.n

    def get_nzeros(self):
        return self.a.get_nzeros() if self.a is not None else self.coo.get_nzeros()

    def get_fmt(self):
        return self.a.get_fmt() if self.a is not None else "COO"

    def free(self):
        if self.a is not None:
            self.a.free()
        self.a = None
        self.coo = None
