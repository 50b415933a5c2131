"""Device vectors: psb_[dsi]_hip_vect, the plug-ins that take the place of
psb_[dsi]_vect_cuda (cuda/psb_d_cuda_vect_mod.F90:48-1327), and the
psb_d_vect_type wrapper whose ``v`` component holds one of them
(base/modules/serial/psb_d_vect_mod.F90).

Dual-memory semantics follow the reference: a host copy ``v.h`` and a
device multivector ``v.dev``; state is host (-1), sync (0) or dev (1)
(cuda/psb_d_cuda_vect_mod.F90:40-42).  Operations on device vectors run on
the device and mark their outputs dev; the host copy is refreshed only by
``sync()``.  There is no host fallback: mixing a device vector with a
non-device operand raises ``PsbError``.
"""
import ctypes as C

import numpy as np

from . import _native as N

IS_HOST, IS_SYNC, IS_DEV = -1, 0, 1


class psb_hip_base_vect:
    _etype = None
    _dtype = None
    fmt = "HIP"

    def __init__(self):
        self.dev = None
        self.h = None
        self.state = IS_HOST
        self.n = 0
        self.count = 1

    # --------------------------------------------------------------- life cycle
    def all(self, n, count=1):
        """d_cuda_all: allocate n entries (x count columns), zero on both sides."""
        self.free()
        h = C.c_void_p()
        N.call("psb_hip_mv_alloc", C.byref(h), count, n, self._etype)
        self.dev, self.n, self.count = h, n, count
        self.h = np.zeros((n,) if count == 1 else (n, count), dtype=self._dtype, order="F")
        self.zero()
        return self

    def asb(self, n):
        """d_cuda_asb_m (cuda/psb_d_cuda_vect_mod.F90:677-702): make room for
        at least n entries, keeping the values (new entries are zero)."""
        if n <= self.n:
            return self
        old = self.get_vect()
        grown = np.zeros((n,) if self.count == 1 else (n, self.count), dtype=self._dtype)
        grown[:self.n] = old
        return self.bld(grown)

    def bld(self, values):
        """d_cuda_bld_x: build from host values (n,) or (n, count)."""
        a = np.asarray(values, dtype=self._dtype)
        count = 1 if a.ndim == 1 else a.shape[1]
        self.all(a.shape[0], count)
        self.h[...] = a
        self.state = IS_HOST
        self.sync()
        return self

    def free(self):
        if self.dt_buf is not None:
            self.free_buffer()
        if self.dev is not None:
            N.fn("psb_hip_mv_free")(self.dev)
            self.dev = None
        self.h = None
        self.n = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def mold(self):
        return type(self)()

    # --------------------------------------------------------------- state
    def is_host(self):
        return self.state == IS_HOST

    def is_dev(self):
        return self.state == IS_DEV

    def is_sync(self):
        return self.state == IS_SYNC

    def set_host(self):
        self.state = IS_HOST

    def set_dev(self):
        self.state = IS_DEV

    def set_sync(self):
        self.state = IS_SYNC

    def sync(self):
        """d_cuda_sync: bring the stale side up to date."""
        if self.dev is None:
            raise N.PsbError(N.INVALID, "sync", "vector not allocated")
        if self.state == IS_HOST:
            buf = np.ascontiguousarray(self.h.T if self.count > 1 else self.h)
            N.call("psb_hip_mv_write", self.dev, buf.ctypes.data)
        elif self.state == IS_DEV:
            if self.count == 1:
                out = np.empty(self.n, dtype=self._dtype)
                N.call("psb_hip_mv_read", self.dev, out.ctypes.data)
                self.h[:] = out
            else:
                out = np.empty((self.count, self.n), dtype=self._dtype)
                N.call("psb_hip_mv_read", self.dev, out.ctypes.data)
                self.h[...] = out.T
        self.state = IS_SYNC

    def ensure_dev(self):
        if self.state == IS_HOST:
            self.sync()

    def get_vect(self):
        if self.state == IS_DEV:
            self.sync()
        return self.h.copy()

    def get_nrows(self):
        return self.n

    # --------------------------------------------------------------- fills
    def zero(self):
        self.set_scal(0)

    def set_scal(self, val, first=None, last=None):
        first = 1 if first is None else first
        last = self.n * self.count if last is None else last
        if last >= first:
            N.call(self._setscal, val, first, last, 1, self.dev)
        # both copies get the same range, so whichever side was current stays so
        if self.h is not None:
            self.h.reshape(-1, order="F")[first - 1:last] = val

    # --------------------------------------------------------------- BLAS-1 methods
    # The type-bound procedures of psb_d_vect_cuda (cuda/psb_d_cuda_vect_mod.F90:
    # 59-103); the receiving object is the one the reference passes (pass(x),
    # pass(y) or pass(z)).  All run on the device and leave outputs dev.
    def _peer(self, *vs):
        for v in vs:
            if not isinstance(v, psb_hip_base_vect) or v._dtype != self._dtype:
                raise N.PsbError(N.INVALID, "vect", "operand is not a device vector of the same "
                                 "type; no host fallback")
            v.ensure_dev()
        self.ensure_dev()

    def _reduce(self, name, n):
        self._peer()
        res = (C.c_double * max(self.count, 1))()
        N.call(name, res, self.n if n is None else n, self.dev)
        return res[0] if self.count == 1 else np.array(res[:self.count])

    def dot_v(self, n, y):
        """d_cuda_dot_v (:831-865): sum_{i<=n} x(i) y(i)."""
        self._peer(y)
        res = (C.c_double * max(self.count, 1))()
        N.call("psb_hip_mv_dot", res, n, self.dev, y.dev)
        return res[0] if self.count == 1 else np.array(res[:self.count])

    def axpby_v(self, m, alpha, x, beta):
        """d_cuda_axpby_v (:880-913), pass(y): y = alpha x + beta y."""
        self._peer(x)
        N.call("psb_hip_mv_axpby", m, float(alpha), x.dev, float(beta), self.dev)
        self.set_dev()

    def upd_xyz(self, m, alpha, beta, gamma, delta, x, y):
        """d_cuda_upd_xyz (:915-975), pass(z): y = alpha x + beta y;
        z = gamma y + delta z."""
        self._peer(x, y)
        N.call("psb_hip_mv_upd_xyz", m, float(alpha), float(beta), float(gamma), float(delta),
               x.dev, y.dev, self.dev)
        y.set_dev()
        self.set_dev()

    def xyzw(self, m, a, b, c, d, e, f, x, y, z):
        """d_cuda_xyzw (:977-1038), pass(w): y = a x + b y; z = c y + d z;
        w = e z + f w."""
        self._peer(x, y, z)
        N.call("psb_hip_mv_xyzw", m, *[float(t) for t in (a, b, c, d, e, f)], x.dev, y.dev,
               z.dev, self.dev)
        for v in (y, z, self):
            v.set_dev()

    def mlt_v(self, x):
        """d_cuda_mlt_v (:1055-1086), pass(y): y = x .* y."""
        self._peer(x)
        N.call("psb_hip_mv_axy", self.n, 1.0, x.dev, self.dev)
        self.set_dev()

    def mlt_v_2(self, alpha, x, y, beta):
        """d_cuda_mlt_v_2 (:1118-1172), pass(z): z = alpha x .* y + beta z."""
        self._peer(x, y)
        N.call("psb_hip_mv_axybz", self.n, float(alpha), x.dev, y.dev, float(beta), self.dev)
        self.set_dev()

    def scal(self, alpha):
        """d_cuda_scal (:1174-1183): x = alpha x."""
        self._peer()
        N.call("psb_hip_mv_scal", float(alpha), self.dev)
        self.set_dev()

    def nrm2(self, n=None):
        """d_cuda_nrm2 (:1186-1196)."""
        return self._reduce("psb_hip_mv_nrm2", n)

    def amax(self, n=None):
        """d_cuda_amax (:1198-1208)."""
        return self._reduce("psb_hip_mv_amax", n)

    def asum(self, n=None):
        """d_cuda_asum: sum |x(i)|."""
        return self._reduce("psb_hip_mv_asum", n)

    def absval(self, y=None):
        """d_cuda_absval1 / absval2: x = |x|, or y = |x| when y is given."""
        out = self if y is None else y
        self._peer(out)
        N.call("psb_hip_mv_abs", out.n, 1.0, self.dev, out.dev)
        out.set_dev()

    @staticmethod
    def device_wait():
        """d_cuda_device_wait: wait for the library stream."""
        N.call("psb_hip_device_sync")

    # --------------------------------------------------------------- insertion
    def ins_v(self, n, irl, val, dupl):
        """d_cuda_ins_v: x(irl(i)) = val(i) (dupl 0 overwrite) or += val(i)
        (dupl 1 add), i = 1..n; irl a psb_i_hip_vect, val of x's type."""
        if not isinstance(irl, psb_i_hip_vect):
            raise N.PsbError(N.INVALID, "ins_v", "irl must be a device index vector")
        self._peer(val)
        idx = irl.get_vect()[:n]
        if n > 0 and (idx.min() < 1 or idx.max() > self.n):
            raise N.PsbError(N.INVALID, "ins_v", "index outside 1..n")
        N.call("psb_hip_mv_geins", n, irl.dev, val.dev, int(dupl), 1, self.dev)
        self.set_dev()

    def ins_a(self, n, irl, val, dupl):
        """d_cuda_ins_a: the same with host arrays (moved to the device)."""
        iv = psb_i_hip_vect().bld(np.asarray(irl[:n], np.int32))
        vv = type(self)().bld(np.asarray(val[:n], self._dtype))
        try:
            self.ins_v(n, iv, vv, dupl)
        finally:
            iv.free()
            vv.free()

    # --------------------------------------------------------------- halo buffers
    # combuf is the host communication buffer; dt_buf its device staging copy
    # (d_cuda_new_buffer, :130-176).  Index vectors are 1-based.
    combuf = None
    dt_buf = None
    dt_buf_sz = 0

    def new_buffer(self, n):
        """d_cuda_new_buffer (:130-176): combuf(n) on the host and a device
        buffer of n entries, each kept if already large enough."""
        if self.combuf is None or len(self.combuf) < n:
            self.combuf = np.zeros(n, dtype=self._dtype)
        if self.dt_buf is not None and self.dt_buf_sz < n:
            N.call("psb_hip_free_buffer", self.dt_buf)
            self.dt_buf, self.dt_buf_sz = None, 0
        if self.dt_buf is None:
            p = C.c_void_p()
            N.call("psb_hip_alloc_buffer", C.byref(p), max(n, 1) * np.dtype(self._dtype).itemsize)
            self.dt_buf, self.dt_buf_sz = p, n

    def free_buffer(self):
        """d_cuda_free_buffer (:190-220)."""
        if self.dt_buf is not None:
            N.fn("psb_hip_free_buffer")(self.dt_buf)
        self.dt_buf, self.dt_buf_sz, self.combuf = None, 0, None

    maybe_free_buffer = free_buffer

    def _index_window(self, name, i, n, idx):
        if not isinstance(idx, psb_i_hip_vect):
            raise N.PsbError(N.INVALID, name, "idx must be a device index vector")
        if i < 1 or i + n - 1 > idx.n:
            raise N.PsbError(N.INVALID, name, "idx(i:i+n-1) outside idx")
        w = idx.get_vect()[i - 1:i + n - 1]
        if n > 0 and (w.min() < 1 or w.max() > self.n):
            raise N.PsbError(N.INVALID, name, "index outside 1..n")
        idx.ensure_dev()
        self.ensure_dev()
        return N.fn("psb_hip_mv_data")(idx.dev)

    def _stage(self, n):
        if self.dt_buf is None or self.dt_buf_sz < n:
            keep = self.combuf
            self.new_buffer(max(n, len(keep) if keep is not None else 0))
            if keep is not None:
                self.combuf[:len(keep)] = keep

    def gthzv_x(self, i, n, idx, y=None):
        """d_cuda_gthzv_x (:222-329): y(1:n) = x(idx(i:i+n-1)), gathered on
        the device and read back once; y is a host array (allocated and
        returned when not given)."""
        didx = self._index_window("gthzv_x", i, n, idx)
        if y is not None and (len(y) < n or np.asarray(y).dtype != self._dtype):
            raise N.PsbError(N.INVALID, "gthzv_x", "y shorter than n or of another type")
        out = np.empty(n, dtype=self._dtype)
        if n == 0:
            return out if y is None else y
        self._stage(n)
        N.call("psb_hip_mv_igath", self.dev, 0, n, i, didx, 1, self.dt_buf, 1)
        N.call("psb_hip_read_buffer", out.ctypes.data, self.dt_buf, n * out.itemsize)
        if y is None:
            return out
        y[:n] = out
        return y

    def gthzbuf(self, i, n, idx):
        """d_cuda_gthzbuf (:331-382): combuf(i:i+n-1) = x(idx(i:i+n-1))."""
        if self.combuf is None:
            raise N.PsbError(N.INVALID, "gthzbuf", "no communication buffer (new_buffer)")
        if i + n - 1 > len(self.combuf):
            raise N.PsbError(N.INVALID, "gthzbuf", "window beyond the buffer")
        didx = self._index_window("gthzbuf", i, n, idx)
        if n == 0:
            return
        self._stage(i + n - 1)
        N.call("psb_hip_mv_igath", self.dev, 0, n, i, didx, i, self.dt_buf, 1)
        isz = np.dtype(self._dtype).itemsize
        tmp = np.empty(n, dtype=self._dtype)
        N.call("psb_hip_read_buffer", tmp.ctypes.data,
               C.c_void_p(self.dt_buf.value + (i - 1) * isz), n * isz)
        self.combuf[i - 1:i + n - 1] = tmp

    def sctb(self, n, idx, x, beta):
        """d_cuda_sctb (:384-399), pass(y): y(idx(k)) = beta y(idx(k)) + x(k)
        with host idx and x; like the reference this runs on the host copy and
        leaves y host."""
        if n == 0:
            return
        if self.state == IS_DEV:
            self.sync()
        ix = np.asarray(idx[:n], np.int64) - 1
        if ix.min() < 0 or ix.max() >= self.n:
            raise N.PsbError(N.INVALID, "sctb", "index outside 1..n")
        for k in range(n):                           # in order: repeated indices accumulate
            self.h[ix[k]] = beta * self.h[ix[k]] + x[k]
        self.set_host()

    def sctb_x(self, i, n, idx, x, beta):
        """d_cuda_sctb_x (:401-504), pass(y): y(idx(i+k-1)) = beta y(..) + x(k)
        on the device, x a host array."""
        didx = self._index_window("sctb_x", i, n, idx)
        if n == 0:
            return
        self._stage(n)
        a = np.ascontiguousarray(x[:n], dtype=self._dtype)
        N.call("psb_hip_write_buffer", self.dt_buf, a.ctypes.data, n * a.itemsize)
        N.call("psb_hip_mv_iscat", self.dev, 0, n, i, didx, 1, self.dt_buf, 1, float(beta))
        self.set_dev()

    def sctb_buf(self, i, n, idx, beta):
        """d_cuda_sctb_buf (:506-555), pass(y): y(idx(i:i+n-1)) = beta y(..) +
        combuf(i:i+n-1) on the device."""
        if self.combuf is None or i + n - 1 > len(self.combuf):
            raise N.PsbError(N.INVALID, "sctb_buf", "no communication buffer or window beyond it")
        didx = self._index_window("sctb_buf", i, n, idx)
        if n == 0:
            return
        self._stage(i + n - 1)
        isz = np.dtype(self._dtype).itemsize
        a = np.ascontiguousarray(self.combuf[i - 1:i + n - 1])
        N.call("psb_hip_write_buffer", C.c_void_p(self.dt_buf.value + (i - 1) * isz),
               a.ctypes.data, n * isz)
        N.call("psb_hip_mv_iscat", self.dev, 0, n, i, didx, i, self.dt_buf, 1, float(beta))
        self.set_dev()


class psb_d_hip_vect(psb_hip_base_vect):
    """Double-precision device vector (psb_d_vect_cuda)."""
    _etype = N.TYPE_DOUBLE
    _dtype = np.float64
    _setscal = "psb_hip_mv_setscal_d"


class psb_s_hip_vect(psb_hip_base_vect):
    """Single-precision device vector (psb_s_vect_cuda)."""
    _etype = N.TYPE_FLOAT
    _dtype = np.float32
    _setscal = "psb_hip_mv_setscal_s"


class psb_i_hip_vect(psb_hip_base_vect):
    """Index vector (psb_i_vect_cuda), used for halo and insertion indices."""
    _etype = N.TYPE_INT
    _dtype = np.int32
    _setscal = "psb_hip_mv_setscal_i"


class psb_d_base_vect:
    """Host-only vector (psb_d_base_vect_type, base/modules/serial/
    psb_d_base_vect_mod.F90:63-150): values in host memory, always in the host
    state.  A product of a device matrix with such a vector runs on the
    matrix's host copy (cuda/impl/psb_d_cuda_hlg_vect_mv.F90:77-108)."""
    _dtype = np.float64
    fmt = "BASE"

    def __init__(self):
        self.h = None
        self.n = 0
        self.count = 1

    def all(self, n, count=1):
        self.h = np.zeros((n,) if count == 1 else (n, count), dtype=self._dtype, order="F")
        self.n, self.count = n, count
        return self

    def bld(self, values):
        a = np.asarray(values, dtype=self._dtype)
        self.all(a.shape[0], 1 if a.ndim == 1 else a.shape[1])
        self.h[...] = a
        return self

    def free(self):
        self.h = None
        self.n = 0

    def mold(self):
        return type(self)()

    def is_host(self):
        return True

    def is_dev(self):
        return False

    def is_sync(self):
        return False

    def set_host(self):
        pass

    def sync(self):
        pass

    def get_vect(self):
        return self.h.copy()

    def get_nrows(self):
        return self.n

    def zero(self):
        self.h[...] = 0

    def set_scal(self, val, first=None, last=None):
        lo = 0 if first is None else first - 1
        hi = self.n if last is None else last
        self.h[lo:hi] = val


class psb_s_base_vect(psb_d_base_vect):
    """Single-precision host-only vector (psb_s_base_vect_type)."""
    _dtype = np.float32


def _same_kind(*vs):
    t = type(vs[0])
    for v in vs[1:]:
        if type(v) is not t:
            raise N.PsbError(N.INVALID, "vect", "operands of different vector types "
                             f"({type(vs[0]).__name__} vs {type(v).__name__}); no host fallback")


class psb_d_vect_type:
    """psb_d_vect_type: wrapper whose ``v`` is the storage plug-in."""

    def __init__(self, v=None):
        self.v = v

    def get_nrows(self):
        return self.v.get_nrows()

    def get_vect(self):
        return self.v.get_vect()

    def cnv(self, mold=None):
        """d_vect_cnv (base/modules/serial/psb_d_vect_mod.F90:590-612): move
        the values into storage of the mold's type."""
        nv = (mold or psb_d_hip_vect()).mold()
        if self.v is not None:
            nv.bld(self.v.get_vect().astype(nv._dtype))
            self.v.free()
        self.v = nv
        return self

    def set(self, val):
        self.v.set_scal(val)

    def zero(self):
        self.v.zero()

    def sync(self):
        self.v.sync()

    def free(self):
        if self.v is not None:
            self.v.free()

    @property
    def dtype(self):
        return self.v._dtype


# psb_s_vect_type shares the wrapper; the plug-in decides the precision
psb_s_vect_type = psb_d_vect_type
