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
