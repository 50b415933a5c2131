"""Environment: the counterpart of psb_init/psb_exit plus psb_cuda_init and
the device queries of cuda/psb_cuda_env_mod.F90:43-300."""
import ctypes as C
import time

from . import _native as N


class psb_ctxt_type:
    """One process, one device (np = 1): the context of psb_init
    (base/modules/penv/psi_penv_mod.F90:811)."""

    def __init__(self, device=0):
        self.device = device
        self.me = 0
        self.np = 1


def psb_init(device=0):
    """psb_init + psb_cuda_init(ctxt, dev): select the device, create the
    library stream and reduction workspace."""
    N.call("psb_hip_init", device)
    return psb_ctxt_type(device)


def psb_exit(ctxt=None):
    N.call("psb_hip_exit")


def psb_info(ctxt):
    return ctxt.me, ctxt.np


def psb_wtime():
    """psb_wtime (psi_penv_mod.F90:1044): wall clock in seconds."""
    return time.perf_counter()


def psb_hip_DeviceSync():
    """psb_cuda_DeviceSync: wait for all queued device work."""
    N.call("psb_hip_device_sync")


def psb_hip_getDeviceCount():
    return N.fn("psb_hip_get_device_count")()


def psb_hip_MultiProcessors():
    return N.fn("psb_hip_get_multiprocessors")()


def psb_hip_MemoryBusWidth():
    return N.fn("psb_hip_get_memory_bus_width")()


def psb_hip_MemoryClockRate():
    return N.fn("psb_hip_get_memory_clock_rate")()


def psb_hip_WarpSize():
    return N.fn("psb_hip_get_warp_size")()


def psb_hip_MemoryPeakBandwidth():
    """psb_cuda_MemoryPeakBandwidth: 2*clock*busWidth/8, MB/s."""
    return N.fn("psb_hip_memory_peak_bandwidth")()


def psb_hip_DeviceName():
    buf = C.create_string_buffer(256)
    N.call("psb_hip_get_device_name", buf, 256)
    return buf.value.decode()


class DeviceTimer:
    """HIP-event timing on the library stream (milliseconds)."""

    def start(self):
        N.call("psb_hip_timer_start")

    def stop_ms(self):
        return N.fn("psb_hip_timer_stop")()
