"""Minimal HIP runtime access through ctypes: events on a given stream,
device memory and synchronisation (the library's kernels run on its own
streams, which torch's current-stream events would not see), and the counter
calibration copy."""
from __future__ import annotations

import ctypes as ct
import os


class Hip:
    """HIP events through the runtime's C API (the kernels run on the
    context's own stream, so torch's current-stream events would not see them)."""

    def __init__(self):
        self.h = ct.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        self.h.hipEventDestroy.argtypes = [ct.c_void_p]

    def event(self):
        e = ct.c_void_p()
        assert self.h.hipEventCreate(ct.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ct.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b) -> float:
        assert self.h.hipEventSynchronize(b) == 0
        f = ct.c_float()
        assert self.h.hipEventElapsedTime(ct.byref(f), a, b) == 0
        return f.value

    def device_sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def malloc(self, nbytes):
        p = ct.c_void_p()
        assert self.h.hipMalloc(ct.byref(p), ct.c_size_t(nbytes)) == 0
        return p

    def free(self, p):
        assert self.h.hipFree(p) == 0



def copy_probe(hip: Hip, nbytes: int) -> None:
    """One plain copy kernel of known size (read nbytes + write nbytes): the
    calibration point for the FETCH_SIZE / WRITE_SIZE counters."""
    from . import _lib

    lib = _lib.load()
    a, b = hip.malloc(nbytes), hip.malloc(nbytes)
    assert lib.gpe_bench_copy(b, a, nbytes, None) == 0
    hip.device_sync()
    hip.free(a)
    hip.free(b)
