"""Device memory, streams and events over the library's own HIP runtime
(include/ntt_amd_runtime.h), so that no second HIP runtime copy is needed in
the process."""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib, NTTError, NTT_OK

_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I = ctypes.c_int


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 64), ("compute_units", ctypes.c_int),
                ("clock_khz", ctypes.c_int), ("l2_bytes", ctypes.c_int), ("lds_per_block", ctypes.c_size_t),
                ("total_global_mem", ctypes.c_size_t)]


_PROTOS = {
    "ntt_rt_device_count": [ctypes.POINTER(_I)],
    "ntt_rt_set_device": [_I],
    "ntt_rt_device_info": [_I, ctypes.POINTER(DeviceInfo)],
    "ntt_rt_malloc": [ctypes.POINTER(_P), _SZ],
    "ntt_rt_free": [_P],
    "ntt_rt_memcpy_htod": [_P, _P, _SZ],
    "ntt_rt_memcpy_dtoh": [_P, _P, _SZ],
    "ntt_rt_memcpy_dtod": [_P, _P, _SZ, _P],
    "ntt_rt_memset": [_P, _I, _SZ],
    "ntt_rt_synchronize": [],
    "ntt_rt_mem_info": [ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)],
    "ntt_rt_stream_create": [ctypes.POINTER(_P)],
    "ntt_rt_stream_destroy": [_P],
    "ntt_rt_stream_synchronize": [_P],
    "ntt_rt_event_create": [ctypes.POINTER(_P)],
    "ntt_rt_event_destroy": [_P],
    "ntt_rt_event_record": [_P, _P],
    "ntt_rt_event_synchronize": [_P],
    "ntt_rt_event_elapsed_ms": [ctypes.POINTER(ctypes.c_float), _P, _P],
    "ntt_rt_fill_uniform": [_P, _P, _SZ, _I, _SZ, ctypes.c_uint64, _P],
    "ntt_rt_butterfly_roof": [_I, ctypes.POINTER(ctypes.c_double), _P],
    "ntt_rt_stream_copy": [_P, _P, _SZ, _P],
    "ntt_rt_graph_begin": [_P],
    "ntt_rt_graph_end": [_P, ctypes.POINTER(_P)],
    "ntt_rt_graph_launch": [_P, _P],
    "ntt_rt_graph_destroy": [_P],
}
for _n, _a in _PROTOS.items():
    _f = getattr(_lib, _n)
    _f.restype = _I
    _f.argtypes = _a


def _ck(rc, what):
    if rc != NTT_OK:
        raise NTTError(rc, what)


def device_count() -> int:
    n = _I(0)
    rc = _lib.ntt_rt_device_count(ctypes.byref(n))
    return n.value if rc == NTT_OK else 0


def is_available() -> bool:
    return device_count() > 0


def device_info(device: int = 0) -> dict:
    d = DeviceInfo()
    _ck(_lib.ntt_rt_device_info(device, ctypes.byref(d)), "device_info")
    return {"name": d.name.decode(), "arch": d.arch.decode(), "compute_units": d.compute_units,
            "clock_khz": d.clock_khz, "l2_bytes": d.l2_bytes, "lds_per_block": d.lds_per_block,
            "total_global_mem": d.total_global_mem}


def synchronize():
    _ck(_lib.ntt_rt_synchronize(), "synchronize")


def mem_info():
    f, t = _SZ(0), _SZ(0)
    _ck(_lib.ntt_rt_mem_info(ctypes.byref(f), ctypes.byref(t)), "mem_info")
    return f.value, t.value


class DeviceBuffer:
    """A device allocation of `nbytes` bytes."""

    def __init__(self, nbytes: int):
        p = _P()
        _ck(_lib.ntt_rt_malloc(ctypes.byref(p), max(int(nbytes), 4)), "malloc")
        self.ptr = int(p.value)
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _ck(_lib.ntt_rt_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes), "memcpy_htod")
        return b

    def upload(self, a: np.ndarray, offset_bytes: int = 0):
        a = np.ascontiguousarray(a)
        assert offset_bytes + a.nbytes <= self.nbytes
        _ck(_lib.ntt_rt_memcpy_htod(self.ptr + offset_bytes, a.ctypes.data, a.nbytes), "memcpy_htod")

    def numpy(self, dtype=np.uint32, shape=None, offset_bytes: int = 0, nbytes: Optional[int] = None) -> np.ndarray:
        nb = self.nbytes - offset_bytes if nbytes is None else nbytes
        out = np.empty(nb // np.dtype(dtype).itemsize, dtype=dtype)
        _ck(_lib.ntt_rt_memcpy_dtoh(out.ctypes.data, self.ptr + offset_bytes, out.nbytes), "memcpy_dtoh")
        return out.reshape(shape) if shape is not None else out

    def copy_from(self, other: "DeviceBuffer", nbytes: Optional[int] = None, stream=None):
        nb = min(self.nbytes, other.nbytes) if nbytes is None else nbytes
        _ck(_lib.ntt_rt_memcpy_dtod(self.ptr, other.ptr, nb, _sh(stream)), "memcpy_dtod")

    def zero(self):
        _ck(_lib.ntt_rt_memset(self.ptr, 0, self.nbytes), "memset")

    def free(self):
        if self.ptr:
            _lib.ntt_rt_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _sh(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream or None
    return stream.handle or None


class Stream:
    def __init__(self):
        p = _P()
        _ck(_lib.ntt_rt_stream_create(ctypes.byref(p)), "stream_create")
        self.handle = int(p.value)

    def synchronize(self):
        _ck(_lib.ntt_rt_stream_synchronize(self.handle), "stream_synchronize")

    def __del__(self):
        try:
            if self.handle:
                _lib.ntt_rt_stream_destroy(self.handle)
                self.handle = 0
        except Exception:
            pass


class Event:
    def __init__(self):
        p = _P()
        _ck(_lib.ntt_rt_event_create(ctypes.byref(p)), "event_create")
        self.handle = int(p.value)

    def record(self, stream=None):
        _ck(_lib.ntt_rt_event_record(self.handle, _sh(stream)), "event_record")

    def synchronize(self):
        _ck(_lib.ntt_rt_event_synchronize(self.handle), "event_synchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        _ck(_lib.ntt_rt_event_elapsed_ms(ctypes.byref(ms), self.handle, end.handle), "event_elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _lib.ntt_rt_event_destroy(self.handle)
                self.handle = 0
        except Exception:
            pass


def fill_uniform(buf, moduli_dev, limbs: int, logN: int, count: int, seed: int, stream=None):
    """Uniform residues per limb, generated on the device (synthetic operands)."""
    _ck(_lib.ntt_rt_fill_uniform(buf.ptr if hasattr(buf, "ptr") else buf,
                                 moduli_dev.ptr if hasattr(moduli_dev, "ptr") else moduli_dev,
                                 limbs, logN, count, seed & ((1 << 64) - 1), _sh(stream)), "fill_uniform")


BUTTERFLY_CT, BUTTERFLY_CT_LAZY, BUTTERFLY_GS, BUTTERFLY_GS_LAZY = 0, 1, 2, 3
BUTTERFLY_CT64, BUTTERFLY_GS64 = 4, 5
BUTTERFLY_CT_LAZY_SHOUP, BUTTERFLY_CT_LAZY_MONT8 = 6, 7


def butterfly_roof(variant: int = BUTTERFLY_CT, stream=None) -> float:
    """Butterflies/s of the kernels' own butterfly code on register-resident
    operands: the compute roof of the transform kernels."""
    v = ctypes.c_double(0)
    _ck(_lib.ntt_rt_butterfly_roof(variant, ctypes.byref(v), _sh(stream)), "butterfly_roof")
    return v.value


def stream_copy(dst, src, nbytes: int, stream=None):
    """Plain streaming copy (each byte read once, written once)."""
    _ck(_lib.ntt_rt_stream_copy(dst.ptr if hasattr(dst, "ptr") else dst, src.ptr if hasattr(src, "ptr") else src,
                                nbytes, _sh(stream)), "stream_copy")


class Graph:
    """A captured sequence of library calls on a stream (hipGraph), replayable.

        with Graph.capture(stream) as g:     # calls enqueued here are recorded
            plan.forward(buf, 1, stream)
        g.launch(stream)
    """

    def __init__(self):
        self.handle = None
        self._stream = None

    @classmethod
    def capture(cls, stream):
        g = cls()
        g._stream = stream
        return g

    def __enter__(self):
        _ck(_lib.ntt_rt_graph_begin(_sh(self._stream)), "graph_begin")
        return self

    def __exit__(self, exc_type, exc, tb):
        h = ctypes.c_void_p()
        rc = _lib.ntt_rt_graph_end(_sh(self._stream), ctypes.byref(h))
        if exc_type is None:
            _ck(rc, "graph_end")
            self.handle = h.value
        return False

    def launch(self, stream=None):
        _ck(_lib.ntt_rt_graph_launch(self.handle, _sh(stream)), "graph_launch")

    def close(self):
        if self.handle:
            _lib.ntt_rt_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
