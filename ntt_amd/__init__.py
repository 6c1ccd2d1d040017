"""ntt_amd — MI355X (gfx950) RNS multi-limb negacyclic NTT/INTT.

Python mirror of the reference's host API over the library's C ABI
(include/ntt_amd.h).  Names follow the reference:

  ModulusConfig, PhaseConfig, TransformConfig  src/ntt_config.h:6-23
  default_moduli()                              src/ntt_config.h:75-94
  default_launch_config(logN)                   DefaultNTTLaunchConfig<LogN>, src/ntt_config.h:62-73
  launch_ntt / launch_intt                      src/v04_montgomery.cu:603-658 (Montgomery)
  launch_ntt_plain / launch_intt_plain          src/v03_template_config.cu:549-603
  make_modulus_runtime, make_twiddle_table      src/v04_montgomery.cu:681-727

Device buffers are ntt_amd.device.DeviceBuffer objects (allocated through the
library's own HIP runtime, include/ntt_amd_runtime.h) or raw integer device
addresses.  The HIP library is required: importing this package raises if
libntt_amd.so is missing.
"""
from __future__ import annotations

import ctypes
import os
import re
from dataclasses import dataclass
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libntt_amd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "ntt_amd.h")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"ntt_amd: HIP library not built ({LIB_PATH} missing); run `python ntt_amd/build.py`")

_lib = ctypes.CDLL(LIB_PATH)

NTT_OK = 0
STATUS = {
    0: "NTT_OK", 1: "NTT_ERR_INVALID_ARGUMENT", 2: "NTT_ERR_MODULUS_CONGRUENCE",
    3: "NTT_ERR_PRIMITIVE_ROOT", 4: "NTT_ERR_MODULUS_RANGE", 5: "NTT_ERR_CONFIG",
    6: "NTT_ERR_HIP", 7: "NTT_ERR_OUT_OF_MEMORY", 8: "NTT_ERR_BUDGET", 9: "NTT_ERR_NOT_FOUND",
    10: "NTT_ERR_BUSY", 11: "NTT_ERR_NO_DEVICE", 12: "NTT_ERR_SCRATCH_LIMIT",
}
PLAIN, MONTGOMERY = 0, 1
FORWARD, INVERSE = 0, 1
ENGINE_AUTO, ENGINE_TWO_PASS, ENGINE_FUSED = 0, 1, 2
TWIDDLE_AUTO, TWIDDLE_TABLE, TWIDDLE_OT = -1, 0, 1


class NTTError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = _lib.ntt_status_string(status).decode()
        if status == 6:
            msg += " — " + (_lib.ntt_last_hip_error() or b"").decode()
        super().__init__(f"{what}: {STATUS.get(status, status)} ({msg})" if what else msg)


# ---- C structs --------------------------------------------------------------
class ModulusConfig(ctypes.Structure):
    """ModulusConfig{qi, primitive_root}, src/ntt_config.h:6-9."""
    _fields_ = [("qi", ctypes.c_uint32), ("primitive_root", ctypes.c_uint32)]

    def __repr__(self):
        return f"ModulusConfig({self.qi}, {self.primitive_root})"

    def __iter__(self):
        yield self.qi
        yield self.primitive_root


class ModulusConfig64(ctypes.Structure):
    """ntt_modulus_config64{q, root}: a prime q < 2^62 and a quadratic non-residue."""
    _fields_ = [("q", ctypes.c_uint64), ("root", ctypes.c_uint64)]

    def __repr__(self):
        return f"ModulusConfig64({self.q}, {self.root})"

    def __iter__(self):
        yield self.q
        yield self.root


class PhaseConfig(ctypes.Structure):
    """PhaseConfig{radix_stages, stage_merging, warp_batching}, src/ntt_config.h:12-16."""
    _fields_ = [("radix_stages", ctypes.c_int32), ("stage_merging", ctypes.c_int32),
                ("warp_batching", ctypes.c_int32)]

    def astuple(self):
        return (self.radix_stages, self.stage_merging, self.warp_batching)

    def __repr__(self):
        return f"PhaseConfig{self.astuple()}"


class TransformConfig(ctypes.Structure):
    """TransformConfig, src/ntt_config.h:18-23."""
    _fields_ = [("ntt_phase1", PhaseConfig), ("ntt_phase2", PhaseConfig),
                ("intt_phase1", PhaseConfig), ("intt_phase2", PhaseConfig)]

    @classmethod
    def make(cls, p1, p2, ip1=None, ip2=None):
        ip1 = p2 if ip1 is None else ip1
        ip2 = p1 if ip2 is None else ip2
        return cls(PhaseConfig(*p1), PhaseConfig(*p2), PhaseConfig(*ip1), PhaseConfig(*ip2))

    def astuple(self):
        return (self.ntt_phase1.astuple(), self.ntt_phase2.astuple(),
                self.intt_phase1.astuple(), self.intt_phase2.astuple())

    def __repr__(self):
        return f"TransformConfig{self.astuple()}"


class ModulusRuntime(ctypes.Structure):
    """ModulusRuntime, src/v04_montgomery.cu:17-26."""
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("qi", "primitive_root", "q_inv", "r2", "psi", "psi_inv", "omega", "degree_inv")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class TuneResult(ctypes.Structure):
    """ntt_tune_result (include/ntt_amd.h)."""
    _fields_ = [("cfg", TransformConfig), ("chunk", ctypes.c_size_t), ("fwd_ms", ctypes.c_double),
                ("inv_ms", ctypes.c_double), ("default_fwd_ms", ctypes.c_double),
                ("default_inv_ms", ctypes.c_double), ("candidates", ctypes.c_int), ("engine", ctypes.c_int),
                ("fused_fwd_ms", ctypes.c_double), ("fused_inv_ms", ctypes.c_double), ("fused_stagger", ctypes.c_int),
                ("fused_eb", ctypes.c_int), ("fused_eb_inv", ctypes.c_int), ("fused_lazy", ctypes.c_int),
                ("fused_lazy_inv", ctypes.c_int), ("fused_words_log2", ctypes.c_int),
                ("fused_words_log2_inv", ctypes.c_int), ("fused_ot", ctypes.c_int), ("fused_ot_inv", ctypes.c_int),
                ("engine_inv", ctypes.c_int)]

    def as_dict(self):
        return {"cfg": self.cfg.astuple(), "chunk": int(self.chunk), "fwd_ms": self.fwd_ms, "inv_ms": self.inv_ms,
                "default_fwd_ms": self.default_fwd_ms, "default_inv_ms": self.default_inv_ms,
                "candidates": self.candidates, "engine": {1: "two_pass", 2: "fused"}.get(self.engine, self.engine),
                "fused_fwd_ms": self.fused_fwd_ms, "fused_inv_ms": self.fused_inv_ms,
                "fused_stagger": self.fused_stagger, "fused_eb": self.fused_eb, "fused_eb_inv": self.fused_eb_inv,
                "fused_lazy": self.fused_lazy, "fused_lazy_inv": self.fused_lazy_inv,
                "fused_words_log2": self.fused_words_log2, "fused_words_log2_inv": self.fused_words_log2_inv,
                "fused_ot": self.fused_ot, "fused_ot_inv": self.fused_ot_inv,
                "engine_inv": {1: "two_pass", 2: "fused"}.get(self.engine_inv, self.engine_inv)}


class BatchJob(ctypes.Structure):
    """ntt_batch_job of include/ntt_amd.h."""
    _fields_ = [("data", ctypes.c_void_p), ("moduli", ctypes.c_void_p), ("limbs", ctypes.c_uint32),
                ("logN", ctypes.c_int32), ("direction", ctypes.c_int32), ("domain", ctypes.c_int32)]


class TableJob(ctypes.Structure):
    """ntt_table_job: a job of ntt_batch_run_static (device pointers)."""
    _fields_ = [("data", ctypes.c_void_p), ("twiddles", ctypes.c_void_p), ("moduli", ctypes.c_void_p),
                ("q_inv", ctypes.c_void_p), ("scale", ctypes.c_void_p), ("limbs", ctypes.c_uint32),
                ("logN", ctypes.c_int32), ("direction", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("slots", ctypes.c_int), ("max_jobs_per_launch", ctypes.c_int)]


class Request(ctypes.Structure):
    """ntt_request of include/ntt_amd.h."""
    _fields_ = [("logN", ctypes.c_int32), ("limbs", ctypes.c_uint32), ("direction", ctypes.c_int32),
                ("domain", ctypes.c_int32), ("moduli", ctypes.c_void_p), ("src_host", ctypes.c_void_p),
                ("src_dev", ctypes.c_void_p), ("seed", ctypes.c_uint64), ("auto_release", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("fused_launches", ctypes.c_uint64), ("two_pass_launches", ctypes.c_uint64),
                ("mean_jobs_per_launch", ctypes.c_double), ("arena_bytes", ctypes.c_size_t),
                ("arena_peak_bytes", ctypes.c_size_t), ("max_slots_busy", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


REQ_QUEUED, REQ_ADMITTED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_UNKNOWN = range(6)


@dataclass
class VersionConfig:
    """VersionConfig{logN, moduli, transform}, src/ntt_config.h:25-29."""
    logN: int
    moduli: List[ModulusConfig]
    transform: TransformConfig


# ---- prototypes ---------------------------------------------------------------
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_SZ = ctypes.c_size_t
_I = ctypes.c_int
_PU32 = ctypes.POINTER(ctypes.c_uint32)
_PCFG = ctypes.POINTER(TransformConfig)

_PROTOS = {
    "ntt_status_string": (ctypes.c_char_p, [_I]),
    "ntt_last_hip_error": (ctypes.c_char_p, []),
    "ntt_version": (ctypes.c_char_p, []),
    "ntt_max_log_n": (_I, []),
    "ntt_default_moduli": (_I, [ctypes.POINTER(ModulusConfig), _I]),
    "ntt_default_transform_config": (_I, [_I, _PCFG]),
    "ntt_reference_transform_config": (_I, [_I, _PCFG]),
    "ntt_validate_transform_config": (_I, [_I, _PCFG]),
    "ntt_generate_moduli": (_I, [_I, _U32, _I, ctypes.POINTER(ModulusConfig)]),
    "ntt_make_modulus_runtime": (_I, [_I, ModulusConfig, ctypes.POINTER(ModulusRuntime)]),
    "ntt_make_twiddle_table": (_I, [_I, _U32, _U32, _PU32]),
    "ntt_make_device_tables_host": (_I, [_I, ctypes.POINTER(ModulusConfig), _SZ, _I, _PU32, _PU32, _PU32,
                                         _PU32, _PU32, _PU32]),
    "ntt_mont_q_inv": (_U32, [_U32]),
    "ntt_to_mont": (_U32, [_U32, _U32]),
    "ntt_mont_r2": (_U32, [_U32]),
    "ntt_from_mont": (_U32, [_U32, _U32, _U32]),
    "ntt_launch_ntt_montgomery": (_I, [_P, _P, _P, _P, _P, _SZ, _I, _PCFG, _P]),
    "ntt_launch_intt_montgomery": (_I, [_P, _P, _P, _P, _P, _SZ, _I, _PCFG, _P]),
    "ntt_launch_ntt_plain": (_I, [_P, _P, _P, _SZ, _I, _PCFG, _P]),
    "ntt_launch_intt_plain": (_I, [_P, _P, _P, _P, _SZ, _I, _PCFG, _P]),
    "ntt_context_create": (_I, [_I, ctypes.POINTER(_P)]),
    "ntt_context_destroy": (_I, [_P]),
    "ntt_context_twiddle_bytes": (_SZ, [_P]),
    "ntt_context_twiddle_entries": (_SZ, [_P]),
    "ntt_plan_create": (_I, [_P, _I, ctypes.POINTER(ModulusConfig), _SZ, _I, _PCFG, ctypes.POINTER(_P)]),
    "ntt_plan_destroy": (_I, [_P]),
    "ntt_plan_get_config": (_I, [_P, _PCFG]),
    "ntt_plan_forward": (_I, [_P, _P, _P]),
    "ntt_plan_inverse": (_I, [_P, _P, _P]),
    "ntt_plan_forward_many": (_I, [_P, _P, _SZ, _P]),
    "ntt_plan_inverse_many": (_I, [_P, _P, _SZ, _P]),
    "ntt_generate_moduli64": (_I, [_I, ctypes.c_uint64, _I, ctypes.POINTER(ModulusConfig64)]),
    "ntt_plan64_create": (_I, [_P, _I, ctypes.POINTER(ModulusConfig64), _SZ, ctypes.POINTER(_P)]),
    "ntt_plan64_destroy": (_I, [_P]),
    "ntt_plan64_forward": (_I, [_P, _P, _SZ, _P]),
    "ntt_plan64_inverse": (_I, [_P, _P, _SZ, _P]),
    "ntt_plan_forward_to": (_I, [_P, _P, _P, _SZ, _P]),
    "ntt_plan_inverse_to": (_I, [_P, _P, _P, _SZ, _P]),
    "ntt_plan_set_mod_down_factors": (_I, [_P, _PU32]),
    "ntt_plan_forward_mod_down": (_I, [_P, _P, _P, _P, _SZ, _P]),
    "ntt_plan_set_mod_down_sub": (_I, [_P, _U32, _U32, _PU32]),
    "ntt_plan_forward_skip": (_I, [_P, _P, _P, _U32, _U32, _SZ, _P]),
    "ntt_plan_inverse_from": (_I, [_P, _P, _P, _U32, _U32, _SZ, _P]),
    "ntt_plan_run_pass": (_I, [_P, _P, _SZ, _I, _I, _P]),
    "ntt_plan_set_engine": (_I, [_P, _I]),
    "ntt_plan_set_fused_shape": (_I, [_P, _I]),
    "ntt_plan_set_fused_words": (_I, [_P, _I]),
    "ntt_plan_set_fused_arith": (_I, [_P, _I]),
    "ntt_plan_set_twiddle_mode": (_I, [_P, _I]),
    "ntt_plan_set_constants": (_I, [_P, _PU32, _PU32]),
    "ntt_plan_get_fused_variant": (_I, [_P, _SZ, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I),
                                        ctypes.POINTER(_I)]),
    "ntt_plan_get_engine": (_I, [_P, _SZ]),
    "ntt_plan_get_engine_for": (_I, [_P, _SZ, _I]),
    "ntt_autotune": (_I, [_P, _I, ctypes.POINTER(ModulusConfig), _SZ, _SZ, _I, _I, ctypes.POINTER(TuneResult)]),
    "ntt_context_set_tuning": (_I, [_P, _I, _SZ, _PCFG, _SZ, _I]),
    "ntt_context_clear_tuning": (_I, [_P]),
    "ntt_context_save_tuning": (_I, [_P, ctypes.c_char_p]),
    "ntt_context_load_tuning": (_I, [_P, ctypes.c_char_p]),
    "ntt_batch_create": (_I, [_P, _P, _SZ, ctypes.POINTER(_P)]),
    "ntt_batch_destroy": (_I, [_P]),
    "ntt_batch_run": (_I, [_P, _P]),
    "ntt_batch_capture": (_I, [_P, _P]),
    "ntt_batch_replay": (_I, [_P, _P]),
    "ntt_batch_set_data": (_I, [_P, _SZ, _P]),
    "ntt_batch_info": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "ntt_batch_run_static": (_I, [ctypes.POINTER(TableJob), _SZ, _P]),
    "ntt_server_create": (_I, [_P, _P, ctypes.POINTER(_P)]),
    "ntt_server_destroy": (_I, [_P]),
    "ntt_server_submit": (_I, [_P, _P, ctypes.POINTER(ctypes.c_uint64)]),
    "ntt_server_state": (_I, [_P, ctypes.c_uint64, ctypes.POINTER(_I)]),
    "ntt_server_wait": (_I, [_P, ctypes.c_uint64]),
    "ntt_server_fetch": (_I, [_P, ctypes.c_uint64, _P, _I, _I]),
    "ntt_server_release": (_I, [_P, ctypes.c_uint64]),
    "ntt_server_stats_get": (_I, [_P, _P]),
    "ntt_server_prepare": (_I, [_P, _I, ctypes.c_uint32, _I]),
    "ntt_server_run_trace": (_I, [_P, _SZ, _P, _P, _P, _P, _P, _I, _P, ctypes.POINTER(ctypes.c_double)]),
}


def _bind(lib):
    for name, (res, args) in _PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


_bind(_lib)

from . import device  # noqa: E402  (binds include/ntt_amd_runtime.h)
from .device import DeviceBuffer, Event, Stream, synchronize  # noqa: E402,F401


def lib():
    return _lib


def header_functions(path: str = HEADER) -> List[str]:
    """Names of every function declared in include/*.h."""
    names = []
    inc = os.path.dirname(path)
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"\b(ntt_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def _check(rc: int, what: str):
    if rc != NTT_OK:
        raise NTTError(rc, what)


def _addr(x) -> Optional[int]:
    """Device address of a DeviceBuffer, an object with data_ptr(), or an int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "ptr"):
        return int(x.ptr)
    return int(x.data_ptr())


def _stream(stream) -> Optional[int]:
    """hipStream_t as an int (None / 0 = the null stream)."""
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream or None
    return int(stream.handle) or None


def _cfg_ptr(cfg: Optional[TransformConfig]):
    return ctypes.byref(cfg) if cfg is not None else None


# ---- configuration API ---------------------------------------------------------
def default_moduli() -> List[ModulusConfig]:
    n = _lib.ntt_default_moduli(None, 0)
    arr = (ModulusConfig * n)()
    _lib.ntt_default_moduli(arr, n)
    return list(arr)


def default_launch_config(logN: int) -> TransformConfig:
    """DefaultNTTLaunchConfig<LogN> with warp_batching re-derived for wave64."""
    c = TransformConfig()
    _check(_lib.ntt_default_transform_config(logN, ctypes.byref(c)), "default_launch_config")
    return c


def reference_launch_config(logN: int) -> TransformConfig:
    """The reference's literal DefaultNTTLaunchConfig<LogN> values."""
    c = TransformConfig()
    _check(_lib.ntt_reference_transform_config(logN, ctypes.byref(c)), "reference_launch_config")
    return c


def validate_config(logN: int, cfg: TransformConfig) -> int:
    return _lib.ntt_validate_transform_config(logN, ctypes.byref(cfg))


def make_version_config(logN: int = 16, moduli=None, transform=None) -> VersionConfig:
    """make_version_config<LC>(), src/ntt_config.h:96-103."""
    return VersionConfig(logN, list(moduli) if moduli is not None else default_moduli(),
                         transform if transform is not None else default_launch_config(logN))


def default_ntt_config() -> VersionConfig:
    """default_ntt_config(), src/ntt_config.h:105-107."""
    return make_version_config(16)


def generate_moduli(logN: int, count: int, bound: int = 1 << 30) -> List[ModulusConfig]:
    """The `count` largest primes q < bound with q = 1 mod 2N, smallest primitive roots."""
    arr = (ModulusConfig * max(count, 1))()
    n = _lib.ntt_generate_moduli(logN, bound, count, arr)
    if n < 0:
        raise NTTError(-n, "generate_moduli")
    return list(arr)[:n]


def generate_moduli64(logN: int, count: int, bound: int = 1 << 62) -> List[ModulusConfig64]:
    """The `count` largest primes q < bound (<= 2^62) with q = 1 mod 2N, each
    with its smallest quadratic non-residue."""
    arr = (ModulusConfig64 * max(count, 1))()
    n = _lib.ntt_generate_moduli64(logN, bound, count, arr)
    if n < 0:
        raise NTTError(-n, "generate_moduli64")
    return list(arr)[:n]


def as_moduli(moduli) -> List[ModulusConfig]:
    return [m if isinstance(m, ModulusConfig) else ModulusConfig(int(m[0]), int(m[1])) for m in moduli]


# ---- host precompute --------------------------------------------------------------
def make_modulus_runtime(logN: int, modulus) -> dict:
    m = as_moduli([modulus])[0]
    rt = ModulusRuntime()
    _check(_lib.ntt_make_modulus_runtime(logN, m, ctypes.byref(rt)), "make_modulus_runtime")
    return rt.as_dict()


def make_twiddle_table(logN: int, q: int, psi: int) -> np.ndarray:
    out = np.empty(1 << logN, dtype=np.uint32)
    _check(_lib.ntt_make_twiddle_table(logN, q, psi, out.ctypes.data_as(_PU32)), "make_twiddle_table")
    return out


@dataclass
class DeviceTables:
    """Host arrays laid out as the reference's main() uploads them (v04:893-915)."""
    twiddles: np.ndarray
    inv_twiddles: np.ndarray
    moduli: np.ndarray
    q_inv: np.ndarray
    r2: np.ndarray
    degree_inv: np.ndarray


def make_tables(logN: int, moduli, domain: int = MONTGOMERY) -> DeviceTables:
    ms = as_moduli(moduli)
    L, N = len(ms), 1 << logN
    arr = (ModulusConfig * L)(*ms)
    t = DeviceTables(*(np.empty(L * N, np.uint32) for _ in range(2)), *(np.empty(L, np.uint32) for _ in range(4)))
    p = lambda a: a.ctypes.data_as(_PU32)
    _check(_lib.ntt_make_device_tables_host(logN, arr, L, domain, p(t.twiddles), p(t.inv_twiddles), p(t.moduli),
                                            p(t.q_inv), p(t.r2), p(t.degree_inv)), "make_tables")
    return t


def mont_q_inv(q): return int(_lib.ntt_mont_q_inv(q))
def to_mont(x, q): return int(_lib.ntt_to_mont(x, q))
def mont_r2(q): return int(_lib.ntt_mont_r2(q))
def from_mont(x, q, q_inv): return int(_lib.ntt_from_mont(x, q, q_inv))


# ---- drop-in launchers (device pointers / tensors) ------------------------------------
def launch_ntt(A, twiddles, moduli, q_inv, r2, limb_count: int, logN: int,
               cfg: Optional[TransformConfig] = None, stream=None):
    """launch_ntt<LC>, src/v04_montgomery.cu:603-626 (Montgomery variant)."""
    _check(_lib.ntt_launch_ntt_montgomery(_addr(A), _addr(twiddles), _addr(moduli), _addr(q_inv), _addr(r2),
                                          limb_count, logN, _cfg_ptr(cfg), _stream(stream)), "launch_ntt")


def launch_intt(A, inv_twiddles, moduli, q_inv, degree_inv, limb_count: int, logN: int,
                cfg: Optional[TransformConfig] = None, stream=None):
    """launch_intt<LC>, src/v04_montgomery.cu:628-658 (Montgomery variant)."""
    _check(_lib.ntt_launch_intt_montgomery(_addr(A), _addr(inv_twiddles), _addr(moduli), _addr(q_inv),
                                           _addr(degree_inv), limb_count, logN, _cfg_ptr(cfg), _stream(stream)),
           "launch_intt")


def launch_ntt_plain(A, twiddles, moduli, limb_count: int, logN: int, cfg=None, stream=None):
    """launch_ntt<LC>, src/v03_template_config.cu:549-571 (plain variant)."""
    _check(_lib.ntt_launch_ntt_plain(_addr(A), _addr(twiddles), _addr(moduli), limb_count, logN,
                                     _cfg_ptr(cfg), _stream(stream)), "launch_ntt_plain")


def launch_intt_plain(A, inv_twiddles, moduli, degree_inv, limb_count: int, logN: int, cfg=None, stream=None):
    """launch_intt<LC>, src/v03_template_config.cu:573-603 (plain variant)."""
    _check(_lib.ntt_launch_intt_plain(_addr(A), _addr(inv_twiddles), _addr(moduli), _addr(degree_inv),
                                      limb_count, logN, _cfg_ptr(cfg), _stream(stream)), "launch_intt_plain")


# ---- context / plans ----------------------------------------------------------------
class Context:
    """Device context owning the (q, logN, direction) twiddle cache."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(_lib.ntt_context_create(device, ctypes.byref(h)), "ntt_context_create")
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            _lib.ntt_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def twiddle_bytes(self) -> int:
        return int(_lib.ntt_context_twiddle_bytes(self.handle))

    @property
    def twiddle_entries(self) -> int:
        return int(_lib.ntt_context_twiddle_entries(self.handle))

    def plan(self, logN: int, moduli, domain: int = MONTGOMERY, cfg: Optional[TransformConfig] = None) -> "Plan":
        return Plan(self, logN, moduli, domain, cfg)

    def plan64(self, logN: int, moduli) -> "Plan64":
        return Plan64(self, logN, moduli)

    def autotune(self, logN: int, moduli, count: int = 1, domain: int = MONTGOMERY, reps: int = 3) -> dict:
        """Measure phase configs for `count` operands of these moduli; the winner
        is used by this context's plans for launches of that size."""
        ms = as_moduli(moduli)
        arr = (ModulusConfig * len(ms))(*ms)
        r = TuneResult()
        _check(_lib.ntt_autotune(self.handle, logN, arr, len(ms), count, domain, reps, ctypes.byref(r)),
               "ntt_autotune")
        return r.as_dict()

    def set_tuning(self, logN: int, total_limbs: int, cfg: TransformConfig, chunk: int = 0,
                   engine: int = ENGINE_TWO_PASS):
        _check(_lib.ntt_context_set_tuning(self.handle, logN, total_limbs, ctypes.byref(cfg), chunk, engine),
               "ntt_context_set_tuning")

    def clear_tuning(self):
        _check(_lib.ntt_context_clear_tuning(self.handle), "ntt_context_clear_tuning")

    def save_tuning(self, path: str):
        _check(_lib.ntt_context_save_tuning(self.handle, path.encode()), "ntt_context_save_tuning")

    def load_tuning(self, path: str):
        _check(_lib.ntt_context_load_tuning(self.handle, path.encode()), "ntt_context_load_tuning")

    def batch(self, jobs) -> "Batch":
        return Batch(self, jobs)

    def server(self, hbm_budget_bytes: int = 64 << 30, slots: int = 4, max_jobs_per_launch: int = 256) -> "Server":
        return Server(self, hbm_budget_bytes, slots, max_jobs_per_launch)


def batch_run_static(jobs, stream=None):
    """ntt_batch_run_static: jobs are (data, twiddles, moduli, q_inv, scale,
    limbs, logN, direction) tuples of device buffers (or addresses) and ints."""
    arr = (TableJob * len(jobs))()
    for k, (data, tw, q, qi, sc, limbs, logN, direction) in enumerate(jobs):
        arr[k] = TableJob(_addr(data), _addr(tw), _addr(q), _addr(qi), _addr(sc), limbs, logN, direction)
    _check(_lib.ntt_batch_run_static(arr, len(jobs), _stream(stream)), "ntt_batch_run_static")


class Server:
    """Request loop with continuous batching (ntt_server_*)."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, slots: int, max_jobs_per_launch: int):
        self.ctx = ctx
        cfg = ServerConfig(hbm_budget_bytes, slots, max_jobs_per_launch)
        h = ctypes.c_void_p()
        _check(_lib.ntt_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)), "ntt_server_create")
        self.handle = h
        self._keep = {}

    def submit(self, logN: int, limbs: int, direction: int = FORWARD, domain: int = MONTGOMERY, moduli=None,
               src_host: Optional[np.ndarray] = None, src_dev=None, seed: int = 0, auto_release: bool = False) -> int:
        r = Request(logN, limbs, direction, domain, None, None, None, seed & ((1 << 64) - 1), int(auto_release))
        keep = []
        if moduli is not None:
            ms = as_moduli(moduli)
            arr = (ModulusConfig * len(ms))(*ms)
            keep.append(arr)
            r.moduli = ctypes.addressof(arr)
        if src_host is not None:
            a = np.ascontiguousarray(src_host, dtype=np.uint32)
            keep.append(a)
            r.src_host = a.ctypes.data
        if src_dev is not None:
            r.src_dev = _addr(src_dev)
        rid = ctypes.c_uint64(0)
        _check(_lib.ntt_server_submit(self.handle, ctypes.byref(r), ctypes.byref(rid)), "ntt_server_submit")
        if keep:
            self._keep[rid.value] = keep      # sources must outlive the request
        return rid.value

    def state(self, rid: int) -> int:
        v = _I(0)
        _check(_lib.ntt_server_state(self.handle, rid, ctypes.byref(v)), "ntt_server_state")
        return v.value

    def wait(self, rid: int):
        _check(_lib.ntt_server_wait(self.handle, rid), "ntt_server_wait")

    def fetch(self, rid: int, logN: int, limbs: int, release: bool = True) -> np.ndarray:
        out = np.empty((limbs, 1 << logN), np.uint32)
        _check(_lib.ntt_server_fetch(self.handle, rid, out.ctypes.data, 0, int(release)), "ntt_server_fetch")
        if release:
            self._keep.pop(rid, None)
        return out

    def fetch_device(self, rid: int, dst, release: bool = True):
        """Copies a finished result into device buffer `dst`."""
        _check(_lib.ntt_server_fetch(self.handle, rid, _addr(dst), 1, int(release)), "ntt_server_fetch")
        if release:
            self._keep.pop(rid, None)

    def release(self, rid: int):
        _check(_lib.ntt_server_release(self.handle, rid), "ntt_server_release")
        self._keep.pop(rid, None)

    def stats(self) -> dict:
        st = ServerStats()
        _check(_lib.ntt_server_stats_get(self.handle, ctypes.byref(st)), "ntt_server_stats_get")
        return st.as_dict()

    def prepare(self, logN: int, limbs: int, domain: int = MONTGOMERY):
        _check(_lib.ntt_server_prepare(self.handle, logN, limbs, domain), "ntt_server_prepare")

    def run_trace(self, requests, domain: int = MONTGOMERY):
        """Replays workloads.Request records (arrival times in seconds from the
        start); returns (latency seconds per request, wall seconds)."""
        n = len(requests)
        arr = np.array([r.arrival_s for r in requests], np.float64)
        ln = np.array([r.job.logN for r in requests], np.int32)
        lb = np.array([r.job.limbs for r in requests], np.uint32)
        dr = np.array([r.job.direction for r in requests], np.int32)
        sd = np.array([r.seed for r in requests], np.uint64)
        lat = np.zeros(n, np.float64)
        wall = ctypes.c_double(0)
        _check(_lib.ntt_server_run_trace(self.handle, n, arr.ctypes.data, ln.ctypes.data, lb.ctypes.data,
                                         dr.ctypes.data, sd.ctypes.data, domain, lat.ctypes.data,
                                         ctypes.byref(wall)), "ntt_server_run_trace")
        return lat, wall.value

    def close(self):
        if self.handle:
            _lib.ntt_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Uneven transforms launched together (ntt_batch_*): jobs are
    (device operand, logN, moduli, direction, domain) tuples."""

    def __init__(self, ctx: Context, jobs):
        self.ctx = ctx
        keep = []
        arr = (BatchJob * len(jobs))()
        for k, (data, logN, moduli, direction, domain) in enumerate(jobs):
            ms = as_moduli(moduli)
            m = (ModulusConfig * len(ms))(*ms)
            keep.append(m)
            arr[k] = BatchJob(_addr(data), ctypes.addressof(m), len(ms), logN, direction, domain)
        h = ctypes.c_void_p()
        _check(_lib.ntt_batch_create(ctx.handle, arr, len(jobs), ctypes.byref(h)), "ntt_batch_create")
        self.handle = h
        self.njobs = len(jobs)

    def run(self, stream=None):
        _check(_lib.ntt_batch_run(self.handle, _stream(stream)), "ntt_batch_run")

    def capture(self, stream=None):
        _check(_lib.ntt_batch_capture(self.handle, _stream(stream)), "ntt_batch_capture")

    def replay(self, stream=None):
        _check(_lib.ntt_batch_replay(self.handle, _stream(stream)), "ntt_batch_replay")

    def set_data(self, index: int, data):
        _check(_lib.ntt_batch_set_data(self.handle, index, _addr(data)), "ntt_batch_set_data")

    def info(self) -> dict:
        a, b, c = _I(0), _I(0), _I(0)
        _check(_lib.ntt_batch_info(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "ntt_batch_info")
        return {"fused_groups": a.value, "two_pass_jobs": b.value, "streams": c.value}

    def close(self):
        if self.handle:
            _lib.ntt_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan64:
    """64-bit word transforms (ntt_plan64_*): normal domain in and out."""

    def __init__(self, ctx, logN: int, moduli):
        ms = [m if isinstance(m, ModulusConfig64) else ModulusConfig64(int(m[0]), int(m[1])) for m in moduli]
        arr = (ModulusConfig64 * len(ms))(*ms)
        h = _P()
        _check(_lib.ntt_plan64_create(ctx.handle, logN, arr, len(ms), ctypes.byref(h)), "ntt_plan64_create")
        self.ctx, self.handle, self.logN, self.limbs = ctx, h, logN, len(ms)

    def forward(self, A, count: int = 1, stream=None):
        _check(_lib.ntt_plan64_forward(self.handle, _addr(A), count, _stream(stream)), "ntt_plan64_forward")

    def inverse(self, A, count: int = 1, stream=None):
        _check(_lib.ntt_plan64_inverse(self.handle, _addr(A), count, _stream(stream)), "ntt_plan64_inverse")

    def close(self):
        if self.handle:
            _lib.ntt_plan64_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    def __init__(self, ctx: Context, logN: int, moduli, domain: int = MONTGOMERY,
                 cfg: Optional[TransformConfig] = None):
        ms = as_moduli(moduli)
        arr = (ModulusConfig * len(ms))(*ms)
        h = ctypes.c_void_p()
        _check(_lib.ntt_plan_create(ctx.handle, logN, arr, len(ms), domain, _cfg_ptr(cfg), ctypes.byref(h)),
               "ntt_plan_create")
        self.ctx, self.handle, self.logN, self.limbs, self.domain = ctx, h, logN, len(ms), domain
        self.moduli = ms

    @property
    def config(self) -> TransformConfig:
        c = TransformConfig()
        _check(_lib.ntt_plan_get_config(self.handle, ctypes.byref(c)), "ntt_plan_get_config")
        return c

    def forward(self, A, count: int = 1, stream=None):
        _check(_lib.ntt_plan_forward_many(self.handle, _addr(A), count, _stream(stream)), "ntt_plan_forward")

    def inverse(self, A, count: int = 1, stream=None):
        _check(_lib.ntt_plan_inverse_many(self.handle, _addr(A), count, _stream(stream)), "ntt_plan_inverse")

    def forward_to(self, dst, src, count: int = 1, stream=None):
        """dst = NTT(src) for `count` operands; src is unchanged unless it is dst."""
        _check(_lib.ntt_plan_forward_to(self.handle, _addr(dst), _addr(src), count, _stream(stream)),
               "ntt_plan_forward_to")

    def inverse_to(self, dst, src, count: int = 1, stream=None):
        """dst = INTT(src) (times the epilogue constants, if set)."""
        _check(_lib.ntt_plan_inverse_to(self.handle, _addr(dst), _addr(src), count, _stream(stream)),
               "ntt_plan_inverse_to")

    def set_mod_down_factors(self, factors=None):
        """Per-limb factors f of forward_mod_down (None releases them)."""
        if factors is None:
            _check(_lib.ntt_plan_set_mod_down_factors(self.handle, None), "ntt_plan_set_mod_down_factors")
            return
        a = (ctypes.c_uint32 * self.limbs)(*[int(x) for x in factors])
        _check(_lib.ntt_plan_set_mod_down_factors(self.handle, ctypes.cast(a, _PU32)),
               "ntt_plan_set_mod_down_factors")

    def forward_mod_down(self, dst, src, sub, count: int = 1, stream=None):
        """dst = (sub - NTT(src)) * f per limb, in the NTT output's order and domain."""
        _check(_lib.ntt_plan_forward_mod_down(self.handle, _addr(dst), _addr(src), _addr(sub), count,
                                              _stream(stream)), "ntt_plan_forward_mod_down")

    def set_mod_down_sub(self, sub_start: int, sub_end: int, factors=None):
        """Mod-down sub operand over limbs [sub_start, sub_end) with optional multipliers."""
        arr = None
        if factors is not None:
            arr = np.ascontiguousarray(np.array([int(f) for f in factors], dtype=np.uint32))
        _check(_lib.ntt_plan_set_mod_down_sub(self.handle, sub_start, sub_end,
                                              arr.ctypes.data_as(_PU32) if arr is not None else None),
               "ntt_plan_set_mod_down_sub")

    def forward_skip(self, dst, src, skip_start: int, skip_end: int, count: int = 1, stream=None):
        _check(_lib.ntt_plan_forward_skip(self.handle, _addr(dst), _addr(src), skip_start, skip_end, count,
                                          _stream(stream)), "ntt_plan_forward_skip")

    def inverse_from(self, dst, src, src_limbs: int, src_first: int, count: int = 1, stream=None):
        _check(_lib.ntt_plan_inverse_from(self.handle, _addr(dst), _addr(src), src_limbs, src_first, count,
                                          _stream(stream)), "ntt_plan_inverse_from")

    def set_engine(self, engine: int):
        _check(_lib.ntt_plan_set_engine(self.handle, engine), "ntt_plan_set_engine")

    def fused_variant(self, count: int = 1, direction: int = FORWARD):
        """(shape, lazy) of the fused kernel a run of `count` operands uses."""
        return self.fused_variant_full(count, direction)[:2]

    def fused_variant_full(self, count: int = 1, direction: int = FORWARD):
        """(shape, lazy, log2 words per workgroup, on-the-fly twiddles) of the fused kernel."""
        eb, lz, lw, ot = _I(0), _I(0), _I(0), _I(0)
        _check(_lib.ntt_plan_get_fused_variant(self.handle, count, direction, ctypes.byref(eb), ctypes.byref(lz),
                                               ctypes.byref(lw), ctypes.byref(ot)), "ntt_plan_get_fused_variant")
        return eb.value, bool(lz.value), lw.value, bool(ot.value)

    def set_constants(self, prologue=None, epilogue=None):
        """Fused per-limb constant multiplies: forward computes NTT(prologue * x),
        inverse computes epilogue * INTT(x) (None = 1)."""
        def arr(v):
            if v is None:
                return None
            a = (ctypes.c_uint32 * self.limbs)(*[int(x) for x in v])
            return ctypes.cast(a, _PU32), a
        p, e = arr(prologue), arr(epilogue)
        _check(_lib.ntt_plan_set_constants(self.handle, p[0] if p else None, e[0] if e else None),
               "ntt_plan_set_constants")

    def set_twiddle_mode(self, mode: int):
        """TWIDDLE_AUTO (default: the tuned choice; untuned, on the fly for the inverse and table twiddles
        for the forward), TWIDDLE_TABLE or TWIDDLE_OT."""
        _check(_lib.ntt_plan_set_twiddle_mode(self.handle, mode), "ntt_plan_set_twiddle_mode")

    def set_fused_arith(self, lazy: int):
        """Fused arithmetic, both directions: 1 lazy (when every q < 2^30), 0 canonical, -1 tuned."""
        _check(_lib.ntt_plan_set_fused_arith(self.handle, lazy), "ntt_plan_set_fused_arith")

    def set_fused_words(self, log2_words: int):
        _check(_lib.ntt_plan_set_fused_words(self.handle, log2_words), "ntt_plan_set_fused_words")

    def set_fused_shape(self, eb: int):
        """Fused workgroup shape, both directions: 7 (512 threads x 128 words), 6 (1024 x 64), 0 = tuned."""
        _check(_lib.ntt_plan_set_fused_shape(self.handle, eb), "ntt_plan_set_fused_shape")

    def engine(self, count: int = 1, direction: int = FORWARD) -> int:
        """Engine a launch of `count` operands uses in `direction` (1 two-pass, 2 fused)."""
        return int(_lib.ntt_plan_get_engine_for(self.handle, count, direction))

    def run_pass(self, A, count: int, direction: int, pass_index: int, stream=None):
        _check(_lib.ntt_plan_run_pass(self.handle, _addr(A), count, direction, pass_index, _stream(stream)),
               "ntt_plan_run_pass")

    def close(self):
        if self.handle:
            _lib.ntt_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- metrics (src/v04_montgomery.cu:808-818) ---------------------------------------
def butterflies(logN: int, limbs: int) -> float:
    """Butterflies per transform = N/2 * logN * limbs."""
    return 0.5 * (1 << logN) * logN * limbs


def algorithmic_bytes(logN: int, limbs: int) -> int:
    """Each uint32 word read once and written once."""
    return 8 * (1 << logN) * limbs


from . import workloads  # noqa: E402,F401
