"""ctypes wrapper around the CPU oracle (oracle/ntt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product package (ntt_amd)
never imports this module.

Every function here forwards to the C restatement of the reference CPU path
(src/v04_montgomery.cu:93-146, 660-797; src/v03_template_config.cu:516-527);
the C file's header lists the exact line ranges and how each is pinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libntt_oracle.so")
_lib = None

U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
P32 = ctypes.POINTER(ctypes.c_uint32)
P64 = ctypes.POINTER(ctypes.c_uint64)


class ModulusRuntime(ctypes.Structure):
    """Field order of ModulusRuntime, src/v04_montgomery.cu:17-26."""

    _fields_ = [(n, ctypes.c_uint32) for n in
                ("qi", "primitive_root", "q_inv", "r2", "psi", "psi_inv", "omega", "degree_inv")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def build() -> str:
    """Compile the oracle with its Makefile (gcc) and return the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name, res, args in [
            ("oracle_add_mod", U32, [U32, U32, U32]),
            ("oracle_sub_mod", U32, [U32, U32, U32]),
            ("oracle_mul_mod", U32, [U32, U32, U32]),
            ("oracle_mont_reduce", U32, [U64, U32, U32]),
            ("oracle_mont_mul", U32, [U32, U32, U32, U32]),
            ("oracle_mont_q_inv", U32, [U32]),
            ("oracle_to_mont", U32, [U32, U32]),
            ("oracle_mont_r2", U32, [U32]),
            ("oracle_from_mont", U32, [U32, U32, U32]),
            ("oracle_pow_mod", U32, [U32, U64, U32]),
            ("oracle_bit_reverse", U32, [U32, ctypes.c_int]),
            ("oracle_make_modulus_runtime", ctypes.c_int,
             [ctypes.c_int, U32, U32, ctypes.POINTER(ModulusRuntime)]),
            ("oracle_make_twiddle_table", None, [ctypes.c_int, U32, U32, P32]),
            ("oracle_cpu_ntt_butterfly", None, [P32, ctypes.c_int, U32, P32]),
            ("oracle_cpu_intt_butterfly", None, [P32, ctypes.c_int, U32, P32, U32]),
            ("oracle_forward", ctypes.c_int, [P32, ctypes.c_int, P32, P32, ctypes.c_size_t, ctypes.c_int]),
            ("oracle_inverse", ctypes.c_int, [P32, ctypes.c_int, P32, P32, ctypes.c_size_t, ctypes.c_int]),
            ("oracle_fill_random", ctypes.c_int, [P32, ctypes.c_size_t, U32, U32]),
            ("oracle_fill_random_limbs", ctypes.c_int, [P32, ctypes.c_int, P32, ctypes.c_size_t, U32]),
            ("oracle_to_mont_inplace", None, [P32, ctypes.c_size_t, U32]),
            ("oracle_from_mont_inplace", None, [P32, ctypes.c_size_t, U32, U32]),
            ("oracle_mul_mod64", U64, [U64, U64, U64]),
            ("oracle_pow_mod64", U64, [U64, U64, U64]),
            ("oracle_psi64", U64, [ctypes.c_int, U64, U64]),
            ("oracle_forward64", ctypes.c_int, [P64, ctypes.c_int, P64, P64, ctypes.c_size_t]),
            ("oracle_inverse64", ctypes.c_int, [P64, ctypes.c_int, P64, P64, ctypes.c_size_t]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P32)


def _moduli_arrays(moduli):
    qs = np.ascontiguousarray([m[0] for m in moduli], dtype=np.uint32)
    gs = np.ascontiguousarray([m[1] for m in moduli], dtype=np.uint32)
    return qs, gs


# ---- scalar helpers -------------------------------------------------------
def add_mod(a, b, q): return int(lib().oracle_add_mod(a, b, q))
def sub_mod(a, b, q): return int(lib().oracle_sub_mod(a, b, q))
def mul_mod(a, b, q): return int(lib().oracle_mul_mod(a, b, q))
def mont_mul(a, b, q, q_inv): return int(lib().oracle_mont_mul(a, b, q, q_inv))
def mont_q_inv(q): return int(lib().oracle_mont_q_inv(q))
def to_mont(x, q): return int(lib().oracle_to_mont(x, q))
def mont_r2(q): return int(lib().oracle_mont_r2(q))
def from_mont(x, q, q_inv): return int(lib().oracle_from_mont(x, q, q_inv))
def pow_mod(b, e, m): return int(lib().oracle_pow_mod(b, e, m))
def bit_reverse(v, w): return int(lib().oracle_bit_reverse(v, w))


def make_modulus_runtime(logN: int, qi: int, root: int):
    """Returns (rc, dict); rc 0 ok, 1 q != 1 mod 2N, 2 bad root."""
    rt = ModulusRuntime()
    rc = lib().oracle_make_modulus_runtime(logN, qi, root, ctypes.byref(rt))
    return rc, rt.as_dict()


def twiddle_table(logN: int, q: int, psi: int) -> np.ndarray:
    out = np.empty(1 << logN, dtype=np.uint32)
    lib().oracle_make_twiddle_table(logN, q, psi, _p(out))
    return out


def cpu_ntt_butterfly(a: np.ndarray, logN: int, q: int, psi_table: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32).copy()
    lib().oracle_cpu_ntt_butterfly(_p(a), logN, q, _p(np.ascontiguousarray(psi_table, dtype=np.uint32)))
    return a


def cpu_intt_butterfly(a: np.ndarray, logN: int, q: int, inv_psi_table: np.ndarray, degree_inv: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32).copy()
    lib().oracle_cpu_intt_butterfly(_p(a), logN, q, _p(np.ascontiguousarray(inv_psi_table, dtype=np.uint32)), degree_inv)
    return a


# ---- whole-operand entry points over A[limb * N + i] ----------------------
PLAIN, MONTGOMERY = 0, 1


def forward(A: np.ndarray, logN: int, moduli, domain: int = MONTGOMERY) -> np.ndarray:
    """Oracle forward NTT of a (limbs, N) or flat uint32 array; returns a new array."""
    out = np.ascontiguousarray(A, dtype=np.uint32).copy()
    qs, gs = _moduli_arrays(moduli)
    rc = lib().oracle_forward(_p(out), logN, _p(qs), _p(gs), len(moduli), domain)
    if rc:
        raise ValueError(f"oracle_forward failed rc={rc}")
    return out


def inverse(A: np.ndarray, logN: int, moduli, domain: int = MONTGOMERY) -> np.ndarray:
    out = np.ascontiguousarray(A, dtype=np.uint32).copy()
    qs, gs = _moduli_arrays(moduli)
    rc = lib().oracle_inverse(_p(out), logN, _p(qs), _p(gs), len(moduli), domain)
    if rc:
        raise ValueError(f"oracle_inverse failed rc={rc}")
    return out


def mod_down(ntt_out: np.ndarray, sub: np.ndarray, factors, moduli, sub_start: int = 0, sub_end=None,
             sub_factors=None) -> np.ndarray:
    """The subtract-and-scale epilogue of Cheddar's NTTForModDown
    (src/NTT_cheddar.cu:566-599): per limb l, s = src2 limb l - sub_start
    when sub_start <= l < sub_end (src2_exists, 575-576), multiplied by
    sub_factors[l - sub_start] when given (src2_padding, 579-584), else 0;
    then res = s - NTT and res * inv_p_prod mod the limb's prime:
    out[l] = (s - ntt_out[l]) * factors[l] mod q_l, canonical.  `sub` holds
    sub_end - sub_start limbs (all limbs by default)."""
    L = len(moduli)
    sub_end = L if sub_end is None else sub_end
    qs = np.array([int(q) for q, _ in moduli], dtype=np.int64)[:, None]
    f = np.array([int(v) for v in factors], dtype=np.int64)[:, None] % qs
    x = np.asarray(ntt_out, dtype=np.int64).reshape(L, -1)
    y = np.zeros_like(x)
    if sub_end > sub_start:
        sv = np.asarray(sub, dtype=np.int64).reshape(sub_end - sub_start, -1)
        for l in range(sub_start, sub_end):
            q = int(qs[l, 0])
            g = 1 if sub_factors is None else int(sub_factors[l - sub_start]) % q
            y[l] = sv[l - sub_start] % q * g % q       # < 2^31 * 2^31: exact in int64
    d = (y - x) % qs                          # in [0, q) < 2^31
    # d * f < 2^62: exact in int64
    return (d * f % qs).astype(np.uint32)


def psi64(logN: int, q: int, root: int) -> int:
    return int(lib().oracle_psi64(logN, q, root))


def _moduli64(moduli):
    qs = np.array([int(q) for q, _ in moduli], dtype=np.uint64)
    rs = np.array([int(r) for _, r in moduli], dtype=np.uint64)
    return qs, rs


def forward64(A: np.ndarray, logN: int, moduli) -> np.ndarray:
    """Normal-domain forward NTT over 64-bit words of a (limbs, N) array;
    moduli = [(q, root)] with q < 2^62 (ntt_oracle.c, 64-bit section)."""
    out = np.ascontiguousarray(A, dtype=np.uint64).copy()
    qs, rs = _moduli64(moduli)
    rc = lib().oracle_forward64(out.ctypes.data_as(P64), logN, qs.ctypes.data_as(P64), rs.ctypes.data_as(P64),
                                len(moduli))
    if rc:
        raise ValueError(f"oracle_forward64 failed rc={rc}")
    return out


def inverse64(A: np.ndarray, logN: int, moduli) -> np.ndarray:
    out = np.ascontiguousarray(A, dtype=np.uint64).copy()
    qs, rs = _moduli64(moduli)
    rc = lib().oracle_inverse64(out.ctypes.data_as(P64), logN, qs.ctypes.data_as(P64), rs.ctypes.data_as(P64),
                                len(moduli))
    if rc:
        raise ValueError(f"oracle_inverse64 failed rc={rc}")
    return out


def fill_random_limbs(logN: int, moduli, seed: int) -> np.ndarray:
    """The reference's fill_random per limb, each limb re-seeded with `seed`."""
    qs, _ = _moduli_arrays(moduli)
    out = np.empty((len(moduli), 1 << logN), dtype=np.uint32)
    rc = lib().oracle_fill_random_limbs(_p(out), logN, _p(qs), len(moduli), seed)
    if rc:
        raise ValueError("fill_random: q must be > 1")
    return out


def fill_random(n: int, q: int, seed: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    if lib().oracle_fill_random(_p(out), n, q, seed):
        raise ValueError("fill_random: q must be > 1")
    return out
