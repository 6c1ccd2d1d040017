"""CPU tests of the ntt_amd library's host side: the C ABI loads and exports
every declared symbol, and the host precompute, moduli generator and phase
config logic agree with the oracle and with independent mathematics.  No
compute call touches a GPU here."""
import ctypes
import json
import os
import random

import numpy as np
import pytest

import ntt_amd
from tests import mathref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "moduli.json")))


def test_library_exports_every_header_symbol():
    names = ntt_amd.header_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(ntt_amd.lib(), n)]
    assert not missing, missing


def test_version_and_limits():
    assert b"gfx950" in ntt_amd.lib().ntt_version()
    assert ntt_amd.lib().ntt_max_log_n() >= 17


def test_default_moduli_matches_reference():
    # src/ntt_config.h:75-94
    ms = [tuple(m) for m in ntt_amd.default_moduli()]
    assert ms[0] == (974258177, 3) and ms[-1] == (1016463361, 38) and len(ms) == 16
    for q, g in ms:
        assert mathref.is_prime(q) and mathref.smallest_primitive_root(q) == g


def test_generate_moduli_matches_fixtures_and_math():
    m16 = [tuple(m) for m in ntt_amd.generate_moduli(16, 16)]
    assert m16 == [tuple(x) for x in FIX["M16"]]
    m48 = [tuple(m) for m in ntt_amd.generate_moduli(17, 48)]
    assert m48 == [tuple(x) for x in FIX["M48"]]
    # independent derivation (trial division roots, Miller-Rabin in Python)
    assert m16 == mathref.largest_ntt_primes(16, 1 << 17)
    assert m48[:12] == mathref.largest_ntt_primes(12, 1 << 18)
    for logN in (12, 13, 14, 15):
        got = [tuple(m) for m in ntt_amd.generate_moduli(logN, 8)]
        assert got == mathref.largest_ntt_primes(8, 2 << logN)


@pytest.mark.parametrize("logN", [1, 8, 12, 16, 17])
def test_modulus_runtime_matches_oracle(oracle_mod, logN):
    for q, g in [tuple(m) for m in ntt_amd.default_moduli()] + [tuple(x) for x in FIX["M48"][:8]]:
        rc, ref = oracle_mod.make_modulus_runtime(logN, q, g)
        if rc:
            with pytest.raises(ntt_amd.NTTError) as e:
                ntt_amd.make_modulus_runtime(logN, (q, g))
            assert e.value.status == {1: 2, 2: 3}[rc]
            continue
        assert ntt_amd.make_modulus_runtime(logN, (q, g)) == ref


def test_modulus_errors():
    with pytest.raises(ntt_amd.NTTError) as e:
        ntt_amd.make_modulus_runtime(17, (974258177, 3))        # v2(q-1) = 17
    assert e.value.status == 2
    with pytest.raises(ntt_amd.NTTError) as e:
        ntt_amd.make_modulus_runtime(10, (998244353, 4))        # not a primitive root
    assert e.value.status == 3
    for bad in [(1, 1), (2, 1), (2147483649, 3), (4294967291, 2)]:
        with pytest.raises(ntt_amd.NTTError) as e:
            ntt_amd.make_modulus_runtime(4, bad)
        assert e.value.status == 4


@pytest.mark.parametrize("logN", [1, 5, 12, 16])
def test_twiddle_tables_match_oracle(oracle_mod, logN):
    q, g = FIX["M48"][3]
    rt = ntt_amd.make_modulus_runtime(logN, (q, g))
    assert np.array_equal(ntt_amd.make_twiddle_table(logN, q, rt["psi"]),
                          oracle_mod.twiddle_table(logN, q, rt["psi"]))


@pytest.mark.parametrize("domain", [0, 1])
def test_device_tables_match_reference_main(oracle_mod, domain):
    """The arrays the reference's main() uploads (v04:893-915)."""
    logN = 12
    moduli = FIX["M16"][:5]
    t = ntt_amd.make_tables(logN, moduli, domain)
    N = 1 << logN
    for l, (q, g) in enumerate(moduli):
        rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
        tw = oracle_mod.twiddle_table(logN, q, rt["psi"])
        itw = oracle_mod.twiddle_table(logN, q, rt["psi_inv"])
        if domain == 1:
            tw = np.array([oracle_mod.to_mont(int(x), q) for x in tw], dtype=np.uint32)
            itw = np.array([oracle_mod.to_mont(int(x), q) for x in itw], dtype=np.uint32)
        assert np.array_equal(t.twiddles[l * N:(l + 1) * N], tw)
        assert np.array_equal(t.inv_twiddles[l * N:(l + 1) * N], itw)
        assert t.moduli[l] == q and t.q_inv[l] == rt["q_inv"] and t.r2[l] == rt["r2"]
        assert t.degree_inv[l] == (oracle_mod.to_mont(rt["degree_inv"], q) if domain else rt["degree_inv"])


def test_montgomery_helpers():
    rng = random.Random(5)
    for q, _ in [tuple(m) for m in ntt_amd.default_moduli()]:
        qi = ntt_amd.mont_q_inv(q)
        assert (q * qi) % (1 << 32) == (1 << 32) - 1
        assert ntt_amd.mont_r2(q) == pow(2, 64, q)
        for _ in range(50):
            x = rng.randrange(q)
            assert ntt_amd.to_mont(x, q) == (x << 32) % q
            assert ntt_amd.from_mont(ntt_amd.to_mont(x, q), q, qi) == x


def test_reference_default_config_values():
    # DefaultNTTLaunchConfig<LogN>, src/ntt_config.h:62-73
    c = ntt_amd.reference_launch_config(16).astuple()
    assert c == ((7, 4, 4), (9, 3, 0), (9, 3, 0), (7, 4, 4))
    c = ntt_amd.reference_launch_config(17).astuple()
    assert c == ((8, 3, 4), (9, 3, 0), (9, 3, 0), (8, 3, 4))


def test_wave64_default_config():
    c = ntt_amd.default_launch_config(16).astuple()
    assert c == ((7, 4, 6), (9, 3, 0), (9, 3, 0), (7, 4, 6))
    for logN in range(1, 21):
        cfg = ntt_amd.default_launch_config(logN)
        assert ntt_amd.validate_config(logN, cfg) == 0, (logN, cfg)


@pytest.mark.parametrize("cfg,ok", [
    (((7, 4, 6), (9, 3, 0)), True),
    (((7, 4, 4), (9, 3, 0)), True),
    (((8, 4, 4), (9, 3, 0)), False),      # radix1 + radix2 != logN (unchecked in the reference)
    (((7, 8, 4), (9, 3, 0)), False),      # stage_merging > radix_stages
    (((7, 6, 0), (9, 3, 0)), False),      # 2^S > 32 words per thread
    (((7, 1, 6), (9, 3, 0)), False),      # 2^(7-1+6) = 4096 threads
    (((7, 4, 10), (9, 3, 0)), False),     # more columns per tile than columns
    (((16, 5, 0), (0, 0, 0)), False),     # 2^11 threads
    (((0, 0, 0), (16, 5, 0)), False),
    (((7, 4, -1), (9, 3, 0)), False),
])
def test_config_validation(cfg, ok):
    c = ntt_amd.TransformConfig.make(*cfg)
    assert (ntt_amd.validate_config(16, c) == 0) == ok


def test_lds_limit_enforced():
    # a 2^15-word column tile is 132 KiB with padding (fits 160 KiB); 2^16 does not
    assert ntt_amd.validate_config(17, ntt_amd.TransformConfig.make((10, 5, 5), (7, 3, 0))) == 0
    assert ntt_amd.validate_config(17, ntt_amd.TransformConfig.make((10, 5, 6), (7, 3, 0))) != 0


def test_generate_moduli64_matches_python_search():
    """ntt_generate_moduli64 against a Python-integer search (Miller-Rabin,
    Euler's criterion): same primes, same smallest quadratic non-residues."""
    from tests.test_oracle import _primes64
    got = [tuple(m) for m in ntt_amd.generate_moduli64(17, 4)]
    assert got == _primes64(4, logN=17)
    for q, g in ntt_amd.generate_moduli64(12, 3, 1 << 40):
        assert q < (1 << 40) and (q - 1) % (1 << 13) == 0 and pow(g, (q - 1) // 2, q) == q - 1


def test_generate_moduli64_arguments():
    with pytest.raises(ntt_amd.NTTError):
        ntt_amd.generate_moduli64(0, 2)
    with pytest.raises(ntt_amd.NTTError):
        ntt_amd.generate_moduli64(30, 2)
    assert ntt_amd.generate_moduli64(12, 0) == []
    # a bound too small for any q = 1 mod 2N yields none; the bound is capped at 2^62
    assert ntt_amd.generate_moduli64(12, 3, 1 << 13) == []
    assert all(m.q < (1 << 62) for m in ntt_amd.generate_moduli64(12, 2, 1 << 63))
