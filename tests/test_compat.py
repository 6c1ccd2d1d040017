"""The header-only C++ mirror of the reference interface
(include/ntt_amd_compat.hpp): a reference-style caller compiled unchanged
against it (tests/compat/drop_in_main.cpp, built by ntt_amd/build.py)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "compat", "_bin", "drop_in_main")


def _bin():
    if not os.path.exists(BIN):
        pytest.fail("tests/compat/_bin/drop_in_main is missing: run __graft_entry__.build()")
    return BIN


def _info():
    out = subprocess.run([_bin(), "info"], capture_output=True, text=True, check=True).stdout.splitlines()
    moduli, cfgs, cur = [], {}, None
    for ln in out:
        f = ln.split()
        if f[0] == "modulus":
            moduli.append((int(f[1]), int(f[2])))
        elif f[0] == "config":
            cur = cfgs.setdefault(f[1], [])
        elif f[0].endswith("_phase1") or f[0].endswith("_phase2"):
            cur.append(tuple(int(v) for v in f[1:]))
    return moduli, cfgs


def test_compat_header_configs_and_moduli():
    import ntt_amd
    moduli, cfgs = _info()
    assert moduli == [tuple(m) for m in ntt_amd.default_moduli()]
    for n in (10, 12, 16, 17):
        assert cfgs[f"default{n}"] == [tuple(p) for p in ntt_amd.default_launch_config(n).astuple()], n
    # the reference's literal DefaultNTTLaunchConfig<16> (src/ntt_config.h:62-73)
    assert cfgs["reference16"] == [(7, 4, 4), (9, 3, 0), (9, 3, 0), (7, 4, 4)]


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["mont-default", "mont-reference", "plain", "plain-runtime"])
def test_compat_drop_in_run(oracle_mod, tmp_path, variant):
    """The reference's workload (default_moduli, logN 16, fill_random with
    seed 1234, v04:28 and 786-797) through the mirrored launchers: forward
    bit-exact against the oracle, inverse back to the input."""
    import ntt_amd
    from tests import gpu_util
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    orc = oracle_mod
    moduli = [tuple(m) for m in ntt_amd.default_moduli()]
    x = orc.fill_random_limbs(16, moduli, 1234)
    inp, fo, bo = tmp_path / "in.bin", tmp_path / "fwd.bin", tmp_path / "back.bin"
    x.astype(np.uint32).tofile(inp)
    print("case compat", variant, flush=True)
    r = subprocess.run([_bin(), "run", variant, str(inp), str(fo), str(bo)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    fwd = np.fromfile(fo, np.uint32).reshape(x.shape)
    back = np.fromfile(bo, np.uint32).reshape(x.shape)
    domain = 1 if variant.startswith("mont") else 0
    assert np.array_equal(fwd, orc.forward(x, 16, moduli, domain))
    assert np.array_equal(back, x)
