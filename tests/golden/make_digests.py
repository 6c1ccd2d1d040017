"""Writes tests/golden/digests.json: SHA-256 per limb and the first/last 16
words per limb of the oracle's outputs on the C1-C3 workloads and on the
reference's own workload.

C1  plain NTT, N = 2^16, moduli M16, fill_random seed 0
C2  Montgomery NTT, same data
C3  Montgomery NTT, N = 2^17, moduli M48, seed 0
REF Montgomery NTT, N = 2^16, default_moduli(), seed 1234 (src/v04_montgomery.cu:28)

Inputs are recorded the same way, so a change in fill_random shows up apart
from a change in the transform.  Run from the repository root:
    python3 tests/golden/make_digests.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

DEFAULT_MODULI = [(974258177, 3), (1081212929, 6), (1196556289, 7), (993263617, 5), (989986817, 3),
                  (1074266113, 5), (1168900097, 5), (1010565121, 7), (957349889, 6), (1073872897, 7),
                  (1209139201, 31), (994705409, 3), (998244353, 3), (1073479681, 11), (1158676481, 3),
                  (1016463361, 38)]


def limb_record(a: np.ndarray):
    return [{"sha256": hashlib.sha256(np.ascontiguousarray(row, dtype="<u4").tobytes()).hexdigest(),
             "head": [int(v) for v in row[:16]], "tail": [int(v) for v in row[-16:]]} for row in a]


def cases():
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "moduli.json")))
    m16 = [tuple(x) for x in fx["M16"]]
    m48 = [tuple(x) for x in fx["M48"]]
    return {
        "C1": (16, m16, 0, orc.PLAIN),
        "C2": (16, m16, 0, orc.MONTGOMERY),
        "C3": (17, m48, 0, orc.MONTGOMERY),
        "REF": (16, DEFAULT_MODULI, 1234, orc.MONTGOMERY),
    }


def main():
    out = {"_comment": "oracle outputs on C1-C3 and the reference workload; made by tests/golden/make_digests.py"}
    for name, (logN, moduli, seed, domain) in cases().items():
        x = orc.fill_random_limbs(logN, moduli, seed)
        y = orc.forward(x, logN, moduli, domain)
        out[name] = {"logN": logN, "moduli": [list(m) for m in moduli], "seed": seed, "domain": domain,
                     "input": limb_record(x), "forward": limb_record(y)}
        print(name, "done", flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "digests.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
