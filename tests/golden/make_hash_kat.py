"""Writes tests/golden/hash_kat.json: Kmer256 limbs of random canonical k-mers and the
three bloom.rs:61-78 hashes of each, as computed by the NumPy transcription
(oracle/hash_np.py).  The C++ oracle must agree with it (checked when the
fixture is written and by tests/test_oracle_bloom.py)."""
import json
import os
import random
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import hash_np as H  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    rng = random.Random(20261203)
    cases = []
    for k in (5, 15, 21, 31, 32, 33, 55, 63, 64, 65, 99, 127):
        for _ in range(24):
            s = bytes(rng.choice(b"ACGT") for _ in range(k))
            w = H.canonical_words_of_string(s)
            cases.append({"k": k, "seq": s.decode(), "limbs": [str(x) for x in w]})
    # the reference's own u64 test keys (bloom.rs:141-142, tests/test_hammer_and_bloom.rs:43-45)
    for x in (123456789, 987654321, 0x123456789ABCDEF0, 0xFEDCBA9876543210, 0xA5A5A5A5A5A5A5A5):
        cases.append({"k": 0, "seq": "", "limbs": [str(x), "0", "0", "0"]})
    limbs = np.array([[int(v) for v in c["limbs"]] for c in cases], dtype=np.uint64)
    for j, seed in enumerate(H.SEEDS):
        hv = H.hash_kmer256(limbs[:, 0], limbs[:, 1], limbs[:, 2], limbs[:, 3], seed)
        cpp = O.hash_many(limbs, int(seed))
        assert np.array_equal(hv, cpp), "NumPy and C++ transcriptions disagree"
        for c, h in zip(cases, hv):
            c[f"h{j + 1}"] = str(int(h))
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "hash_kat.json"), "w") as fh:
        json.dump({"seeds": [str(int(s)) for s in H.SEEDS], "cases": cases}, fh, separators=(",", ":"))


if __name__ == "__main__":
    main()
