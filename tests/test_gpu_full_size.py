"""GPU parity at full size: the engine's outputs on the 5 Mb / 100x workload (k = 21, 33,
55, 2^29-bit tiers) and on the 1 Mb / 30x workload (k = 21, Hammer off and on) against
SHA-256 digests of the sequential oracle's outputs on the same seeded reads
(tests/golden/full_size_digests.json, written by tests/golden/make_full_digests.py):
both Bloom tiers, the sorted count table, the counts, the solid mask, the cutoff and, with
Hammer on, the corrected reads.  Integer work: every digest must match exactly."""
import json
import os
import sys

import pytest

from kc import AssemblerConfig, SpectrumEngine, synth
from kc.digest import diff, engine_digests

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

DIGESTS = json.load(open(os.path.join(GOLDEN, "full_size_digests.json")))
BITS = 1 << 29


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def store_5mb():
    return synth.workload_5mb_100x(seed=2)[0]


@pytest.fixture(scope="module")
def store_1mb():
    return synth.workload_1mb_30x(seed=1)[0]


def run_digests(store, k, ec, tune=None):
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=BITS, error_correct=ec))
    try:
        if tune:
            eng.set_tune(**tune)
        eng.upload(store)
        info = eng.run()
        first = engine_digests(eng, info, with_reads=ec)
        info = eng.run()  # the same context again (warm sizes): same outputs
        return first, engine_digests(eng, info, with_reads=ec)
    finally:
        eng.close()


@pytest.mark.parametrize("k", [21, 33, 55])
def test_5mb_100x_matches_oracle(store_5mb, k):
    want = DIGESTS[f"5mb_100x_k{k}"]
    say(f"case 5mb_100x k={k} bits={BITS} reads={len(store_5mb)}")
    for got in run_digests(store_5mb, k, False):
        assert not diff(got, want), (diff(got, want), got, want)


@pytest.mark.parametrize("k,ec", [(21, False), (21, True), (33, False), (55, False)])
def test_1mb_30x_matches_oracle(store_1mb, k, ec):
    want = DIGESTS[f"1mb_30x_k{k}" + ("_hammer" if ec else "")]
    say(f"case 1mb_30x k={k} ec={ec} reads={len(store_1mb)}")
    for got in run_digests(store_1mb, k, ec):
        assert not diff(got, want), (diff(got, want), got, want)


@pytest.mark.parametrize("k", [21, 55])
def test_5mb_100x_with_full_bin_tables_matches_oracle(store_5mb, k):
    """Bins sized for 4096 distinct k-mers (2,600-3,000 on average here, against the LDS
    table's 768): nearly every bin fills its table, keeps half of its keys by the next
    key-hash bit and hands the other half to a pass of its own, often several times."""
    want = DIGESTS[f"5mb_100x_k{k}"]
    say(f"case 5mb_100x k={k} kmers_per_bin=4096")
    for got in run_digests(store_5mb, k, False, tune={"kmers_per_bin": 4096}):
        assert not diff(got, want), (diff(got, want), got, want)


def test_1mb_30x_hammer_with_full_bin_tables_matches_oracle(store_1mb):
    """The same for Hammer's recount over the corrected reads, which also takes the
    binned count."""
    want = DIGESTS["1mb_30x_k21_hammer"]
    say("case 1mb_30x k=21 ec=True kmers_per_bin=4096")
    for got in run_digests(store_1mb, 21, True, tune={"kmers_per_bin": 4096}):
        assert not diff(got, want), (diff(got, want), got, want)
