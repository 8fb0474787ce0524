"""GPU: kc_library_verify's closed-form checks on runs whose outputs equal the oracle's:
the seen tier equals the OR of every window's bits, the Bloom-solid windows of the counted
reads sum to the count table's total, the table is strictly ascending with consistent
solid flags, and its checksum equals the one computed on the host from the oracle's
table.  Also the metagenome sample at 1/20 scale (20 genomes, lognormal abundances,
k = 55) against the oracle's digests."""
import json
import os
import sys

import numpy as np
import pytest

from kc import AssemblerConfig, SpectrumEngine, synth
from kc.digest import diff, engine_digests, table_checksum

from conftest import GOLDEN
from test_gpu_graph import _uneven_library

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_verify(v, info, okeys, ocnt, priors=False):
    assert v["windows"] == info["n_occ"]
    assert v["seen_words_differing"] == 0
    assert v["unordered_pairs"] == 0 and v["mask_mismatches"] == 0
    assert v["sum_counts"] == int(ocnt.astype(np.uint64).sum())
    if not priors:
        assert v["solid_windows"] == v["sum_counts"]
    assert v["checksum"] == table_checksum(okeys, ocnt)


@pytest.mark.parametrize("k,ec,bits", [(21, False, 1 << 20), (33, True, 1 << 16), (55, False, 4096),
                                       (63, True, (1 << 18) + 64)])
def test_verify_on_one_library(oracle, k, ec, bits):
    store, _ = synth.uniform_genome_pairs(seed=70 + k, genome_size=20000, n_pairs=2000, read_len=150, p_start=0.005)
    d, o, l = store.arrays()
    say(f"case verify k={k} ec={ec} bits={bits}")
    S = oracle.Spectrum(d, o, l, k, bits, ec, 5.0, 1)
    okeys, ocnt, _ = S.counts()
    for events64 in (False, True):
        eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec), events64=events64)
        eng.upload(store)
        info = eng.run()
        check_verify(eng.verify(), info, okeys, ocnt)
        eng.close()


@pytest.mark.parametrize("k,ec", [(31, False), (55, True)])
def test_verify_on_several_libraries(oracle, k, ec):
    libs, _ = zip(*[_uneven_library(900 + i, 3000 + 1000 * i, 150 + 40 * i) for i in range(3)])
    bits = 1 << 16
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec))
    for st in libs:
        eng.add_library(st, bloom_bits=bits)
    say(f"case verify libraries k={k} ec={ec}")
    eng.run()
    for i, st in enumerate(libs):
        d, o, l = st.arrays()
        S = oracle.Spectrum(d, o, l, k, bits, ec, 5.0, 1)
        okeys, ocnt, _ = S.counts()
        check_verify(eng.verify(i), eng.info(i), okeys, ocnt)
    eng.close()


def test_checksum_sees_one_changed_count():
    rng = np.random.default_rng(1)
    keys = np.sort(rng.integers(0, 2**62, 1000, dtype=np.uint64))
    vals = np.zeros((1000, 4), np.uint64)
    vals[:, 0] = keys
    cnt = rng.integers(1, 50, 1000).astype(np.uint32)
    a = table_checksum(vals, cnt)
    cnt[500] += 1
    assert table_checksum(vals, cnt) != a


DIGESTS = json.load(open(os.path.join(GOLDEN, "full_size_digests.json")))


def test_metagenome_twentieth_matches_oracle():
    want = DIGESTS["meta_0.05_k55"]
    store, meta = synth.workload_metagenome(seed=4, scale=0.05)
    say(f"case metagenome 1/20 reads={len(store)} genomes={len(meta['pairs'])}")
    eng = SpectrumEngine(AssemblerConfig(k=55, bloom_bits=want["bloom_bits"]))
    eng.upload(store)
    info = eng.run()
    got = engine_digests(eng, info)
    assert not diff(got, want), diff(got, want)
    keys, cnt, _ = eng.counts()
    v = eng.verify()
    eng.close()
    check_verify(v, info, keys, cnt)
