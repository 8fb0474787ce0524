"""GPU: the Bloom first setters taken by sorting (k_fs_items, fs_sort, k_fs_min: three
(entry, setter) items per distinct k-mer radix-sorted by partition, minima in LDS) give the
same seen / solid tiers, count table and cutoff as the device atomics (setter_path 1) and as
the oracle: one and several libraries (first-setter arrays of uneven sizes side by side in
one group, so partitions straddle library boundaries), sizes that are not a multiple of the
16384-entry partition, Hammer on, k > 63, graph replay, and the sparse mode, which keeps the
atomics."""
import sys

import numpy as np
import pytest

from kc import AssemblerConfig, SpectrumEngine, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sorted_launched(eng):
    return eng.kernel_times().get("k_fs_min", (0.0, 0))[1] > 0


@pytest.mark.parametrize("k,ec,bits,ip", [(21, False, 1 << 20, 0), (33, False, (1 << 18) + 64, 0),
                                          (55, True, 4096, 0), (31, False, 64, 0), (21, True, 1 << 16, 1),
                                          (77, False, (1 << 17) + 192, 0), (127, True, 1 << 18, 0),
                                          (45, False, 1 << 19, 1)])
def test_sorted_setters_match_atomics_and_oracle(oracle, k, ec, bits, ip):
    store, _ = synth.uniform_genome_pairs(seed=900 + k, genome_size=20000, n_pairs=2000, read_len=150,
                                          p_start=0.005)
    d, o, l = store.arrays()
    S = oracle.Spectrum(d, o, l, k, bits, ec, 5.0, 1)
    okeys, ocnt, omask = S.counts()
    oseen, osolid = S.bloom()
    distinct = set()
    for path in (0, 1):
        say(f"case setter path={path} k={k} ec={ec} bits={bits} insert path={ip}")
        eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec), timing=True)
        try:
            eng.set_tune(insert_path=ip, setter_path=path)
            eng.upload(store)
            for rep in range(2):  # the second run reuses every buffer
                info = eng.run()
                assert sorted_launched(eng) == (path == 0)
                seen, solid = eng.bloom()
                keys, cnt, mask = eng.counts()
                assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid)
                assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt) and np.array_equal(mask, omask)
                assert info["cutoff"] == S.cutoff and info["n_solid"] == S.n_solid
                distinct.add(info["n_distinct"])
        finally:
            eng.close()
    assert len(distinct) == 1  # the item pass counts the distinct k-mers as the atomics do


@pytest.mark.parametrize("k,ec", [(33, False), (25, True), (99, False)])
def test_several_libraries_uneven_arrays(oracle, k, ec):
    """Three libraries with 2^16 + 64, 2^15 and 4096 + 64-bit tiers share one first-setter
    array; eager runs and graph replays equal the oracle library by library."""
    bits = [(1 << 16) + 64, 1 << 15, 4096 + 64]
    stores = [synth.uniform_genome_pairs(seed=950 + i, genome_size=6000 + 3000 * i, n_pairs=600 + 200 * i,
                                         read_len=150, p_start=0.004)[0] for i in range(3)]
    eng = SpectrumEngine(AssemblerConfig(k=k, error_correct=ec), timing=True)
    try:
        for st, b in zip(stores, bits):
            eng.add_library(st, bloom_bits=b)
        for mode in ("eager", "graph plan", "graph replay", "graph replay"):
            say(f"case libraries k={k} ec={ec} {mode}")
            if mode == "eager":
                eng.run()
                assert sorted_launched(eng)
            else:  # graphs take the atomics (no radix-sort memset nodes in a graph)
                eng.graph_run()
            for i, (st, b) in enumerate(zip(stores, bits)):
                d, o, l = st.arrays()
                S = oracle.Spectrum(d, o, l, k, b, ec, 5.0, 1)
                seen, solid = eng.bloom(i)
                oseen, osolid = S.bloom()
                assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid)
                keys, cnt, mask = eng.counts(i)
                okeys, ocnt, omask = S.counts()
                assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt) and np.array_equal(mask, omask)
                assert eng.info(i)["cutoff"] == S.cutoff
    finally:
        eng.close()


def test_sparse_mode_keeps_atomics(oracle):
    """A 2^26-bit tier for 4000 reads: every k-mer touches under 1/32 of the entries, so the
    first setters stay device atomics that also set the seen bits."""
    k, bits = 21, 1 << 26
    store, _ = synth.uniform_genome_pairs(seed=977, genome_size=15000, n_pairs=2000, read_len=150, p_start=0.004)
    d, o, l = store.arrays()
    S = oracle.Spectrum(d, o, l, k, bits, False, 5.0, 1)
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits), timing=True)
    try:
        eng.upload(store)
        eng.run()
        assert not sorted_launched(eng)
        seen, solid = eng.bloom()
        oseen, osolid = S.bloom()
        assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid)
        assert np.array_equal(eng.counts()[0], S.counts()[0])
    finally:
        eng.close()


def test_1mb_30x_paths_agree():
    """The full 1 Mb genome / 30x workload on a 2^26-bit tier, dense: 4096 partitions; both
    setter paths give identical tiers and tables."""
    store, _ = synth.workload_1mb_30x()
    out = []
    for path in (0, 1):
        say(f"case 1mb_30x setter path={path}")
        eng = SpectrumEngine(AssemblerConfig(k=33, bloom_bits=1 << 26), timing=True)
        try:
            eng.set_tune(setter_path=path)
            eng.upload(store)
            info = eng.run()
            assert sorted_launched(eng) == (path == 0)
            out.append((info["n_distinct"], info["n_solid"], info["cutoff"], *eng.bloom(), *eng.counts()))
        finally:
            eng.close()
    a, b = out
    assert a[:3] == b[:3]
    assert all(np.array_equal(x, y) for x, y in zip(a[3:], b[3:]))


def test_group_over_2_31_entries_without_the_mark():
    """Two libraries with 2^31 - 2^20-bit tiers share one first-setter array of 2^32 - 2^21
    entries (2^18 - 2^7 partitions): the partition keys then reach bit 63, so the items
    carry no counted-twice mark (k_fs_min writes cleared solid words, k_promote takes every
    k-mer).  The 5 Mb / 100x reads in both libraries keep it dense (98M distinct k-mers);
    both setter paths give identical tiers, tables and cutoffs."""
    store, _ = synth.workload_5mb_100x(seed=3)
    bits = (1 << 31) - (1 << 20)
    out = []
    for path in (0, 1):
        say(f"case two 2^31-bit libraries setter path={path}")
        eng = SpectrumEngine(AssemblerConfig(k=21), timing=True)
        try:
            eng.set_tune(setter_path=path)
            eng.add_library(store, bloom_bits=bits)
            eng.add_library(store, bloom_bits=bits)
            eng.run()
            assert sorted_launched(eng) == (path == 0)
            res = []
            for i in range(2):
                info = eng.info(i)
                res.append(((info["n_distinct"], info["n_solid"], info["cutoff"]), eng.bloom(i), eng.counts(i)))
            out.append(res)
        finally:
            eng.close()
    for (sa, ba, ca), (sb, bb, cb) in zip(*out):
        assert sa == sb
        assert all(np.array_equal(x, y) for x, y in zip(ba, bb))
        assert all(np.array_equal(x, y) for x, y in zip(ca, cb))
    assert out[0][0][0] == out[0][1][0]  # the same reads give the same summary in both libraries
