"""GPU: tuning settings, the autotuner and HIP-graph replay give results identical to the
oracle (and to the eager run): every launch shape and load factor, several libraries of
uneven read lengths with priors, Hammer on and off.  A graph is replayed several times
with eager exports and synchronisations between replays, and every replay is checked."""
import sys

import numpy as np
import pytest

from kc import AssemblerConfig, PackedReads, SpectrumEngine, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _uneven_library(seed, genome, pairs, min_len=50, max_len=250, err=0.003):
    rng = np.random.default_rng(seed)
    st, meta = synth.uniform_genome_pairs(seed=seed, genome_size=genome, n_pairs=pairs, read_len=max_len,
                                          insert_mean=max_len * 2, insert_sd=20, p_start=err)
    out = PackedReads()
    for i in range(len(st)):
        r = st.get_read(i)
        out.pack_clean(r[: int(rng.integers(min_len, max_len + 1))])
    return out, "".join("ACGT"[int(x)] for x in meta["genome"])


def _snapshot(eng, n):
    return [(eng.info(i), eng.bloom(i), eng.counts(i)) for i in range(n)]


def _same(a, b):
    for (ia, ba, ca), (ib, bb, cb) in zip(a, b):
        for key in ("n_occ", "n_distinct", "n_counts", "n_solid", "cutoff", "top_cov", "corrected_reads"):
            assert ia[key] == ib[key], key
        assert all(np.array_equal(x, y) for x, y in zip(ba, bb))
        assert all(np.array_equal(x, y) for x, y in zip(ca, cb))


def _oracle_check(oracle, eng, libs, k, bits, ec, mcs, priors):
    for i, st in enumerate(libs):
        d, o, l = st.arrays()
        S = oracle.Spectrum(d, o, l, k, bits, ec, mcs[i], 1, priors=priors.get(i))
        keys, cnt, mask = eng.counts(i)
        okeys, ocnt, omask = S.counts()
        seen, solid = eng.bloom(i)
        oseen, osolid = S.bloom()
        assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid)
        assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt) and np.array_equal(mask, omask)
        assert eng.info(i)["cutoff"] == S.cutoff


@pytest.mark.parametrize("k,ec", [(33, False), (21, True), (55, True), (99, True), (127, False)])
def test_graph_replays_match_eager_and_oracle(oracle, k, ec):
    libs, genomes = zip(*[_uneven_library(700 + i, 2000 + 1700 * i, 120 + 60 * i) for i in range(4)])
    libs = list(libs)
    bits = 1 << 17
    mcs = [5.0, 5.0, 14.0, 5.0]
    priors = {1: [genomes[1][50:700], genomes[1][900:1000].lower()]}
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec))
    for i, st in enumerate(libs):
        eng.add_library(st, bloom_bits=bits, min_coverage=mcs[i])
    eng.set_priors(priors[1], library=1)
    say(f"case graph k={k} ec={ec}")
    eng.run()
    eager = _snapshot(eng, len(libs))
    _oracle_check(oracle, eng, libs, k, bits, ec, mcs, priors)
    for rep in range(4):  # call 0 plans and captures, calls 1.. replay
        say(f"  graph call {rep}")
        info = eng.graph_run()
        assert info["n_occ"] == eager[0][0]["n_occ"]
        _same(_snapshot(eng, len(libs)), eager)
    # an eager run in between drops the graph; the next call plans again
    eng.run()
    _same(_snapshot(eng, len(libs)), eager)
    eng.graph_run()
    eng.graph_run()
    _same(_snapshot(eng, len(libs)), eager)
    eng.close()


@pytest.mark.parametrize("wpt,bpc,load,kpb,ip", [(1, 0, 0.5, 0, 1), (2, 8, 0.35, 0, 1), (4, 16, 0.8, 0, 1),
                                               (4, 1, 0.65, 0, 1), (2, 0, 0.9, 0, 1), (1, 0, 0.0, 64, 0),
                                               (1, 4, 0.0, 192, 0), (1, 16, 0.0, 640, 0), (1, 1, 0.0, 4096, 0)])
def test_tune_settings_are_exact(oracle, wpt, bpc, load, kpb, ip):
    """Launch shapes and load factors of the table insert (ip 1), and bin sizes and
    workgroups per CU of the binned count (ip 0): every setting equals the oracle."""
    libs, genomes = zip(*[_uneven_library(40 + i, 3000 + 900 * i, 200, err=0.006) for i in range(3)])
    libs = list(libs)
    k, bits = 31, (1 << 16) + 64
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits))
    for st in libs:
        eng.add_library(st, bloom_bits=bits)
    eng.set_tune(windows_per_thread=wpt, blocks_per_cu=bpc, table_load=load, kmers_per_bin=kpb, insert_path=ip)
    assert eng.tune()["windows_per_thread"] == wpt and eng.tune()["kmers_per_bin"] == kpb
    say(f"case tune wpt={wpt} bpc={bpc} load={load} kpb={kpb} ip={ip}")
    for rep in range(2):  # the second run sizes its bins from the first's exact count
        eng.run()
        _oracle_check(oracle, eng, libs, k, bits, False, [5.0] * 3, {})
    eng.close()


def test_tune_rejects_bad_bin_size():
    eng = SpectrumEngine(AssemblerConfig(k=21, bloom_bits=1 << 16))
    for bad in (1, 63, 4097):
        with pytest.raises(Exception, match="kmers_per_bin"):
            eng.set_tune(kmers_per_bin=bad)
    eng.close()


def test_autotune_keeps_results(oracle):
    libs, _ = zip(*[_uneven_library(90 + i, 4000, 300) for i in range(2)])
    libs = list(libs)
    k, bits = 33, 1 << 18
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=True))
    for st in libs:
        eng.add_library(st, bloom_bits=bits)
    say("case autotune")
    best = eng.autotune(reps=1)
    assert best["windows_per_thread"] in (1, 2, 4) and best["ms"] > 0
    assert eng.tune() == best
    eng.run()
    _oracle_check(oracle, eng, libs, k, bits, True, [5.0] * 2, {})
    eng.graph_run()
    eng.graph_run()
    _oracle_check(oracle, eng, libs, k, bits, True, [5.0] * 2, {})
    eng.close()


def test_autotune_on_the_binned_path(oracle):
    """Without Hammer the input takes the binned count: the search covers its bin size and
    workgroups per CU and the first-setter path, and the chosen setting stays exact,
    eagerly and replayed."""
    libs, _ = zip(*[_uneven_library(95 + i, 5000, 300) for i in range(3)])
    libs = list(libs)
    k, bits = 33, 1 << 18
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits), timing=True)
    for st in libs:
        eng.add_library(st, bloom_bits=bits)
    say("case autotune binned")
    best = eng.autotune(reps=1)
    assert best["ms"] > 0 and best["kmers_per_bin"] in (0, 192, 256, 352, 448, 640)
    assert best["minimizer_len"] in (0, 11, 13, 16, 20, 24, 28, 32)
    assert best["blocks_per_cu"] in (0, 4, 8, 16) and best["setter_path"] in (0, 1)
    assert eng.tune() == best
    eng.run()
    assert "k_bin_count" in eng.kernel_times()
    _oracle_check(oracle, eng, libs, k, bits, False, [5.0] * 3, {})
    eng.graph_run()
    eng.graph_run()
    _oracle_check(oracle, eng, libs, k, bits, False, [5.0] * 3, {})
    eng.close()


@pytest.mark.parametrize("k,ec,bits", [(21, False, 1 << 20), (21, True, 1 << 16), (11, False, 1 << 14),
                                       (31, False, (1 << 18) + 64), (32, True, 1 << 20), (33, False, 1 << 20),
                                       (55, True, 4096), (63, False, 1 << 19), (5, False, 64 * 3), (1, False, 64)])
def test_insert_paths_are_exact(oracle, k, ec, bits):
    """Stage 2 through the minimizer-binned count (super-k-mers sorted by bin, each bin
    counted in LDS, the table packed) and per-window table atomics: both equal the oracle,
    on one library.
    With Hammer, Stage 3's recount re-lays the packed table out for hashing."""
    store, _ = synth.uniform_genome_pairs(seed=300 + k, genome_size=20000, n_pairs=2000, read_len=150, p_start=0.005)
    d, o, l = store.arrays()
    S = oracle.Spectrum(d, o, l, k, bits, ec, 5.0, 1)
    okeys, ocnt, omask = S.counts()
    oseen, osolid = S.bloom()
    distinct = set()
    for path in (3, 1):
        say(f"case insert path={path} k={k} ec={ec}")
        eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec), timing=True)
        eng.set_tune(insert_path=path)
        eng.upload(store)
        info = eng.run()
        binned = eng.kernel_times().get("k_bin_count", (0.0, 0))[1] > 0
        assert binned == (path == 3)
        keys, cnt, mask = eng.counts()
        seen, solid = eng.bloom()
        distinct.add(info["n_distinct"])
        assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid)
        assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt) and np.array_equal(mask, omask)
        assert info["cutoff"] == S.cutoff and info["n_solid"] == S.n_solid
        if ec:
            assert np.array_equal(eng.reads(), S.corrected_data())
        # a second run on the same context (buffers reused) gives the same tables
        eng.run()
        assert np.array_equal(eng.counts()[0], okeys) and np.array_equal(eng.bloom()[1], osolid)
        eng.close()
    assert len(distinct) == 1  # both paths hold the same distinct k-mers


def test_tune_rejects_bad_settings():
    from kc import KcError
    eng = SpectrumEngine(AssemblerConfig(k=21, bloom_bits=1 << 12))
    for bad in ({"windows_per_thread": 3}, {"table_load": 0.99}, {"blocks_per_cu": 5000}, {"insert_path": 4},
                {"insert_path": 2}):
        with pytest.raises(KcError):
            eng.set_tune(**bad)
    eng.close()


def test_multik_over_resident_reads(oracle):
    """multik.rs:109-178: k = 21, 33, 55, 77, 127 over one upload, with the previous k's
    sequences injected as priors; then k back to 21 (buffers sized for larger k reused)."""
    store, meta = synth.uniform_genome_pairs(seed=77, genome_size=6000, n_pairs=500, read_len=150,
                                             insert_mean=300, insert_sd=20, p_start=0.004)
    g = "".join("ACGT"[int(x)] for x in meta["genome"])
    d, o, l = store.arrays()
    bits = 1 << 18
    eng = SpectrumEngine(AssemblerConfig(k=21, bloom_bits=bits, error_correct=True))
    eng.upload(store)
    for k, priors in ((21, None), (33, [g[0:800]]), (55, [g[0:800], g[3000:4000]]), (77, [g[100:900]]),
                      (127, [g[2000:2600], g[5000:5200]]), (21, None)):
        say(f"case multik k={k}")
        eng.set_k(k)
        eng.set_priors(priors or [])
        info = eng.run()
        S = oracle.Spectrum(d, o, l, k, bits, True, 5.0, 1, priors=priors)
        assert info["n_occ"] == S.n_occ
        keys, cnt, mask = eng.counts()
        okeys, ocnt, omask = S.counts()
        assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt) and np.array_equal(mask, omask)
        assert np.array_equal(eng.reads(), S.corrected_data())
    eng.close()


def test_sparse_first_setters(oracle):
    """Libraries whose k-mers touch a small part of large Bloom tiers take the sparse
    first-setter path (the first-setter array stays all ~0 between runs, only touched
    entries are reset): repeated eager runs, a change of k, graph replays, and a switch to
    small tiers (the dense path) and back all match the oracle."""
    libs, genomes = zip(*[_uneven_library(500 + i, 2000 + 800 * i, 80 + 30 * i) for i in range(4)])
    libs = list(libs)
    big, small = 1 << 22, 1 << 12
    eng = SpectrumEngine(AssemblerConfig(k=33, bloom_bits=big, error_correct=True))
    for st in libs:
        eng.add_library(st, bloom_bits=big)
    mcs = [5.0] * len(libs)
    for k in (33, 33, 21):
        say(f"case sparse k={k}")
        eng.set_k(k)
        eng.run()
        _oracle_check(oracle, eng, libs, k, big, True, mcs, {})
    for rep in range(3):
        say(f"  sparse graph call {rep}")
        eng.graph_run()
        _oracle_check(oracle, eng, libs, 21, big, True, mcs, {})
    eng.close()
    # the same libraries on small tiers (dense), then large again in a fresh context
    eng = SpectrumEngine(AssemblerConfig(k=21, bloom_bits=small))
    for st in libs:
        eng.add_library(st, bloom_bits=small)
    eng.run()
    _oracle_check(oracle, eng, libs, 21, small, False, mcs, {})
    eng.close()


@pytest.mark.parametrize("k,bits,nlib", [(21, 1 << 18, 1), (33, 1 << 17, 4), (55, (1 << 16) + 64, 3), (63, 1 << 19, 2)])
def test_graph_replays_the_binned_count(oracle, k, bits, nlib):
    """Without Hammer or priors a captured graph runs the minimizer-binned count and the
    sorted first setters with the sizes of the run that planned it (guarded on the device;
    rocPRIM's memset nodes become fill kernels): every replay equals the eager run and the
    oracle."""
    libs, _ = zip(*[_uneven_library(1200 + 7 * i + k, 3000 + 2500 * i, 200 + 60 * i) for i in range(nlib)])
    libs = list(libs)
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits), timing=True)
    for st in libs:
        eng.add_library(st, bloom_bits=bits)
    say(f"case binned graph k={k} libraries={nlib}")
    eng.run()
    assert eng.kernel_times().get("k_bin_count", (0.0, 0))[1] > 0
    eager = _snapshot(eng, nlib)
    _oracle_check(oracle, eng, libs, k, bits, False, [5.0] * nlib, {})
    for rep in range(4):
        say(f"  graph call {rep}")
        eng.graph_run()
        _same(_snapshot(eng, nlib), eager)
    eng.close()
