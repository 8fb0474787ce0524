"""GPU parity of the 64-bit event path (the one a library of 2^30 - 1 windows or more
takes, assemble.rs:136-157 order over u64 window numbers), forced on smaller inputs with
KC_FLAG_EVENTS64: every 64-bit instantiation (table slots with u64 first events, the
binned count, the atomic first setters into a u64 array, Hammer's recount, graph replay)
against the oracle, and at full size against the recorded digests."""
import json
import os
import sys

import numpy as np
import pytest

from kc import AssemblerConfig, SpectrumEngine, synth
from kc.digest import diff, engine_digests

from conftest import GOLDEN
from test_gpu_graph import _oracle_check, _same, _snapshot, _uneven_library

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("k,ec,bits", [(21, False, 1 << 20), (21, True, 1 << 16), (33, False, (1 << 18) + 64),
                                       (55, True, 4096), (55, False, 1 << 19), (77, True, 1 << 18),
                                       (127, False, 1 << 17), (31, False, 64)])
def test_events64_matches_oracle(oracle, k, ec, bits):
    store, _ = synth.uniform_genome_pairs(seed=600 + k, genome_size=20000, n_pairs=2000, read_len=150, p_start=0.005)
    d, o, l = store.arrays()
    say(f"case events64 k={k} ec={ec} bits={bits}")
    S = oracle.Spectrum(d, o, l, k, bits, ec, 5.0, 1)
    for ip in (0, 1):  # binned count (k <= 63) and the table insert
        eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec), events64=True)
        eng.set_tune(insert_path=ip)
        eng.upload(store)
        for rep in range(2):
            info = eng.run()
            seen, solid = eng.bloom()
            keys, cnt, mask = eng.counts()
            oseen, osolid = S.bloom()
            okeys, ocnt, omask = S.counts()
            assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid)
            assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt) and np.array_equal(mask, omask)
            assert info["cutoff"] == S.cutoff and info["n_solid"] == S.n_solid
            if ec:
                assert info["corrected_reads"] == S.corrected_reads
                assert np.array_equal(eng.reads(), S.corrected_data())
        eng.close()


@pytest.mark.parametrize("k,ec", [(33, False), (55, True)])
def test_events64_libraries_and_graph_replay(oracle, k, ec):
    libs, genomes = zip(*[_uneven_library(800 + i, 2500 + 1500 * i, 150 + 50 * i) for i in range(3)])
    libs = list(libs)
    bits = 1 << 17
    mcs = [5.0, 12.0, 5.0]
    priors = {2: [genomes[2][100:900]]}
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec), events64=True)
    for i, st in enumerate(libs):
        eng.add_library(st, bloom_bits=bits, min_coverage=mcs[i])
    eng.set_priors(priors[2], library=2)
    say(f"case events64 libraries k={k} ec={ec}")
    eng.run()
    eager = _snapshot(eng, len(libs))
    _oracle_check(oracle, eng, libs, k, bits, ec, mcs, priors)
    for rep in range(3):
        say(f"  graph call {rep}")
        eng.graph_run()
        _same(_snapshot(eng, len(libs)), eager)
    eng.close()


DIGESTS = json.load(open(os.path.join(GOLDEN, "full_size_digests.json")))


@pytest.mark.parametrize("case", ["1mb_30x_k21", "1mb_30x_k21_hammer", "5mb_100x_k55"])
def test_events64_full_size_digests(case):
    want = DIGESTS[case]
    store = (synth.workload_5mb_100x(seed=2) if want["workload"] == "5mb_100x" else synth.workload_1mb_30x(seed=1))[0]
    say(f"case events64 full size {case}")
    eng = SpectrumEngine(AssemblerConfig(k=want["k"], bloom_bits=want["bloom_bits"],
                                         error_correct=want["error_correct"]), events64=True)
    eng.upload(store)
    info = eng.run()
    got = engine_digests(eng, info, with_reads=want["error_correct"])
    eng.close()
    assert not diff(got, want), diff(got, want)
