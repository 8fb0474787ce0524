"""PackedReads layout (packed_reads.rs): the reference's own tests
(packed_reads.rs:700-891, tests/test_packed_reads.rs) run against the host
store in kc.packed_reads (product) and the oracle's pack_clean/get_read
restatement, and the vectorised packers are checked against pack_clean."""
import numpy as np
import pytest

from kc import packed_reads as P


def lcg_workload():
    # tests/test_packed_reads.rs:41-79
    state = 0x12345678
    M = (1 << 64) - 1
    reads = []
    for _ in range(2000):
        state = (state * 6364136223846793005 + 1) & M
        ln = 50 + state % 200
        seq = bytearray()
        for j in range(ln):
            state = (state * 6364136223846793005 + 1) & M
            seq.append(b"ACGT"[(state ^ j) % 4])
        reads.append(bytes(seq))
    return reads


def modulo4_reads():
    # tests/test_packed_reads.rs:5-38
    return [bytes(b"ACGT"[i % 4] for i in range(n)) for n in range(65)]


@pytest.mark.parametrize("reads", [modulo4_reads(), lcg_workload(),
                                   [b"", b"G", b"TA", b"CGC", b"ACGT", b"ACGTA"],
                                   [b"ACGTACGT", b"TTTTAAAACCCCGGGG", b"A", b"CGA", b"GATTACA"]])
def test_roundtrip_host_and_oracle(oracle, reads):
    store = P.PackedReads()
    for r in reads:
        store.add_read(r)
    assert len(store) == len(reads)
    for i, r in enumerate(reads):
        assert store.get_read(i) == r
    d, o, l = oracle.pack(reads)
    assert np.array_equal(store.data, d) and np.array_equal(store.offsets, o) and np.array_equal(store.lengths, l)
    for i, r in enumerate(reads):
        assert oracle.get_read(d, o[i], l[i]) == r


def test_vectorised_packers_match_pack_clean(oracle):
    reads = lcg_workload()
    codes = np.concatenate([np.frombuffer(r, np.uint8) for r in reads])
    codes = P._ASCII_TO_2BIT[codes]
    d, o = P.pack_codes(codes, [len(r) for r in reads])
    d2, o2, _ = oracle.pack(reads)
    assert np.array_equal(d, d2) and np.array_equal(o, o2)
    rng = np.random.default_rng(3)
    fixed = rng.integers(0, 4, (300, 150), dtype=np.uint8)
    d3, o3 = P.pack_fixed_length(fixed)
    d4, o4, _ = oracle.pack([bytes(b"ACGT"[c] for c in row) for row in fixed])
    assert np.array_equal(d3, d4) and np.array_equal(o3, o4)


def test_lowercase_and_invalid_bytes(oracle):
    # pack_clean skips non-ACGT bytes in release builds (packed_reads.rs:555-562)
    store = P.PackedReads()
    store.pack_clean(b"acgtNacgt")
    assert store.get_read(0) == b"ACGTACGT"
    d, o, l = oracle.pack([b"acgtNacgt"])
    assert np.array_equal(store.data, d) and int(l[0]) == 8


def test_n_gates():
    # packed_reads.rs:724-840, tests/test_packed_reads.rs:120-258
    assert P._trim_n_ends(b"NNNACGT") == b"ACGT"
    assert P._trim_n_ends(b"ACGTNNN") == b"ACGT"
    assert P._trim_n_ends(b"NNACGTNN") == b"ACGT"
    assert P._trim_n_ends(b"NNNNN") == b""
    assert P._split_on_n(b"ACGTNNNACGT") == [b"ACGT", b"ACGT"]
    assert P._split_on_n(b"ACGTNACGTNACGT") == [b"ACGT"] * 3
    s = P.PackedReads()
    s.ingest_read(b"ACGTACGTACGTACGTACGTANNN")
    assert len(s) == 1 and not s.ambiguous_sidecar and s.get_read(0) == b"ACGTACGTACGTACGTACGTA"
    s = P.PackedReads()
    s.ingest_read(b"NNNACGTACGTACGTACGTACGTA")
    assert len(s) == 1 and s.get_read(0) == b"ACGTACGTACGTACGTACGTA"
    s = P.PackedReads()
    s.ingest_read(b"ACGTACGTACGTACGTACGTANNNACGTACGTACGTACGTACGTA")
    assert len(s) == 2 and not s.ambiguous_sidecar
    s = P.PackedReads()
    s.ingest_read(b"ACGTN" + b"ACGTACGTACGTACGTACGTA")
    assert len(s) == 1 and not s.ambiguous_sidecar
    s = P.PackedReads()
    s.ingest_read(b"NNNNNN")
    assert len(s) == 0 and not s.ambiguous_sidecar
    s = P.PackedReads()
    s.ingest_read(b"ACGTNNCGTA")
    assert len(s) == 0 and len(s.ambiguous_sidecar) == 1
    s = P.PackedReads()
    s.ingest_read(b"ACGTNACGT")
    assert len(s) == 0 and b"N" in s.ambiguous_sidecar[0]


def test_three_block_paired_synchronization():
    # packed_reads.rs:847-890
    r1_0, r2_0 = b"AAAAACCCCCGGGGGTTTTTAAAA", b"TTTTTAAAAAGGGGGCCCCCTTTT"
    r1_1 = b"ACGTACGTACGTACGTACGTACNNACGTACGTACGTACGTACGTAC"
    r2_1 = b"GGGGGGGGGGGGCCCCCCCCCCCC"
    r1_2, r2_2 = b"NNNCCCCCGGGGTTTTTAAAAACCCCCNN", b"NNNGGGGGCCCCTTTTTAAAAAGGGGGNN"
    s = P.PackedReads.from_paired_slices([r1_0, r1_1, r1_2], [r2_0, r2_1, r2_2])
    assert s.pe_count == 2 and len(s) == 7
    assert s.get_pair(0) == (r1_0, r2_0)
    assert s.get_pair(1) == (b"CCCCCGGGGTTTTTAAAAACCCCC", b"GGGGGCCCCTTTTTAAAAAGGGGG")
    assert s.get_read(4) == b"ACGTACGTACGTACGTACGTAC"
    assert s.get_read(5) == b"ACGTACGTACGTACGTACGTAC"
    assert s.get_read(6) == r2_1


def test_three_block_paired_synchronization_and_trimming():
    # tests/test_packed_reads.rs:259-318
    p0 = (b"ACGTACGTACGTACGTACGTACGT", b"TGCATGCATGCATGCATGCATGCA")
    p1 = (b"NNAAACCCGGGTTTAAACCCGGGTTTNN", b"TTTGGGCCCAAATTTGGGCCCAAA")
    p2 = (b"ACGTACGTACGTACGTACGTACNNNGTGCATGCATGCATGCATGCAT", b"CCCCCCCCCCCCCCCCCCCCCCCC")
    p3 = (b"GGGGGGGGGGGGGGGGGGGGGGGG", b"AAAAAAAAAAAAAAAAAAAAAAAA")
    s = P.PackedReads.from_paired_slices([p[0] for p in (p0, p1, p2, p3)], [p[1] for p in (p0, p1, p2, p3)])
    assert s.pe_count == 3 and len(s) == 9
    assert s.get_pair(0) == p0
    assert s.get_pair(1) == (b"AAACCCGGGTTTAAACCCGGGTTT", p1[1])
    assert s.get_pair(2) == p3
