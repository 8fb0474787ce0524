"""Pins the oracle's FASTQ ingestion (packed_reads.rs from_paired_fastq / append_from_file:
FastxReader, the three N gates and the three-block layout) against the reference's own
three-block test vectors and against the host mirror (kc.packed_reads), which follows the
same source independently, on random files with every gate exercised."""
import numpy as np
import pytest

from kc import packed_reads as P


def fastq(reads, crlf=False, blank_between=False, lead_blank=False):
    eol = b"\r\n" if crlf else b"\n"
    out = [eol] if lead_blank else []
    for i, r in enumerate(reads):
        if blank_between and i:
            out.append(eol)
        out.append(b"@r%d extra" % i + eol + r + eol + b"+" + eol + b"I" * len(r) + eol)
    return b"".join(out)


def fasta(reads, width=0, crlf=False, blank_inside=False, lead_blank=False):
    """FASTA text of the reads: a '>' header each, the sequence in lines of `width` bases
    (one line when 0; none for an empty read), optionally with empty lines inside."""
    eol = b"\r\n" if crlf else b"\n"
    out = [eol, b"  " + eol] if lead_blank else []
    for i, r in enumerate(reads):
        out.append(b">r%d extra" % i + eol)
        w = width or max(len(r), 1)
        for j in range(0, len(r), w):
            out.append(r[j:j + w] + eol)
            if blank_inside and j % (3 * w) == 0:
                out.append(eol)
    return b"".join(out)


def fasta_reads(reads):
    """The records FastxReader returns for fasta(reads): a last record without bases is none."""
    return reads[:-1] if reads and not reads[-1] else reads


def store_of(res):
    data, offs, lens, pe, amb, _, _ = res
    return P.PackedReads.from_arrays(data, offs, lens, pe_count=pe), amb


def mirror(r1, r2=None):
    if r2 is None:
        s = P.PackedReads()
        for r in r1:
            s.ingest_read(r)
        return s
    return P.PackedReads.from_paired_slices(r1, r2)


def test_reference_three_block_vectors(oracle):
    # packed_reads.rs:847-890 and tests/test_packed_reads.rs:259-318, through FASTQ text
    r1 = [b"AAAAACCCCCGGGGGTTTTTAAAA", b"ACGTACGTACGTACGTACGTACNNACGTACGTACGTACGTACGTAC", b"NNNCCCCCGGGGTTTTTAAAAACCCCCNN"]
    r2 = [b"TTTTTAAAAAGGGGGCCCCCTTTT", b"GGGGGGGGGGGGCCCCCCCCCCCC", b"NNNGGGGGCCCCTTTTTAAAAAGGGGGNN"]
    s, amb = store_of(oracle.pack_fastq(fastq(r1), fastq(r2)))
    assert s.pe_count == 2 and len(s) == 7 and amb == 0
    assert s.get_pair(0) == (r1[0], r2[0])
    assert s.get_pair(1) == (b"CCCCCGGGGTTTTTAAAAACCCCC", b"GGGGGCCCCTTTTTAAAAAGGGGG")
    assert [s.get_read(i) for i in (4, 5, 6)] == [b"ACGTACGTACGTACGTACGTAC"] * 2 + [r2[1]]
    p = [(b"ACGTACGTACGTACGTACGTACGT", b"TGCATGCATGCATGCATGCATGCA"),
         (b"NNAAACCCGGGTTTAAACCCGGGTTTNN", b"TTTGGGCCCAAATTTGGGCCCAAA"),
         (b"ACGTACGTACGTACGTACGTACNNNGTGCATGCATGCATGCATGCAT", b"CCCCCCCCCCCCCCCCCCCCCCCC"),
         (b"GGGGGGGGGGGGGGGGGGGGGGGG", b"AAAAAAAAAAAAAAAAAAAAAAAA")]
    s, _ = store_of(oracle.pack_fastq(fastq([a for a, _ in p]), fastq([b for _, b in p])))
    assert s.pe_count == 3 and len(s) == 9
    assert s.get_pair(0) == p[0] and s.get_pair(2) == p[3]
    assert s.get_pair(1) == (b"AAACCCGGGTTTAAACCCGGGTTT", p[1][1])


def random_reads(rng, n):
    out = []
    for _ in range(n):
        L = int(rng.integers(0, 160))
        r = bytearray(rng.choice(list(b"ACGTacgt"), L).astype(np.uint8).tobytes())
        kind = int(rng.integers(0, 6))
        if kind == 1 and L:
            m = int(rng.integers(1, min(L, 4) + 1))
            r[:m] = b"N" * m
        if kind == 2 and L > 30:
            r[int(rng.integers(5, L - 5))] = ord("N")
        if kind == 3 and L:
            m = int(rng.integers(1, min(L, 6) + 1))
            r[L - m:] = b"n" * m
        if kind == 4 and L > 40:
            r[int(rng.integers(0, L))] = int(rng.choice(list(b"RYKMnx.")))
        out.append(bytes(r))
    return out


@pytest.mark.parametrize("seed,crlf,blank,lead,unequal", [
    (1, False, False, False, 0), (2, True, False, False, 0), (3, False, True, True, 0), (4, False, False, False, 7),
    (5, True, True, False, -5)])
def test_random_pairs_match_host_mirror(oracle, seed, crlf, blank, lead, unequal):
    rng = np.random.default_rng(seed)
    r1 = random_reads(rng, 400)
    r2 = random_reads(rng, 400 + unequal)
    res = oracle.pack_fastq(fastq(r1, crlf, blank, lead), fastq(r2, crlf, blank, lead))
    s, amb = store_of(res)
    m = mirror(r1, r2)
    assert res[5] == len(r1) and res[6] == len(r2)
    assert s.pe_count == m.pe_count and len(s) == len(m) and amb == len(m.ambiguous_sidecar)
    d, o, l = m.arrays()
    assert np.array_equal(res[0], d) and np.array_equal(res[1], o) and np.array_equal(res[2], l)


def test_single_file_ingest_matches_host_mirror(oracle):
    rng = np.random.default_rng(9)
    r = random_reads(rng, 500)
    res = oracle.pack_fastq(fastq(r))
    m = mirror(r)
    d, o, l = m.arrays()
    assert res[3] == 0 and np.array_equal(res[0], d) and np.array_equal(res[1], o) and np.array_equal(res[2], l)


@pytest.mark.parametrize("text", [b"@r\nACGT\n+\nII\n", b"@r\nACGT\n-\nIIII\n", b"@r\nACGT\n", b"xr\nACGT\n+\nIIII\n",
                                  b"@r\nACGT\n+\nIIII\nACGT\n+\nIIII\n"])
def test_malformed_is_an_error(oracle, text):
    with pytest.raises(ValueError):
        oracle.pack_fastq(text, fastq([b"ACGT"]))


def fastq_q(reads, rng, crlf=False):
    """FASTQ with a random quality character per base; returns (text, [(seq, qual), ...])."""
    eol = b"\r\n" if crlf else b"\n"
    out, recs = [], []
    for i, r in enumerate(reads):
        q = bytes(rng.integers(33, 75, len(r)).astype(np.uint8))
        out.append(b"@r%d" % i + eol + r + eol + b"+" + eol + q + eol)
        recs.append((r.upper(), q))
    return b"".join(out), recs


def test_quality_side_array_follows_the_packed_bases(oracle):
    # Not a reference feature (the reference drops qualities): every packed base carries
    # its own quality - 33.  Pinned by locating each output read in its source records.
    rng = np.random.default_rng(31)
    r1, r2 = random_reads(rng, 300), random_reads(rng, 305)
    a, ra = fastq_q(r1, rng, crlf=True)
    b, rb = fastq_q(r2, rng)
    plain = oracle.pack_fastq(a, b)
    data, offs, lens, pe, amb, rec1, rec2, q, qo = oracle.pack_fastq(a, b, quals=True)
    for x, y in zip(plain, (data, offs, lens, pe, amb, rec1, rec2)):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    assert len(q) == int(lens.astype(np.int64).sum())
    assert np.array_equal(qo, np.concatenate([[0], np.cumsum(lens.astype(np.uint64))[:-1]]).astype(np.uint64))
    s = P.PackedReads.from_arrays(data, offs, lens, pe_count=pe)
    recs = ra + rb
    for i in range(len(s)):
        seq = s.get_read(i)
        qual = bytes((q[int(qo[i]):int(qo[i]) + len(seq)] + 33).astype(np.uint8))
        assert any(p >= 0 and src[1][p:p + len(seq)] == qual
                   for src in recs for p in [src[0].find(seq)]), i
    # single-file mode keeps the same pairing of bases and qualities
    d1, o1, l1, *_, q1, qo1 = oracle.pack_fastq(a, quals=True)
    assert len(q1) == int(l1.astype(np.int64).sum())


@pytest.mark.parametrize("seed,width,crlf,blank,lead", [(31, 0, False, False, False), (32, 60, True, True, True),
                                                        (33, 7, False, True, False), (34, 1, True, False, True)])
def test_fasta_pairs_match_host_mirror(oracle, seed, width, crlf, blank, lead):
    """FastxReader's FASTA branch (packed_reads.rs:217-235): single- and multi-line records,
    empty lines inside a sequence, empty records, CRLF, leading blank lines; an empty last
    record is dropped.  Pinned against the host mirror fed the reads themselves."""
    rng = np.random.default_rng(seed)
    r1 = random_reads(rng, 300) + [b""]
    r2 = random_reads(rng, 280)
    a, b = fasta(r1, width, crlf, blank, lead), fasta(r2, width, crlf, blank, lead)
    res = oracle.pack_fastq(a, b)
    assert (res[5], res[6]) == (len(r1) - 1, len(r2))  # records read from each file
    s, _ = store_of(res)
    m = mirror(fasta_reads(r1), fasta_reads(r2))
    assert s.pe_count == m.pe_count and len(s) == len(m)
    assert all(s.get_read(i) == m.get_read(i) for i in range(len(s)))
    # the reference test's multi-line records (tests/test_gfa_and_fastq.rs:72-90), made long
    # enough to pass the gates: lines concatenate in order
    text = b">seq1\n" + b"ACGT" * 6 + b"\n" + b"ACGT" * 6 + b"\n>seq2\nGGGGGGGGGG\nTTTTTTTTTT\nAAAAAAAAAA\n"
    s, amb = store_of(oracle.pack_fastq(text))
    assert [s.get_read(i) for i in range(len(s))] == [b"ACGT" * 12, b"G" * 10 + b"T" * 10 + b"A" * 10] and amb == 0
