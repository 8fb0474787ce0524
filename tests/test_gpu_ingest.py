"""GPU: FASTQ ingestion on the device against the oracle's restatement of
from_paired_fastq / append_from_file (FastxReader, the three N gates, the three-block
layout): bit-exact data, offsets, lengths, paired-block size and sidecar count."""
import gzip
import os
import sys

import numpy as np
import pytest

from kc import KcError, _abi
from kc.ingest import fastq_pack
from test_oracle_fastq import fasta, fastq, random_reads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check(oracle, a, b=None):
    data, offs, lens, pe, amb, rec1, rec2 = oracle.pack_fastq(a, b)
    store, info = fastq_pack(a, b)
    d, o, l = store.arrays()
    assert info["r1_records"] == rec1 and info["r2_records"] == rec2
    assert info["pe_count"] == pe and info["n_ambiguous"] == amb
    assert np.array_equal(d, data) and np.array_equal(o, offs) and np.array_equal(l, lens)
    return info


@pytest.mark.parametrize("seed,crlf,lead,unequal,n", [(1, False, False, 0, 500), (2, True, True, 0, 800),
                                                       (3, False, False, 9, 300), (4, True, False, -4, 1000),
                                                       (5, False, True, 0, 0), (6, False, False, 0, 1)])
def test_pairs_match_oracle(oracle, seed, crlf, lead, unequal, n):
    rng = np.random.default_rng(seed)
    r1 = random_reads(rng, n)
    r2 = random_reads(rng, max(n + unequal, 0))
    a = fastq(r1, crlf, False, lead) + (b"\n\n" if seed % 2 else b"")  # trailing blank lines
    b = fastq(r2, crlf, False, lead)
    say(f"case ingest seed={seed} n={n}")
    info = check(oracle, a, b)
    assert info["r1_records"] == len(r1)


def test_single_file_and_no_final_newline(oracle):
    rng = np.random.default_rng(8)
    a = fastq(random_reads(rng, 700))
    check(oracle, a)
    check(oracle, a[:-1])  # the last quality line without its newline


def test_empty_last_record(oracle):
    # the last record's sequence and quality lines are empty, with and without blank lines after
    a = fastq([b"ACGT" * 10, b"CCGGA" * 9, b""])
    b = fastq([b"TTGCA" * 8, b"GGCA" * 11, b"AC"])
    check(oracle, a, b)
    check(oracle, a + b"\n\n", b + b"\r\n")
    check(oracle, b"@only\n\n+\n\n")


def test_ecoli_fixture(oracle):
    from conftest import GOLDEN
    a = gzip.open(os.path.join(GOLDEN, "ecoli_1K_1.fq.gz"), "rb").read()
    b = gzip.open(os.path.join(GOLDEN, "ecoli_1K_2.fq.gz"), "rb").read()
    info = check(oracle, a, b)
    assert info["pe_count"] > 0


def test_large_synthetic_pairs(oracle):
    rng = np.random.default_rng(11)
    n = 60000
    r1 = [bytes(x) for x in rng.choice(list(b"ACGT"), (n, 150)).astype(np.uint8)]
    r2 = [bytes(x) for x in rng.choice(list(b"ACGT"), (n, 150)).astype(np.uint8)]
    for i in range(0, n, 97):
        r1[i] = r1[i][:70] + b"N" + r1[i][71:]
    check(oracle, fastq(r1), fastq(r2))


def test_context_reuse_shrinking_inputs(oracle):
    # one context, a large input then smaller ones: the grown buffers keep no stale state
    from kc.ingest import FastqIngest
    ctx = FastqIngest()
    for seed, n in ((21, 3000), (22, 40), (23, 700), (24, 0)):
        rng = np.random.default_rng(seed)
        a, b = fastq(random_reads(rng, n)), fastq(random_reads(rng, n + seed % 3))
        say(f"case ingest context seed={seed} n={n}")
        data, offs, lens, pe, amb, rec1, rec2 = oracle.pack_fastq(a, b)
        store, info = ctx.pack(a, b)
        d, o, l = store.arrays()
        assert (info["pe_count"], info["n_ambiguous"], info["r1_records"], info["r2_records"]) == (pe, amb, rec1, rec2)
        assert np.array_equal(d, data) and np.array_equal(o, offs) and np.array_equal(l, lens)
        assert info["ms_device"] >= 0
    # single-file mode after paired calls on the same context
    a = fastq(random_reads(np.random.default_rng(25), 500))
    data, offs, lens, *_ = oracle.pack_fastq(a)
    d, o, l = ctx.pack(a)[0].arrays()
    assert np.array_equal(d, data) and np.array_equal(o, offs) and np.array_equal(l, lens)
    ctx.close()


def test_quality_side_array_matches_oracle(oracle):
    from kc.ingest import FastqIngest
    from test_oracle_fastq import fastq_q
    rng = np.random.default_rng(41)
    ctx = FastqIngest(keep_quals=True)
    for n, crlf in ((700, False), (1500, True), (0, False)):
        a, _ = fastq_q(random_reads(rng, n), rng, crlf)
        b, _ = fastq_q(random_reads(rng, n + 3), rng)
        say(f"case ingest quals n={n}")
        data, offs, lens, pe, amb, rec1, rec2, q, qo = oracle.pack_fastq(a, b, quals=True)
        store, info = ctx.pack(a, b)
        d, o, l = store.arrays()
        assert np.array_equal(d, data) and np.array_equal(o, offs) and np.array_equal(l, lens)
        assert info["n_quals"] == len(q)
        assert np.array_equal(info["quals"], q) and np.array_equal(info["qoff"], qo)
    a, _ = fastq_q(random_reads(rng, 400), rng)
    *_, q, qo = oracle.pack_fastq(a, quals=True)
    _, info = ctx.pack(a)
    assert np.array_equal(info["quals"], q) and np.array_equal(info["qoff"], qo)
    ctx.close()
    # without keep_quals there is no quality array
    _, info = fastq_pack(a)
    assert info["n_quals"] == 0


def test_unsupported_and_malformed():
    for bad in (b"@r\nACGT\n+\nII\n", b"@r\nACGT\n-\nIIII\n", b"xr\nACGT\n+\nIIII\n", b"@r\nACGT\n+\n"):
        with pytest.raises(KcError) as e:
            fastq_pack(bad, fastq([b"ACGT"]))
        assert e.value.status == _abi.KC_EINVAL, bad


@pytest.mark.parametrize("seed,width,crlf,blank,lead", [(41, 0, False, False, False), (42, 60, True, True, True),
                                                        (43, 7, False, True, False), (44, 150, False, False, True)])
def test_fasta_matches_oracle(oracle, seed, width, crlf, blank, lead):
    """FASTA on the device: each record's sequence lines copied together, then the same gates
    and packing; pairs, single files, an empty last record, and a FASTA R1 with a FASTQ R2."""
    say(f"case fasta seed={seed} width={width}")
    rng = np.random.default_rng(seed)
    r1 = random_reads(rng, 700) + [b""]
    r2 = random_reads(rng, 650)
    a, b = fasta(r1, width, crlf, blank, lead), fasta(r2, width, crlf, blank, lead)
    info = check(oracle, a, b)
    assert info["r1_records"] == len(r1) - 1 and info["r2_records"] == len(r2)
    check(oracle, a)
    check(oracle, b, fastq(r1[:-1], crlf))
    check(oracle, b">only\n")  # one record without bases: none
    if width == 0:  # unwrapped records of several kilobases (one workgroup a line)
        longs = [bytes(np.frombuffer(b"ACGTN", dtype=np.uint8)[rng.integers(0, 5, int(L))])
                 for L in (4096, 4097, 9000, 65536 + 33)]
        check(oracle, fasta(longs + random_reads(rng, 50), width, crlf, blank, lead))
    check(oracle, b">a\n>b\n" + b"ACGT" * 8 + b"\n")  # an empty record, then one
    with pytest.raises(KcError) as e:  # no qualities to keep
        from kc.ingest import FastqIngest
        ctx = FastqIngest(keep_quals=True)
        try:
            ctx.pack(a)
        finally:
            ctx.close()
    assert e.value.status == _abi.KC_EUNSUPPORTED


@pytest.mark.parametrize("seed,crlf,n", [(51, False, 600), (52, True, 900), (53, False, 1)])
def test_blank_lines_between_records_match_oracle(oracle, seed, crlf, n):
    """Blank lines between FASTQ records (skipped where FastxReader expects a header): the
    header chain found on the device by pointer doubling; records with empty sequence and
    quality lines, runs of blank lines, leading and trailing blanks."""
    say(f"case blank lines seed={seed} n={n}")
    rng = np.random.default_rng(seed)
    eol = b"\r\n" if crlf else b"\n"

    def blanky(reads):
        out = [eol]
        for i, r in enumerate(reads):
            out.append(eol * int(rng.integers(0, 3)))
            out.append(b"@r%d" % i + eol + r + eol + b"+" + eol + b"I" * len(r) + eol)
        out.append(eol * 2)
        return b"".join(out)

    r1 = random_reads(rng, n) + [b""]
    r2 = random_reads(rng, n)
    a, b = blanky(r1), blanky(r2)
    info = check(oracle, a, b)
    assert info["r1_records"] == len(r1) and info["r2_records"] == len(r2)
    check(oracle, a)
    check(oracle, fastq([b"ACGT" * 10, b"GGGG" * 10], blank_between=True), fastq([b"ACGT" * 10, b"TTTT" * 10]))
    for bad in (b"@r\nACGT\n+\nIIII\n\n  \n@s\nACGT\n+\nIIII\n",  # a whitespace line where a header is due
                b"@r\nACGT\n+\nIIII\n\n@s\nACGT\n+\n"):  # the file ends inside a record
        with pytest.raises(KcError) as e:
            fastq_pack(bad)
        assert e.value.status == _abi.KC_EINVAL, bad
        with pytest.raises(ValueError):
            oracle.pack_fastq(bad)
