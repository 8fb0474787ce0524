"""Host-side PackedReads store, the layout the engine consumes.

Mirrors spades-rs `PackedReads` (packed_reads.rs:42-59): one contiguous byte
buffer of 2-bit bases (A=0, C=1, G=2, T=3, dna.rs:11-22), four bases per byte
with the first base in bits 7..6, every read starting on a fresh byte
(pack_clean, packed_reads.rs:547-578); `offsets[i]` is the byte start of read
i and `lengths[i]` its base count.  Paired input is laid out in three blocks:
R1 mates, R2 mates, then unpaired / N-split sub-reads (finalize_three_blocks,
packed_reads.rs:481-489), with the three-gate N handling of
packed_reads.rs:509-540,633-690.
"""
import gzip

import numpy as np

MIN_K = 21  # packed_reads.rs:38

_ASCII_TO_2BIT = np.full(256, 0xFF, dtype=np.uint8)
for _c, _v in ((b"A", 0), (b"C", 1), (b"G", 2), (b"T", 3), (b"a", 0), (b"c", 1), (b"g", 2), (b"t", 3)):
    _ASCII_TO_2BIT[_c[0]] = _v
_BIT2_TO_ASCII = np.frombuffer(b"ACGT", dtype=np.uint8)


def base_to_2bit(b):
    """dna.rs:30-37: 2-bit code of an ASCII base, or None."""
    v = int(_ASCII_TO_2BIT[b])
    return None if v == 0xFF else v


def bit2_to_base(v):
    """dna.rs:41-43."""
    return int(_BIT2_TO_ASCII[v & 3])


def pack_codes(codes, lengths):
    """Pack reads given as concatenated 2-bit codes (uint8) and per-read lengths.

    Returns (data, offsets) in the pack_clean layout; vectorised, processed in
    chunks of reads so memory stays proportional to the output."""
    codes = np.asarray(codes, dtype=np.uint8)
    lengths = np.asarray(lengths, dtype=np.int64)
    n = lengths.shape[0]
    nbytes = (lengths + 3) // 4
    offsets = np.zeros(n, dtype=np.uint64)
    if n:
        offsets[1:] = np.cumsum(nbytes)[:-1].astype(np.uint64)
    data = np.zeros(int(nbytes.sum()), dtype=np.uint8)
    starts = np.zeros(n, dtype=np.int64)
    if n:
        starts[1:] = np.cumsum(lengths)[:-1]
    chunk = 1 << 18
    for r0 in range(0, n, chunk):
        r1 = min(n, r0 + chunk)
        L = lengths[r0:r1]
        tot = int(L.sum())
        if tot == 0:
            continue
        rid = np.repeat(np.arange(r0, r1), L)
        j = np.arange(tot, dtype=np.int64) - np.repeat(starts[r0:r1] - starts[r0], L)
        src = codes[starts[r0]:starts[r0] + tot]
        byte = offsets[rid].astype(np.int64) + (j >> 2)
        shifted = (src.astype(np.uint8) << (6 - 2 * (j & 3)).astype(np.uint8)).astype(np.uint8)
        # each byte receives at most 4 contributions from distinct bit fields: sum == or
        acc = np.bincount(byte - int(offsets[r0]), weights=shifted, minlength=int(nbytes[r0:r1].sum()))
        data[int(offsets[r0]):int(offsets[r0]) + acc.shape[0]] = acc.astype(np.uint8)
    return data, offsets


def pack_fixed_length(codes2d):
    """Fast path for reads of one length: codes2d (n, L) uint8 -> (data, offsets)."""
    codes2d = np.asarray(codes2d, dtype=np.uint8)
    n, L = codes2d.shape
    nb = (L + 3) // 4
    pad = np.zeros((n, nb * 4), dtype=np.uint8)
    pad[:, :L] = codes2d
    q = pad.reshape(n, nb, 4)
    data = ((q[:, :, 0] << 6) | (q[:, :, 1] << 4) | (q[:, :, 2] << 2) | q[:, :, 3]).astype(np.uint8).reshape(-1)
    offsets = (np.arange(n, dtype=np.uint64) * np.uint64(nb))
    return data, offsets


def _trim_n_ends(read):
    """packed_reads.rs:635-648."""
    codes = _ASCII_TO_2BIT[np.frombuffer(read, dtype=np.uint8)] if len(read) else np.zeros(0, np.uint8)
    ok = np.nonzero(codes != 0xFF)[0]
    if ok.size == 0:
        return b""
    return bytes(read[ok[0]:ok[-1] + 1])


def _has_internal_n(trimmed):
    """packed_reads.rs:652-654."""
    return any(_ASCII_TO_2BIT[b] == 0xFF for b in trimmed)


def _split_on_n(read):
    """packed_reads.rs:659-662: maximal non-empty ACGT runs."""
    out, cur = [], bytearray()
    for b in read:
        if _ASCII_TO_2BIT[b] == 0xFF:
            if cur:
                out.append(bytes(cur))
                cur = bytearray()
        else:
            cur.append(b)
    if cur:
        out.append(bytes(cur))
    return out


class PackedReads:
    """packed_reads.rs:42-59."""

    def __init__(self):
        self._data = bytearray()
        self._offsets = []
        self._lengths = []
        self.pe_count = 0
        self.ambiguous_sidecar = []
        self._frozen = None

    # --- construction --------------------------------------------------------
    @classmethod
    def from_arrays(cls, data, offsets, lengths, pe_count=0):
        s = cls()
        s._frozen = (np.ascontiguousarray(data, dtype=np.uint8), np.ascontiguousarray(offsets, dtype=np.uint64),
                     np.ascontiguousarray(lengths, dtype=np.uint32))
        s.pe_count = int(pe_count)
        return s

    def _thaw(self):
        if self._frozen is not None:
            d, o, l = self._frozen
            self._data = bytearray(d.tobytes())
            self._offsets = [int(x) for x in o]
            self._lengths = [int(x) for x in l]
            self._frozen = None

    def pack_clean(self, read):
        """packed_reads.rs:547-578 (non-ACGT bytes are skipped)."""
        self._thaw()
        self._offsets.append(len(self._data))
        clean_len, byte, shift = 0, 0, 6
        for b in bytes(read):
            v = _ASCII_TO_2BIT[b]
            if v == 0xFF:
                continue
            clean_len += 1
            byte |= int(v) << shift
            if shift == 0:
                self._data.append(byte)
                byte, shift = 0, 6
            else:
                shift -= 2
        if shift != 6:
            self._data.append(byte)
        self._lengths.append(clean_len)

    add_read = pack_clean  # packed_reads.rs:582-584

    def _ingest_trimmed_unpaired(self, trimmed):
        """packed_reads.rs:521-540."""
        if not trimmed:
            return
        any_packed = False
        for seg in _split_on_n(trimmed):
            if len(seg) >= MIN_K:
                self.pack_clean(seg)
                any_packed = True
        if not any_packed:
            self.ambiguous_sidecar.append(bytes(trimmed))

    def ingest_read(self, read):
        """packed_reads.rs:513-517."""
        self._ingest_trimmed_unpaired(_trim_n_ends(bytes(read)))

    def _append(self, other):
        """packed_reads.rs:291-298."""
        self._thaw()
        other._thaw()
        base = len(self._data)
        self._data += other._data
        self._offsets += [o + base for o in other._offsets]
        self._lengths += other._lengths
        self.ambiguous_sidecar += other.ambiguous_sidecar

    @classmethod
    def from_paired_slices(cls, r1_reads, r2_reads):
        """packed_reads.rs:438-461 with partition_pair_streaming (packed_reads.rs:668-690)."""
        b1, b2, b3 = cls(), cls(), cls()
        common = min(len(r1_reads), len(r2_reads))
        for i in range(common):
            t1, t2 = _trim_n_ends(bytes(r1_reads[i])), _trim_n_ends(bytes(r2_reads[i]))
            if (not _has_internal_n(t1) and not _has_internal_n(t2) and len(t1) >= MIN_K and len(t2) >= MIN_K):
                b1.pack_clean(t1)
                b2.pack_clean(t2)
            else:
                b3._ingest_trimmed_unpaired(t1)
                b3._ingest_trimmed_unpaired(t2)
        for r in r1_reads[common:]:
            b3.ingest_read(r)
        for r in r2_reads[common:]:
            b3.ingest_read(r)
        pe = len(b1._lengths)
        b1._append(b2)
        b1._append(b3)
        b1.pe_count = pe
        return b1

    @classmethod
    def from_paired_fastq(cls, r1_path, r2_path):
        """from_paired_files (packed_reads.rs:324-381) over FASTQ (optionally gzipped)."""
        r1 = read_fastq_sequences(r1_path)
        r2 = read_fastq_sequences(r2_path)
        return cls.from_paired_slices(r1, r2)

    # --- access --------------------------------------------------------------
    def arrays(self):
        """(data uint8, offsets uint64, lengths uint32) views of the store."""
        if self._frozen is None:
            self._frozen = (np.frombuffer(bytes(self._data), dtype=np.uint8).copy(),
                            np.array(self._offsets, dtype=np.uint64),
                            np.array(self._lengths, dtype=np.uint32))
            self._data, self._offsets, self._lengths = bytearray(), [], []
        return self._frozen

    @property
    def data(self):
        return self.arrays()[0]

    @property
    def offsets(self):
        return self.arrays()[1]

    @property
    def lengths(self):
        return self.arrays()[2]

    def __len__(self):
        return int(self.arrays()[2].shape[0])

    def is_empty(self):
        return len(self) == 0

    def get_read(self, idx):
        """packed_reads.rs:592-613: ASCII of read idx."""
        data, offs, lens = self.arrays()
        L = int(lens[idx])
        if L == 0:
            return b""
        o = int(offs[idx])
        b = data[o:o + (L + 3) // 4]
        codes = np.stack([(b >> 6) & 3, (b >> 4) & 3, (b >> 2) & 3, b & 3], axis=1).reshape(-1)[:L]
        return _BIT2_TO_ASCII[codes].tobytes()

    def get_pair(self, p):
        """packed_reads.rs:278-288."""
        return self.get_read(p), self.get_read(p + self.pe_count)


def read_fastq_sequences(path):
    """Sequence lines of a 4-line FASTQ file, validated as FastxReader::next_seq does
    (packed_reads.rs:141-216); quality lines are checked for length and dropped."""
    opener = gzip.open if str(path).endswith(".gz") else open
    seqs = []
    with opener(path, "rb") as fh:
        lines = [ln.rstrip(b"\r\n") for ln in fh]
    i = 0
    while i < len(lines):
        if not lines[i]:
            i += 1
            continue
        if not lines[i].startswith(b"@"):
            raise ValueError(f"Malformed FASTQ record: expected '@' header, found {lines[i][:32]!r}")
        if i + 3 >= len(lines):
            raise ValueError("Malformed FASTQ record: unexpected EOF inside a record")
        seq, plus, qual = lines[i + 1], lines[i + 2], lines[i + 3]
        if not plus.startswith(b"+"):
            raise ValueError(f"Malformed FASTQ record: expected '+' separator line, found {plus[:32]!r}")
        if len(qual) != len(seq):
            raise ValueError(f"Malformed FASTQ record: sequence length ({len(seq)}) != quality length ({len(qual)})")
        seqs.append(seq)
        i += 4
    return seqs


def read_sequences(path):
    """parse_reads_from_file (fastq.rs:156-358, no quality filter): the format is taken from
    the first non-whitespace byte ('@' FASTQ, '>' FASTA, anything else is an error).  FASTA
    records are the concatenation of their lines (CR/LF stripped); empty records and empty
    FASTQ sequences are dropped."""
    opener = gzip.open if str(path).lower().endswith(".gz") else open
    with opener(path, "rb") as fh:
        raw = fh.read()
    body = raw.lstrip()
    if not body:
        return []
    if body[:1] == b"@":
        return [s for s in read_fastq_sequences(path) if s]
    if body[:1] != b">":
        raise ValueError(f"Malformed FASTQ/FASTA file {path!r}: expected '@' or '>' header, found {body[:1]!r}")
    seqs, cur = [], []
    lines = body.split(b"\n")
    for line in lines[1:]:  # the rest of the first header line is discarded
        line = line.rstrip(b"\r")
        if line.startswith(b">"):
            if cur:
                seqs.append(b"".join(cur))
            cur = []
        else:
            cur.append(line)
    if cur and b"".join(cur):
        seqs.append(b"".join(cur))
    return [q for q in seqs if q]
