"""ctypes wrapper around oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: the checker for the MI355X engine.  Imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product package.  Every function forwards to oracle/oracle.cpp, which restates
the reference (spades-rs) Rust code it cites.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build():
    """Compile oracle/oracle.cpp with its Makefile (g++)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "or_ascii_to_2bit": (ctypes.c_uint8, [ctypes.c_uint8]),
            "or_bit2_to_base": (ctypes.c_uint8, [ctypes.c_uint8]),
            "or_revcomp_u64": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint32]),
            "or_canonical_u64": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]),
            "or_hamming_2bit": (ctypes.c_uint32, [ctypes.c_uint64, ctypes.c_uint64]),
            "or_kmer_from_bytes": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32, _u64p]),
            "or_kmer_to_string": (None, [_u64p, ctypes.c_uint32, ctypes.c_char_p]),
            "or_kmer_revcomp": (None, [_u64p, ctypes.c_uint32, _u64p]),
            "or_kmer_canonical": (ctypes.c_int, [_u64p, ctypes.c_uint32, _u64p]),
            "or_kmer_extend_right": (None, [_u64p, ctypes.c_uint8, ctypes.c_uint32, _u64p]),
            "or_kmer_prepend_left": (None, [_u64p, ctypes.c_uint8, ctypes.c_uint32, _u64p]),
            "or_hash_kmer256": (ctypes.c_uint64, [_u64p, ctypes.c_uint64]),
            "or_hash_many": (None, [_u64p, ctypes.c_uint64, ctypes.c_uint64, _u64p]),
            "or_filter_new": (ctypes.c_void_p, [ctypes.c_uint64]),
            "or_filter_free": (None, [ctypes.c_void_p]),
            "or_filter_num_words": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_filter_insert": (ctypes.c_int, [ctypes.c_void_p, _u64p]),
            "or_filter_is_solid": (ctypes.c_int, [ctypes.c_void_p, _u64p]),
            "or_filter_insert_many": (None, [ctypes.c_void_p, _u64p, ctypes.c_uint64, _u8p]),
            "or_filter_is_solid_many": (None, [ctypes.c_void_p, _u64p, ctypes.c_uint64, _u8p]),
            "or_filter_export": (None, [ctypes.c_void_p, _u64p, _u64p]),
            "or_correct_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _u8p, ctypes.c_uint64]),
            "or_cutoff": (ctypes.c_uint32, [_u32p, ctypes.c_uint64, ctypes.c_double, _u32p]),
            "or_pack": (ctypes.c_uint64, [_u8p, _u64p, _u64p, ctypes.c_uint64, _u8p, _u64p, _u32p]),
            "or_get_read": (None, [_u8p, ctypes.c_uint64, ctypes.c_uint32, _u8p]),
            "or_spectrum_run": (ctypes.c_void_p, [_u8p, _u64p, _u32p, ctypes.c_uint64, ctypes.c_uint32,
                                                  ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "or_spectrum_run_priors": (ctypes.c_void_p, [_u8p, _u64p, _u32p, ctypes.c_uint64, ctypes.c_uint32,
                                                         ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                         _u8p, _u64p, _u64p, ctypes.c_uint64]),
            "or_prior_boost": (ctypes.c_uint32, [ctypes.c_double]),
            "or_pack_fastq": (ctypes.c_void_p, [_u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64]),
            "or_anchor_run": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_uint32, _u8p, _u64p, _u64p, ctypes.c_uint64]),
            "or_anchor_size": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_anchor_export": (None, [ctypes.c_void_p, _u64p, _u32p, _u32p, _u8p]),
            "or_anchor_free": (None, [ctypes.c_void_p]),
            "or_graph_build": (ctypes.c_void_p, [ctypes.c_uint32, _u64p, ctypes.c_uint64, _u64p, _u32p,
                                                 ctypes.c_uint64]),
            "or_graph_n": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_graph_bases": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_graph_export": (None, [ctypes.c_void_p, _u8p, _u64p, _u64p, ctypes.POINTER(ctypes.c_double)]),
            "or_graph_free": (None, [ctypes.c_void_p]),
            "or_pack_fastq_q": (ctypes.c_void_p, [_u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64]),
            "or_pack_n_quals": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_pack_export_quals": (None, [ctypes.c_void_p, _u8p, _u64p]),
            "or_pack_free": (None, [ctypes.c_void_p]),
            "or_pack_sizes": (ctypes.c_int, [ctypes.c_void_p, _u64p]),
            "or_pack_export": (None, [ctypes.c_void_p, _u8p, _u64p, _u32p]),
            "or_spectrum_free": (None, [ctypes.c_void_p]),
            "or_spectrum_num_words": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_spectrum_num_counts": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_spectrum_n_occ": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_spectrum_cutoff": (ctypes.c_uint32, [ctypes.c_void_p]),
            "or_spectrum_top_cov": (ctypes.c_uint32, [ctypes.c_void_p]),
            "or_spectrum_n_solid": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_spectrum_corrected_reads": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_spectrum_stage2_secs": (ctypes.c_double, [ctypes.c_void_p]),
            "or_spectrum_stage3_secs": (ctypes.c_double, [ctypes.c_void_p]),
            "or_spectrum_export_bloom": (None, [ctypes.c_void_p, _u64p, _u64p]),
            "or_spectrum_export_counts": (None, [ctypes.c_void_p, _u64p, _u32p, _u8p]),
            "or_spectrum_corrected_bytes": (ctypes.c_uint64, [ctypes.c_void_p]),
            "or_spectrum_export_corrected": (None, [ctypes.c_void_p, _u8p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def words(w):
    a = (ctypes.c_uint64 * 4)(*[int(x) & 0xFFFFFFFFFFFFFFFF for x in w])
    return a


# --- dna.rs -----------------------------------------------------------------

def kmer_from_bytes(s, k):
    """Kmer256::from_bytes (dna.rs:299-321) -> 4 limbs or None."""
    out = (ctypes.c_uint64 * 4)()
    ok = lib().or_kmer_from_bytes(bytes(s), len(s), k, out)
    return tuple(out) if ok else None


def kmer_to_string(w, k):
    buf = ctypes.create_string_buffer(k + 1)
    lib().or_kmer_to_string(words(w), k, buf)
    return buf.raw[:k]


def kmer_revcomp(w, k):
    out = (ctypes.c_uint64 * 4)()
    lib().or_kmer_revcomp(words(w), k, out)
    return tuple(out)


def kmer_canonical(w, k):
    out = (ctypes.c_uint64 * 4)()
    rc = lib().or_kmer_canonical(words(w), k, out)
    return tuple(out), bool(rc)


def kmer_extend_right(w, b, k):
    out = (ctypes.c_uint64 * 4)()
    lib().or_kmer_extend_right(words(w), b, k, out)
    return tuple(out)


def kmer_prepend_left(w, a, k):
    out = (ctypes.c_uint64 * 4)()
    lib().or_kmer_prepend_left(words(w), a, k, out)
    return tuple(out)


def revcomp_u64(x, k):
    return lib().or_revcomp_u64(x, k)


def canonical_u64(x, k):
    rc = ctypes.c_int(0)
    v = lib().or_canonical_u64(x, k, ctypes.byref(rc))
    return v, bool(rc.value)


def hamming_2bit(a, b):
    return lib().or_hamming_2bit(a, b)


# --- bloom.rs ---------------------------------------------------------------

SEEDS = (0x517CC1B727220A95, 0x9E3779B97F4A7C15, 0x6C62272E07BB0142)


def hash_kmer256(w, seed):
    return lib().or_hash_kmer256(words(w), seed)


def hash_many(keys4, seed):
    keys4 = np.ascontiguousarray(keys4, dtype=np.uint64)
    out = np.empty(keys4.shape[0], dtype=np.uint64)
    lib().or_hash_many(_p(keys4, _u64p), keys4.shape[0], seed, _p(out, _u64p))
    return out


class TwoTierFilter:
    """bloom.rs:13-131, sequential."""

    def __init__(self, capacity_bits):
        self._h = lib().or_filter_new(capacity_bits)
        self.num_words = lib().or_filter_num_words(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_filter_free(self._h)
            self._h = None

    def insert_kmer256(self, w):
        return bool(lib().or_filter_insert(self._h, words(w)))

    def is_solid_kmer256(self, w):
        return bool(lib().or_filter_is_solid(self._h, words(w)))

    def insert(self, x):
        return self.insert_kmer256((x, 0, 0, 0))

    def is_solid(self, x):
        return self.is_solid_kmer256((x, 0, 0, 0))

    def insert_many(self, keys4):
        keys4 = np.ascontiguousarray(keys4, dtype=np.uint64)
        ret = np.zeros(keys4.shape[0], dtype=np.uint8)
        lib().or_filter_insert_many(self._h, _p(keys4, _u64p), keys4.shape[0], _p(ret, _u8p))
        return ret

    def is_solid_many(self, keys4):
        keys4 = np.ascontiguousarray(keys4, dtype=np.uint64)
        out = np.zeros(keys4.shape[0], dtype=np.uint8)
        lib().or_filter_is_solid_many(self._h, _p(keys4, _u64p), keys4.shape[0], _p(out, _u8p))
        return out

    def export(self):
        seen = np.zeros(self.num_words, dtype=np.uint64)
        solid = np.zeros(self.num_words, dtype=np.uint64)
        lib().or_filter_export(self._h, _p(seen, _u64p), _p(solid, _u64p))
        return seen, solid


# --- hammer.rs --------------------------------------------------------------

def correct_read(filt, k, seq):
    """ErrorCorrector::correct_read_in_place (hammer.rs:34-120). Returns (modified, new_seq)."""
    buf = np.frombuffer(bytearray(seq), dtype=np.uint8).copy()
    m = lib().or_correct_read(filt._h, k, _p(buf, _u8p), len(buf))
    return bool(m), bytes(buf)


# --- assemble.rs ------------------------------------------------------------

def cutoff(counts, min_cov=5.0):
    """determine_solid_kmer_cutoff (assemble.rs:692-735). Returns (cutoff, top_cov)."""
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    top = ctypes.c_uint32(0)
    v = lib().or_cutoff(_p(c, _u32p), c.shape[0], min_cov, ctypes.byref(top))
    return v, top.value


# --- packed_reads.rs --------------------------------------------------------

def pack(reads):
    """pack_clean for each read (packed_reads.rs:547-578) -> (data, offsets, lengths)."""
    n = len(reads)
    lens = np.array([len(r) for r in reads], dtype=np.uint64)
    starts = np.zeros(n, dtype=np.uint64)
    if n:
        starts[1:] = np.cumsum(lens)[:-1]
    ascii = np.frombuffer(b"".join(bytes(r) for r in reads) + b"\0", dtype=np.uint8).copy()
    data = np.zeros(int(((lens + 3) // 4).sum()) + 1, dtype=np.uint8)
    offs = np.zeros(n, dtype=np.uint64)
    lns = np.zeros(n, dtype=np.uint32)
    nb = lib().or_pack(_p(ascii, _u8p), _p(starts, _u64p), _p(lens, _u64p), n,
                       _p(data, _u8p), _p(offs, _u64p), _p(lns, _u32p))
    return data[:nb].copy(), offs, lns


def get_read(data, off, ln):
    out = np.zeros(max(int(ln), 1), dtype=np.uint8)
    d = np.ascontiguousarray(data, dtype=np.uint8)
    if d.size == 0:
        d = np.zeros(1, dtype=np.uint8)
    lib().or_get_read(_p(d, _u8p), int(off), int(ln), _p(out, _u8p))
    return bytes(out[: int(ln)])


def concat_ascii(seqs):
    """Concatenate ASCII sequences: (buffer, starts, lengths) as numpy arrays."""
    bs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    lens = np.array([len(b) for b in bs], dtype=np.uint64)
    starts = np.zeros(len(bs), dtype=np.uint64)
    if len(bs) > 1:
        starts[1:] = np.cumsum(lens)[:-1]
    buf = np.frombuffer(b"".join(bs) + b"\0", dtype=np.uint8).copy()
    return buf, starts, lens


def prior_boost(min_cov):
    """assemble.rs:270: (min_coverage as u32 * 4).max(50)."""
    return lib().or_prior_boost(float(min_cov))


class Spectrum:
    """Stage 2 + Stage 3 of run_assembly_with_packed_reads (assemble.rs:129-298).

    threads == 1: the sequential oracle; threads > 1: the reference's parallel structure
    (timed CPU baseline; its solid tier may legitimately differ, bloom.rs:95-103)."""

    def __init__(self, data, offsets, lengths, k, bloom_bits, error_correct=False, min_cov=5.0, threads=1,
                 priors=None):
        """priors: prior / trusted contig sequences (ASCII bytes or str), assemble.rs:244-284."""
        data = np.ascontiguousarray(data, dtype=np.uint8)
        if data.size == 0:
            data = np.zeros(1, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
        L = lib()
        if priors:
            buf, starts, lens = concat_ascii(priors)
            self._h = L.or_spectrum_run_priors(_p(data, _u8p), _p(offsets, _u64p), _p(lengths, _u32p),
                                               lengths.shape[0], k, bloom_bits, int(bool(error_correct)),
                                               float(min_cov), int(threads), _p(buf, _u8p), _p(starts, _u64p),
                                               _p(lens, _u64p), len(priors))
        else:
            self._h = L.or_spectrum_run(_p(data, _u8p), _p(offsets, _u64p), _p(lengths, _u32p), lengths.shape[0],
                                        k, bloom_bits, int(bool(error_correct)), float(min_cov), int(threads))
        self.k = k
        self.num_words = L.or_spectrum_num_words(self._h)
        self.n_counts = L.or_spectrum_num_counts(self._h)
        self.n_occ = L.or_spectrum_n_occ(self._h)
        self.cutoff = L.or_spectrum_cutoff(self._h)
        self.top_cov = L.or_spectrum_top_cov(self._h)
        self.n_solid = L.or_spectrum_n_solid(self._h)
        self.corrected_reads = L.or_spectrum_corrected_reads(self._h)
        self.stage2_secs = L.or_spectrum_stage2_secs(self._h)
        self.stage3_secs = L.or_spectrum_stage3_secs(self._h)

    def bloom(self):
        seen = np.zeros(self.num_words, dtype=np.uint64)
        solid = np.zeros(self.num_words, dtype=np.uint64)
        lib().or_spectrum_export_bloom(self._h, _p(seen, _u64p), _p(solid, _u64p))
        return seen, solid

    def counts(self):
        n = self.n_counts
        keys = np.zeros((max(n, 1), 4), dtype=np.uint64)
        cnt = np.zeros(max(n, 1), dtype=np.uint32)
        mask = np.zeros(max(n, 1), dtype=np.uint8)
        lib().or_spectrum_export_counts(self._h, _p(keys, _u64p), _p(cnt, _u32p), _p(mask, _u8p))
        return keys[:n], cnt[:n], mask[:n]

    def corrected_data(self):
        nb = lib().or_spectrum_corrected_bytes(self._h)
        out = np.zeros(max(nb, 1), dtype=np.uint8)
        lib().or_spectrum_export_corrected(self._h, _p(out, _u8p))
        return out[:nb]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_spectrum_free(self._h)
            self._h = None


def pack_fastq(r1, r2=None, quals=False):
    """packed_reads.rs from_paired_fastq (or append_from_file when r2 is None) over file
    contents: (data, offsets, lengths, pe_count, n_ambiguous, r1_records, r2_records), and
    with quals=True also (quals, qoff): the Phred value of every packed base (quality
    character - 33), read i's at quals[qoff[i]:].  Raises ValueError for a malformed file."""
    L = lib()
    a = np.frombuffer(bytes(r1) + b"\0", dtype=np.uint8)
    b = None if r2 is None else np.frombuffer(bytes(r2) + b"\0", dtype=np.uint8)
    fn = L.or_pack_fastq_q if quals else L.or_pack_fastq
    h = fn(_p(a, _u8p), len(r1), _p(b, _u8p) if b is not None else None, 0 if r2 is None else len(r2))
    try:
        sizes = np.zeros(6, dtype=np.uint64)
        if L.or_pack_sizes(h, _p(sizes, _u64p)) != 0:
            raise ValueError("malformed FASTQ/FASTA")
        n, dl = int(sizes[0]), int(sizes[1])
        data = np.zeros(max(dl, 1), dtype=np.uint8)
        offs = np.zeros(max(n, 1), dtype=np.uint64)
        lens = np.zeros(max(n, 1), dtype=np.uint32)
        L.or_pack_export(h, _p(data, _u8p), _p(offs, _u64p), _p(lens, _u32p))
        out = (data[:dl], offs[:n], lens[:n], int(sizes[2]), int(sizes[3]), int(sizes[4]), int(sizes[5]))
        if not quals:
            return out
        nq = int(L.or_pack_n_quals(h))
        q = np.zeros(max(nq, 1), dtype=np.uint8)
        qo = np.zeros(max(n, 1), dtype=np.uint64)
        L.or_pack_export_quals(h, _p(q, _u8p), _p(qo, _u64p))
        return out + (q[:nq], qo[:n])
    finally:
        L.or_pack_free(h)


ANCHOR_KINDS = {"polisher": 0, "paired_info": 1, "splitter": 2}


def anchor_index(kind, k, seqs):
    """The unique-anchor k-mer index of polisher.rs:33-60, paired_info.rs:34-70 or
    splitter.rs:57-70 over ASCII sequences: (keys (n, 4) Kmer256 limbs sorted, sequence
    index, position, is_rc) of every k-mer the reference's map keeps."""
    L = lib()
    data, starts, lens = concat_ascii(seqs)
    h = L.or_anchor_run(ANCHOR_KINDS[kind], k, _p(data, _u8p), _p(starts, _u64p), _p(lens, _u64p), len(seqs))
    try:
        n = int(L.or_anchor_size(h))
        keys = np.zeros((max(n, 1), 4), dtype=np.uint64)
        seq = np.zeros(max(n, 1), dtype=np.uint32)
        pos = np.zeros(max(n, 1), dtype=np.uint32)
        rc = np.zeros(max(n, 1), dtype=np.uint8)
        L.or_anchor_export(h, _p(keys, _u64p), _p(seq, _u32p), _p(pos, _u32p), _p(rc, _u8p))
        return keys[:n], seq[:n], pos[:n], rc[:n]
    finally:
        L.or_anchor_free(h)


def compacted_graph(k, solid_keys4, cov_keys4, cov_counts):
    """CompactedGraph::build (graph.rs:65-234) over the solid k-mers (Kmer256 limbs) with
    coverage from the count table; iterates the solid k-mers in ascending order.  Returns
    a list of (sequence bytes, kmers_count, mean_coverage) in build order."""
    L = lib()
    s = np.ascontiguousarray(solid_keys4, dtype=np.uint64).reshape(-1, 4)
    c = np.ascontiguousarray(cov_keys4, dtype=np.uint64).reshape(-1, 4)
    n = np.ascontiguousarray(cov_counts, dtype=np.uint32)
    s1 = s if s.shape[0] else np.zeros((1, 4), np.uint64)
    c1 = c if c.shape[0] else np.zeros((1, 4), np.uint64)
    n1 = n if n.shape[0] else np.zeros(1, np.uint32)
    h = L.or_graph_build(k, _p(s1, _u64p), s.shape[0], _p(c1, _u64p), _p(n1, _u32p), c.shape[0])
    try:
        m, nb = int(L.or_graph_n(h)), int(L.or_graph_bases(h))
        seq = np.zeros(max(nb, 1), np.uint8)
        off = np.zeros(m + 1, np.uint64)
        km = np.zeros(max(m, 1), np.uint64)
        cov = np.zeros(max(m, 1), np.float64)
        L.or_graph_export(h, _p(seq, _u8p), _p(off, _u64p), _p(km, _u64p),
                          cov.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return [(seq[int(off[i]):int(off[i + 1])].tobytes(), int(km[i]), float(cov[i])) for i in range(m)]
    finally:
        L.or_graph_free(h)

