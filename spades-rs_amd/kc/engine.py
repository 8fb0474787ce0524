"""Host interface of the gfx950 engine, mirroring the reference's Rust API.

  TwoTierFilter        bloom.rs:13-131      (device-resident two-tier filter)
  ErrorCorrector       hammer.rs:20-144     (Hammer against a TwoTierFilter)
  AssemblerConfig      assemble.rs:35-103   (the fields the k-mer path reads)
  count_kmers          assemble.rs:122-298  (Stage 2 + Stage 3 of
                                             run_assembly_with_packed_reads)

Everything runs through libkc.so (include/kc.h); there is no host fallback.
A k-mer crosses the boundary as the four little-endian limbs of the
reference's Kmer256(lo: u128, hi: u128) (dna.rs:137-138).
"""
import ctypes
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _abi
from ._abi import KcError, lib, ptr, u8p, u32p, u64p


def _check(status, obj, errfn):
    if status != _abi.KC_OK:
        msg = errfn(obj).decode() if obj else ""
        raise KcError(status, msg)


def _keys4(keys):
    """Accept u64 k-mers (k <= 32, as the reference's insert(u64)) or (n, 4) limbs."""
    a = np.asarray(keys, dtype=np.uint64)
    if a.ndim == 1:
        out = np.zeros((a.shape[0], 4), dtype=np.uint64)
        out[:, 0] = a
        return out
    return np.ascontiguousarray(a.reshape(-1, 4))


class TwoTierFilter:
    """bloom.rs:13-131 on one GPU.

    insert_kmer256 keeps the reference's sequential semantics: a batch is applied as if
    its keys were inserted one after the other, after all earlier batches."""

    def __init__(self, capacity_bits, device=0):
        h = ctypes.c_void_p()
        st = lib().kc_filter_create(ctypes.byref(h), device, int(capacity_bits))
        if st != _abi.KC_OK:
            raise KcError(st, "kc_filter_create failed")
        self._h = h
        self.num_words = lib().kc_filter_num_words(h)
        self.size_bits = self.num_words * 64

    def close(self):
        if getattr(self, "_h", None):
            lib().kc_filter_destroy(self._h)
            self._h = None

    __del__ = close

    def memory_usage_bytes(self):
        """bloom.rs:129-131."""
        return lib().kc_filter_memory_usage_bytes(self._h)

    def insert_kmer256_many(self, keys4):
        k4 = _keys4(keys4)
        prom = np.zeros(k4.shape[0], dtype=np.uint8)
        _check(lib().kc_filter_insert_kmer256(self._h, ptr(k4, u64p), k4.shape[0], ptr(prom, u8p)),
               self._h, lib().kc_filter_last_error)
        return prom.astype(bool)

    def is_solid_kmer256_many(self, keys4):
        k4 = _keys4(keys4)
        out = np.zeros(k4.shape[0], dtype=np.uint8)
        _check(lib().kc_filter_is_solid_kmer256(self._h, ptr(k4, u64p), k4.shape[0], ptr(out, u8p)),
               self._h, lib().kc_filter_last_error)
        return out.astype(bool)

    def insert_kmer256(self, limbs):
        """bloom.rs:83-107."""
        return bool(self.insert_kmer256_many(np.array([limbs], dtype=np.uint64))[0])

    def is_solid_kmer256(self, limbs):
        """bloom.rs:111-126."""
        return bool(self.is_solid_kmer256_many(np.array([limbs], dtype=np.uint64))[0])

    def insert(self, kmer):
        """bloom.rs:48-50: Kmer256(kmer as u128, 0)."""
        return self.insert_kmer256((int(kmer), 0, 0, 0))

    def is_solid(self, kmer):
        """bloom.rs:54-56."""
        return self.is_solid_kmer256((int(kmer), 0, 0, 0))

    def export(self):
        seen = np.zeros(self.num_words, dtype=np.uint64)
        solid = np.zeros(self.num_words, dtype=np.uint64)
        _check(lib().kc_filter_export(self._h, ptr(seen, u64p), ptr(solid, u64p), self.num_words),
               self._h, lib().kc_filter_last_error)
        return seen, solid


def hamming_distance_2bit(a, b, device=0):
    """hammer.rs:13-18 on the GPU for arrays of 2-bit packed u64 words: the number of bases
    in which a[i] and b[i] differ."""
    x = np.ascontiguousarray(np.atleast_1d(a), dtype=np.uint64)
    y = np.ascontiguousarray(np.atleast_1d(b), dtype=np.uint64)
    if x.shape != y.shape:
        raise ValueError("a and b must have the same shape")
    out = np.zeros(max(x.size, 1), dtype=np.uint32)
    st = lib().kc_hamming_distance_2bit(device, ptr(x, u64p), ptr(y, u64p), x.size, ptr(out, u32p))
    if st != _abi.KC_OK:
        raise KcError(st, "kc_hamming_distance_2bit failed")
    return out[:x.size]


class ErrorCorrector:
    """hammer.rs:20-144: single-base correction of reads against a TwoTierFilter."""

    def __init__(self, k):
        self.k = int(k)

    def correct_packed(self, packed, filt):
        """Correct every read of a PackedReads store; returns (corrected data, n_corrected)."""
        data, offs, lens = packed.arrays()
        out = np.zeros(max(data.shape[0], 1), dtype=np.uint8)
        n_corr = ctypes.c_uint64(0)
        d = data if data.shape[0] else np.zeros(1, np.uint8)
        _check(lib().kc_filter_correct_reads(filt._h, self.k, ptr(d, u8p), data.shape[0], ptr(offs, u64p),
                                             ptr(lens, u32p), lens.shape[0], ptr(out, u8p), ctypes.byref(n_corr)),
               filt._h, lib().kc_filter_last_error)
        return out[: data.shape[0]], int(n_corr.value)

    def correct_reads(self, reads, filt):
        """hammer.rs:123-143: (corrected ASCII reads, number of reads corrected)."""
        from .packed_reads import PackedReads
        store = PackedReads()
        for r in reads:
            store.pack_clean(r)
        data, n = self.correct_packed(store, filt)
        fixed = PackedReads.from_arrays(data, store.offsets, store.lengths)
        return [fixed.get_read(i) for i in range(len(fixed))], n

    def correct_read_in_place(self, seq, filt):
        """hammer.rs:34-120 for one ASCII read: returns (modified, corrected read)."""
        out, n = self.correct_reads([seq], filt)
        return n > 0, out[0]


@dataclass
class AssemblerConfig:
    """The AssemblerConfig fields read by Stage 2-3 (assemble.rs:35-103 defaults).

    prior_contigs: sequences (str or bytes) whose k-mers get the prior boost
    (assemble.rs:244-284); trusted_contigs: FASTA/FASTQ paths whose sequences are added
    to them (parse_reads_from_file, fastq.rs:156)."""
    k: int = 31
    min_coverage: float = 5.0
    bloom_bits: int = 512 * 1024 * 1024
    error_correct: bool = False
    prior_contigs: Optional[List] = None
    trusted_contigs: Optional[List[str]] = None

    def all_priors(self):
        """assemble.rs:244-264: prior contigs, then every trusted file's sequences (a file
        that fails to parse is skipped with a warning, as the reference does)."""
        from .packed_reads import read_sequences
        out = [p.encode() if isinstance(p, str) else bytes(p) for p in (self.prior_contigs or [])]
        for path in self.trusted_contigs or []:
            try:
                out.extend(read_sequences(path))
            except (OSError, ValueError) as e:
                import sys
                print(f"  [Warning] Failed to parse trusted contigs from {path!r}: {e}", file=sys.stderr)
        return out


@dataclass
class KmerSpectrum:
    """Products of Stage 2-3: both Bloom tiers, the sorted count table (keys as Kmer256
    limbs), the cutoff and the solid set (assemble.rs:286-292)."""
    k: int
    seen: np.ndarray
    solid_bits: np.ndarray
    keys: np.ndarray
    counts: np.ndarray
    solid_mask: np.ndarray
    cutoff: int
    info: dict

    @property
    def solid_kmers(self):
        return self.keys[self.solid_mask.astype(bool)]


class SpectrumEngine:
    """One engine context: reads stay resident in HBM across runs."""

    def __init__(self, config, device=0, table_load=0.0, timing=False, events64=False):
        """events64: 64-bit event indices on any input (the path libraries of 2^30 - 1
        windows or more take; the results are the same)."""
        cfg = config
        if not 1 <= cfg.k <= _abi.KC_MAX_K:
            raise KcError(_abi.KC_EUNSUPPORTED if cfg.k == 128 else _abi.KC_EINVAL,
                          f"k={cfg.k} outside 1..{_abi.KC_MAX_K}")
        o = _abi.kc_opts(device=device, k=cfg.k, bloom_bits=int(cfg.bloom_bits), min_coverage=float(cfg.min_coverage),
                         error_correct=int(bool(cfg.error_correct)), table_load=float(table_load),
                         flags=(_abi.KC_FLAG_TIMING if timing else 0) | (_abi.KC_FLAG_EVENTS64 if events64 else 0))
        h = ctypes.c_void_p()
        st = lib().kc_create(ctypes.byref(h), ctypes.byref(o))
        if st != _abi.KC_OK:
            raise KcError(st, "kc_create failed")
        self._h = h
        self.config = cfg
        self._data_len = []

    def close(self):
        if getattr(self, "_h", None):
            lib().kc_destroy(self._h)
            self._h = None

    __del__ = close

    def _chk(self, st):
        _check(st, self._h, lib().kc_last_error)

    def upload(self, packed):
        """Replace every library with `packed` (library 0)."""
        data, offs, lens = packed.arrays()
        d = data if data.shape[0] else np.zeros(1, np.uint8)
        self._chk(lib().kc_reads_upload(self._h, ptr(d, u8p), data.shape[0], ptr(offs, u64p), ptr(lens, u32p),
                                        lens.shape[0]))
        self._data_len = [data.shape[0]]

    def add_library(self, packed, bloom_bits=None, min_coverage=None):
        """Append an independent library; returns its index."""
        data, offs, lens = packed.arrays()
        d = data if data.shape[0] else np.zeros(1, np.uint8)
        idx = ctypes.c_uint32(0)
        bb = self.config.bloom_bits if bloom_bits is None else bloom_bits
        mc = self.config.min_coverage if min_coverage is None else min_coverage
        self._chk(lib().kc_library_add(self._h, ptr(d, u8p), data.shape[0], ptr(offs, u64p), ptr(lens, u32p),
                                       lens.shape[0], int(bb), float(mc), ctypes.byref(idx)))
        if idx.value == 0:
            self._data_len = []
        self._data_len.append(data.shape[0])
        return idx.value

    def set_priors(self, seqs, library=0):
        """Prior / trusted contig sequences of a library (assemble.rs:244-284); applied by the
        next run().  An empty list clears them."""
        bs = [q.encode() if isinstance(q, str) else bytes(q) for q in seqs]
        lens = np.array([len(b) for b in bs], dtype=np.uint64)
        starts = np.zeros(len(bs), dtype=np.uint64)
        if len(bs) > 1:
            starts[1:] = np.cumsum(lens)[:-1]
        buf = np.frombuffer(b"".join(bs) + b"\0", dtype=np.uint8).copy()
        self._chk(lib().kc_library_priors_set(self._h, library, ptr(buf, u8p), int(lens.sum()), ptr(starts, u64p),
                                              ptr(lens, u64p), len(bs)))

    def n_libraries(self):
        n = ctypes.c_uint32(0)
        self._chk(lib().kc_library_count(self._h, ctypes.byref(n)))
        return n.value

    def bloom_build(self):
        self._chk(lib().kc_bloom_build(self._h))

    def count(self):
        info = _abi.kc_info()
        self._chk(lib().kc_count(self._h, ctypes.byref(info)))
        return info.as_dict()

    def run(self):
        info = _abi.kc_info()
        self._chk(lib().kc_run(self._h, ctypes.byref(info)))
        return info.as_dict()

    def set_k(self, k):
        """Next runs count k-mers of size k over the same resident reads (multik.rs:109-178)."""
        self._chk(lib().kc_set_k(self._h, int(k)))
        self.config = AssemblerConfig(k=int(k), min_coverage=self.config.min_coverage,
                                      bloom_bits=self.config.bloom_bits, error_correct=self.config.error_correct,
                                      prior_contigs=self.config.prior_contigs,
                                      trusted_contigs=self.config.trusted_contigs)

    def graph_run(self):
        """Stage 2 + 3 as one HIP graph launch (captured on the first call; re-planned
        eagerly whenever a data-dependent size no longer matches).  Same results as run()."""
        info = _abi.kc_info()
        self._chk(lib().kc_graph_run(self._h, ctypes.byref(info)))
        return info.as_dict()

    def graph_release(self):
        self._chk(lib().kc_graph_release(self._h))

    def tune(self):
        t = _abi.kc_tune()
        self._chk(lib().kc_tune_get(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tune(self, windows_per_thread=1, blocks_per_cu=0, table_load=0.0, insert_path=0, setter_path=0,
                 kmers_per_bin=0, minimizer_len=0):
        """insert_path: 0 (default) or 3 minimizer-binned count where it applies (k <= 63,
        reads under 2^24 bases, no prior contigs or Hammer recount inside a graph capture),
        else per-window table atomics; 1 per-window table atomics.  setter_path: 0 (default)
        Bloom first setters by sorting (entry, setter) items by partition and taking minima
        in LDS where it applies (dense mode, 32-bit events), else device atomics; 1 device
        atomics.  kmers_per_bin: distinct k-mers a bin of the binned count is sized for (0:
        352).  minimizer_len: m of the binned count's minimizers (0: the default for k, 11 up
        to k = 32, then k - 17 between 16 and 32; else 1..32, at most k).  Every setting
        gives the same results."""
        t = _abi.kc_tune(windows_per_thread=windows_per_thread, blocks_per_cu=blocks_per_cu,
                         table_load=float(table_load), insert_path=insert_path, setter_path=setter_path,
                         kmers_per_bin=kmers_per_bin, minimizer_len=minimizer_len)
        self._chk(lib().kc_tune_set(self._h, ctypes.byref(t)))

    def autotune(self, reps=2):
        """Coordinate search over the knobs of the Stage 2 path the run takes (binned
        count: k-mers per bin and workgroups per CU; table insert: load factor, windows per
        thread and workgroups per CU), then the first-setter path, on the uploaded
        libraries; every setting must give the same results; keeps and returns the fastest."""
        t = _abi.kc_tune()
        self._chk(lib().kc_autotune(self._h, int(reps), ctypes.byref(t)))
        return t.as_dict()

    def set_quals(self, quals, qoffsets, library=0):
        """Per-base Phred qualities (features of the learned first pass)."""
        q = np.ascontiguousarray(quals, dtype=np.uint8)
        o = np.ascontiguousarray(qoffsets, dtype=np.uint64)
        qq = q if q.shape[0] else np.zeros(1, np.uint8)
        self._chk(lib().kc_library_quals_set(self._h, library, ptr(qq, u8p), q.shape[0], ptr(o, u64p)))

    def features(self, library=0):
        """(n_counts, 16) float32 features of the count-table entries, in their sorted order."""
        n = self.info(library)["n_counts"]
        f = np.zeros((max(n, 1), 16), dtype=np.float32)
        self._chk(lib().kc_library_features(self._h, library, ptr(f, _abi.f32p), n))
        return f[:n]

    def learned_select(self, mlp, margin=0.05, library=0):
        """(labels, fallback mask, device ms) of the learned first pass over the entries."""
        n = self.info(library)["n_counts"]
        lab = np.zeros(max(n, 1), dtype=np.uint8)
        fb = np.zeros(max(n, 1), dtype=np.uint8)
        ms = ctypes.c_double(0)
        self._chk(lib().kc_library_learned_select(self._h, library, mlp._h, float(margin), ptr(lab, u8p), ptr(fb, u8p),
                                                  n, ctypes.byref(ms)))
        return lab[:n].astype(bool), fb[:n].astype(bool), ms.value

    def verify(self, library=0):
        """Closed-form checks of the last run recomputed window by window on the device
        (kc_library_verify): seen-tier recount, Bloom-solid windows against the sum of the
        counts, sortedness, solid flags, and a checksum of the sorted table."""
        v = _abi.kc_verify()
        self._chk(lib().kc_library_verify(self._h, library, ctypes.byref(v)))
        return v.as_dict()

    def cutoff_fragile(self, library=0, resamples=32, seed=0):
        """The learned pass's library-level margin (kc_library_cutoff_fragile): True when the
        library's cutoff is too fragile for a learned selection."""
        f = ctypes.c_int32(0)
        self._chk(lib().kc_library_cutoff_fragile(self._h, library, int(resamples), int(seed), ctypes.byref(f)))
        return bool(f.value)

    def info(self, library=0):
        info = _abi.kc_info()
        self._chk(lib().kc_library_info(self._h, library, ctypes.byref(info)))
        return info.as_dict()

    def bloom(self, library=0):
        nw = self.info(library)["num_words"]
        seen = np.zeros(nw, dtype=np.uint64)
        solid = np.zeros(nw, dtype=np.uint64)
        self._chk(lib().kc_library_bloom_export(self._h, library, ptr(seen, u64p), ptr(solid, u64p), nw))
        return seen, solid

    def query(self, keys4, library=0):
        k4 = _keys4(keys4)
        out = np.zeros(k4.shape[0], dtype=np.uint8)
        self._chk(lib().kc_library_bloom_query(self._h, library, ptr(k4, u64p), k4.shape[0], ptr(out, u8p)))
        return out.astype(bool)

    def counts(self, library=0):
        n = self.info(library)["n_counts"]
        keys = np.zeros((max(n, 1), 4), dtype=np.uint64)
        cnt = np.zeros(max(n, 1), dtype=np.uint32)
        mask = np.zeros(max(n, 1), dtype=np.uint8)
        self._chk(lib().kc_library_counts_export(self._h, library, ptr(keys, u64p), ptr(cnt, u32p), ptr(mask, u8p),
                                                 n))
        return keys[:n], cnt[:n], mask[:n]

    def reads(self, library=0):
        dl = self._data_len[library]
        out = np.zeros(max(dl, 1), dtype=np.uint8)
        self._chk(lib().kc_library_reads_export(self._h, library, ptr(out, u8p), dl))
        return out[:dl]

    def unitigs(self, library=0):
        """CompactedGraph::build (graph.rs:65-234) over the library's solid set, on the GPU."""
        from .graph import CompactedGraph
        return CompactedGraph.build(self, library)

    def kernel_times(self):
        names = (ctypes.c_char_p * 32)()
        ms = (ctypes.c_double * 32)()
        launches = (ctypes.c_uint32 * 32)()
        n = lib().kc_kernel_times(self._h, names, ms, launches, 32)
        return {names[i].decode(): (ms[i], launches[i]) for i in range(max(n, 0))}


def run_many(engines):
    """kc_run of several engines at once (kc_run_many): one host thread and stream per
    engine, each engine's results those of its run() alone.  Returns their infos."""
    engines = list(engines)
    n = len(engines)
    hs = (_abi._vp * n)(*[e._h for e in engines])
    infos = (_abi.kc_info * n)()
    st = lib().kc_run_many(hs, n, infos)
    if st != _abi.KC_OK:
        msgs = [lib().kc_last_error(e._h) for e in engines]
        raise KcError(st, "; ".join(m.decode() for m in msgs if m))
    return [infos[i].as_dict() for i in range(n)]


def count_kmers(packed, config, device=0):
    """Stage 2 + Stage 3 of run_assembly_with_packed_reads (assemble.rs:129-298) on the GPU."""
    eng = SpectrumEngine(config, device=device)
    try:
        eng.upload(packed)
        priors = config.all_priors()
        if priors:
            eng.set_priors(priors)
        info = eng.run()
        seen, solid = eng.bloom()
        keys, cnt, mask = eng.counts()
        return KmerSpectrum(k=config.k, seen=seen, solid_bits=solid, keys=keys, counts=cnt, solid_mask=mask,
                            cutoff=info["cutoff"], info=info)
    finally:
        eng.close()
