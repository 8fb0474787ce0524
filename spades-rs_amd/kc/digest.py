"""SHA-256 digests of a spectrum run's outputs, for comparing full-size runs with a
recorded result without shipping the arrays.

Each output is hashed in a fixed byte layout: Bloom tiers as little-endian u64 words,
the sorted count table as (n, 4) little-endian u64 Kmer256 limbs (dna.rs:137-138, limb
order of bloom.rs:62-65), counts as little-endian u32, the solid mask as one byte per
entry, corrected reads (when Hammer ran) as the packed bytes.
"""
import hashlib

import numpy as np


def sha(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def spectrum_digests(seen, solid, keys, counts, mask, cutoff, n_occ, corrected=None, corrected_reads=None):
    d = {"seen": sha(seen, "<u8"), "solid": sha(solid, "<u8"), "keys": sha(np.asarray(keys).reshape(-1, 4), "<u8"),
         "counts": sha(counts, "<u4"), "mask": sha(mask, "u1"), "n_counts": int(len(counts)),
         "n_solid": int(np.count_nonzero(mask)), "cutoff": int(cutoff), "n_occ": int(n_occ)}
    if corrected is not None:
        d["corrected"] = sha(corrected, "u1")
        d["corrected_reads"] = int(corrected_reads)
    return d


def engine_digests(eng, info, library=0, with_reads=False):
    """Digests of a finished SpectrumEngine run (one library)."""
    seen, solid = eng.bloom(library)
    keys, cnt, mask = eng.counts(library)
    corr = eng.reads(library) if with_reads else None
    return spectrum_digests(seen, solid, keys, cnt, mask, info["cutoff"], info["n_occ"], corr,
                            info.get("corrected_reads") if with_reads else None)


OUTPUTS = ("seen", "solid", "keys", "counts", "mask", "n_counts", "n_solid", "cutoff", "n_occ", "corrected",
           "corrected_reads")


def diff(got, want):
    """Names of the output fields recorded in `want` that differ in `got` (empty when equal)."""
    return [f for f in OUTPUTS if f in want and got.get(f) != want[f]]


M64 = (1 << 64) - 1


def _fmix(x):
    x = x ^ (x >> np.uint64(33))
    x = x * np.uint64(0xff51afd7ed558ccd)
    x = x ^ (x >> np.uint64(33))
    x = x * np.uint64(0xc4ceb9fe1a85ec53)
    return x ^ (x >> np.uint64(33))


def table_checksum(values, counts):
    """The checksum kc_library_verify reports for a sorted table: the sum mod 2^64 over
    entries i of mix(i, V words, count), with V the k-mer value in four words, lowest first
    (for k <= 64 the Kmer256 limbs)."""
    v = np.ascontiguousarray(values, dtype=np.uint64).reshape(-1, 4)
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    i = np.arange(v.shape[0], dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = _fmix(i + np.uint64(0x9E3779B97F4A7C15) * (v[:, 0] + np.uint64(1)))
        for w in (1, 2, 3):
            x = _fmix(x ^ v[:, w])
        x = _fmix(x ^ c)
        return int(x.sum(dtype=np.uint64))
