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
