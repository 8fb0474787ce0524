"""Second, independent transcription of the reference Bloom arithmetic in NumPy.

TEST INFRASTRUCTURE ONLY.  It restates bloom.rs:61-126 (hash_kmer256, the
h % size_bits -> (word, bit) split, insert_kmer256 and is_solid_kmer256) and
dna.rs:65-78/118-132/299-321 (2-bit encoding, reverse complement, canonical
form) directly from the Rust source text, without sharing code with
oracle/oracle.cpp, so that the two transcriptions can be checked against each
other on random keys.  NumPy uint64 arithmetic wraps modulo 2**64, which is the
reference's wrapping_mul.
"""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
SEEDS = (np.uint64(0x517CC1B727220A95), np.uint64(0x9E3779B97F4A7C15), np.uint64(0x6C62272E07BB0142))


def _rotl(x, r):
    r = np.uint64(r)
    return (x << r) | (x >> (np.uint64(64) - r))


def hash_kmer256(w0, w1, w2, w3, seed):
    """bloom.rs:61-78 over arrays of limbs."""
    with np.errstate(over="ignore"):
        w0, w1, w2, w3 = (np.asarray(a, dtype=np.uint64) for a in (w0, w1, w2, w3))
        s33 = np.uint64(33)
        x = w0 ^ np.uint64(seed)
        x = (x ^ (_rotl(w1, 21) * np.uint64(0x9E3779B97F4A7C15))) * np.uint64(0xFF51AFD7ED558CCD)
        x ^= x >> s33
        x = (x ^ (_rotl(w2, 43) * np.uint64(0xBF58476D1CE4E5B9))) * np.uint64(0xC4CEB9FE1A85EC53)
        x ^= x >> s33
        x = (x ^ (_rotl(w3, 11) * np.uint64(0x94D049BB133111EB))) * np.uint64(0xFF51AFD7ED558CCD)
        x ^= x >> s33
        return x


def size_bits_for(capacity_bits):
    """bloom.rs:26-28: num_words = max(ceil(bits/64), 1); size_bits = 64 * num_words."""
    nw = max((int(capacity_bits) + 63) // 64, 1)
    return nw * 64, nw


def positions(keys4, size_bits):
    """Bit positions h_j % size_bits for j = 1, 2, 3 (bloom.rs:84-86). keys4: (n, 4) uint64."""
    keys4 = np.asarray(keys4, dtype=np.uint64).reshape(-1, 4)
    out = np.empty((3, keys4.shape[0]), dtype=np.uint64)
    for j, s in enumerate(SEEDS):
        out[j] = hash_kmer256(keys4[:, 0], keys4[:, 1], keys4[:, 2], keys4[:, 3], s) % np.uint64(size_bits)
    return out


def seen_tier(occ_keys4, capacity_bits):
    """Closed form of the seen tier: OR over every occurrence of its 3 bits (order-free)."""
    size_bits, nw = size_bits_for(capacity_bits)
    pos = positions(occ_keys4, size_bits).reshape(-1)
    words = np.zeros(nw, dtype=np.uint64)
    np.bitwise_or.at(words, (pos >> np.uint64(6)).astype(np.int64), np.uint64(1) << (pos & np.uint64(63)))
    return words


def sequential_two_tier(occ_keys4, capacity_bits):
    """Pure-Python sequential insert_kmer256 (bloom.rs:83-107) over occurrences in order.

    Small inputs only.  Returns (seen, solid, promoted flags)."""
    size_bits, nw = size_bits_for(capacity_bits)
    pos = positions(occ_keys4, size_bits)
    seen = [0] * nw
    solid = [0] * nw
    promoted = []
    for t in range(pos.shape[1]):
        prev = []
        for j in range(3):
            p = int(pos[j, t])
            w, b = p >> 6, 1 << (p & 63)
            prev.append(seen[w] & b)
            seen[w] |= b
        if all(prev):
            for j in range(3):
                p = int(pos[j, t])
                solid[p >> 6] |= 1 << (p & 63)
            promoted.append(True)
        else:
            promoted.append(False)
    return (np.array(seen, dtype=np.uint64), np.array(solid, dtype=np.uint64), np.array(promoted))


def is_solid(solid_words, keys4, capacity_bits):
    """bloom.rs:111-126 over arrays."""
    size_bits, _ = size_bits_for(capacity_bits)
    pos = positions(keys4, size_bits)
    ok = np.ones(pos.shape[1], dtype=bool)
    for j in range(3):
        w = solid_words[(pos[j] >> np.uint64(6)).astype(np.int64)]
        ok &= ((w >> (pos[j] & np.uint64(63))) & np.uint64(1)).astype(bool)
    return ok


# --- dna.rs, independent string-level forms --------------------------------

_CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3,
         ord("a"): 0, ord("c"): 1, ord("g"): 2, ord("t"): 3}
_COMP = {ord("A"): ord("T"), ord("C"): ord("G"), ord("G"): ord("C"), ord("T"): ord("A")}


def encode_kmer(s):
    """2-bit big-endian integer of an ACGT string (dna.rs:299-321 for k <= 64)."""
    v = 0
    for ch in s:
        v = (v << 2) | _CODE[ch]
    return v


def kmer_int_to_words(v, k):
    """Kmer256 limbs for k <= 64: lo = v (u128), hi = 0; for k > 64 split first 64 bases / rest."""
    if k <= 64:
        return (v & 0xFFFFFFFFFFFFFFFF, (v >> 64) & 0xFFFFFFFFFFFFFFFF, 0, 0)
    lo = v >> (2 * (k - 64))
    hi = v & ((1 << (2 * (k - 64))) - 1)
    return (lo & 0xFFFFFFFFFFFFFFFF, lo >> 64, hi & 0xFFFFFFFFFFFFFFFF, hi >> 64)


def revcomp_string(s):
    return bytes(_COMP[c] for c in reversed(s))


def canonical_words_of_string(s):
    """Canonical Kmer256 limbs of an ACGT window: min over (lo, hi) order of fwd/rc (dna.rs:185-192)."""
    k = len(s)
    f = kmer_int_to_words(encode_kmer(s), k)
    r = kmer_int_to_words(encode_kmer(revcomp_string(s)), k)
    fk = ((f[1] << 64) | f[0], (f[3] << 64) | f[2])
    rk = ((r[1] << 64) | r[0], (r[3] << 64) | r[2])
    return f if fk <= rk else r
