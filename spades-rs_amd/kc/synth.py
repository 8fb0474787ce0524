"""Seeded synthetic read sets in the PackedReads layout.

Workloads (all seeded with numpy's PCG64 so a seed fixes every base):
  uniform_genome_pairs  — a random genome (independent bases, given GC fraction),
                          paired reads from fragments of length ~ N(insert_mean,
                          insert_sd), R2 the reverse complement of the fragment's
                          last read_len bases, uniform or position-dependent
                          substitution errors (the substituted base is uniform over
                          the other three), Phred qualities.
The store is laid out as the reference builds it from paired files: all R1 mates,
then all R2 mates (packed_reads.rs:481-489).
"""
import numpy as np

from .packed_reads import PackedReads, pack_codes, pack_fixed_length


def random_genome(rng, size, gc=0.5):
    """Independent bases with P(G)=P(C)=gc/2, P(A)=P(T)=(1-gc)/2; codes A0 C1 G2 T3."""
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    if abs(gc - 0.5) < 1e-12:
        return rng.integers(0, 4, size, dtype=np.uint8)
    return rng.choice(4, size=size, p=p).astype(np.uint8)


def error_profile(read_len, p_start, p_end=None):
    """Per-position substitution probability: constant, or linear from p_start to p_end."""
    if p_end is None:
        return np.full(read_len, p_start, dtype=np.float64)
    return np.linspace(p_start, p_end, read_len)


def _mutate(rng, codes, perr):
    mask = rng.random(codes.shape) < perr[None, :]
    sub = rng.integers(1, 4, codes.shape, dtype=np.uint8)
    return np.where(mask, (codes + sub) & 3, codes).astype(np.uint8)


def uniform_genome_pairs(seed, genome_size, n_pairs, read_len=150, insert_mean=300.0, insert_sd=30.0,
                         p_start=0.005, p_end=None, gc=0.5, qual=(30, 40), chunk=1 << 17, genome=None,
                         with_quals=False):
    """Returns (PackedReads, meta).  meta holds the genome and, with_quals, the Phred
    quality sum of every read (R1 block then R2 block)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if genome is None:
        genome = random_genome(rng, genome_size, gc)
    G = genome.shape[0]
    perr = error_profile(read_len, p_start, p_end)
    r1_parts, r2_parts, q1, q2 = [], [], [], []
    ar = np.arange(read_len, dtype=np.int64)
    for c0 in range(0, n_pairs, chunk):
        m = min(chunk, n_pairs - c0)
        frag = np.rint(rng.normal(insert_mean, insert_sd, m)).astype(np.int64)
        frag = np.clip(frag, read_len, G)
        start = (rng.random(m) * (G - frag + 1)).astype(np.int64)
        r1 = genome[start[:, None] + ar[None, :]]
        tail = start + frag - read_len
        r2 = 3 - genome[tail[:, None] + ar[None, ::-1]]
        r1 = _mutate(rng, r1, perr)
        r2 = _mutate(rng, r2, perr)
        r1_parts.append(pack_fixed_length(r1)[0])
        r2_parts.append(pack_fixed_length(r2)[0])
        if with_quals:
            if p_end is None:
                lo, hi = qual
                q1.append(rng.integers(lo, hi + 1, (m, read_len)).sum(axis=1))
                q2.append(rng.integers(lo, hi + 1, (m, read_len)).sum(axis=1))
            else:
                qpos = np.rint(-10.0 * np.log10(perr)).astype(np.int64)
                q1.append(np.full(m, qpos.sum()))
                q2.append(np.full(m, qpos.sum()))
    nb = (read_len + 3) // 4
    data = np.concatenate(r1_parts + r2_parts) if n_pairs else np.zeros(0, np.uint8)
    offsets = np.arange(2 * n_pairs, dtype=np.uint64) * np.uint64(nb)
    lengths = np.full(2 * n_pairs, read_len, dtype=np.uint32)
    store = PackedReads.from_arrays(data, offsets, lengths, pe_count=n_pairs)
    meta = {"genome": genome, "seed": seed, "n_pairs": n_pairs, "read_len": read_len}
    if with_quals:
        meta["qual_sum"] = np.concatenate(q1 + q2) if n_pairs else np.zeros(0, np.int64)
    return store, meta


def workload_1mb_30x(seed=1, scale=1.0):
    """1 Mb uniform genome, 2x150 bp pairs at 30x (100,000 pairs), insert N(300, 30),
    0.5% uniform substitutions, Phred 30-40; `scale` shrinks genome and pairs together."""
    g = max(int(1_000_000 * scale), 2000)
    return uniform_genome_pairs(seed, g, max(int(100_000 * scale), 1), p_start=0.005)


def workload_5mb_100x(seed=2, scale=1.0):
    """5 Mb genome at 50% GC, 2x150 bp pairs at 100x (1,666,667 pairs), substitution rate
    rising linearly 0.1% -> 1.0% along each read with matching Phred qualities."""
    g = max(int(5_000_000 * scale), 2000)
    return uniform_genome_pairs(seed, g, max(int(1_666_667 * scale), 1), p_start=0.001, p_end=0.010)


def workload_metagenome(seed=4, scale=1.0, n_genomes=20, n_pairs=30_000_000, sigma=1.0, p_err=0.005, progress=None):
    """A metagenome sample: n_genomes random genomes of 1-6 Mb (uniform; x scale), cell
    abundances lognormal(0, sigma), n_pairs x scale 2x150 bp pairs drawn from the genomes
    in proportion to abundance x genome length (insert N(300, 30), uniform substitutions
    at p_err), the pairs in a seeded random order; one library, R1 mates then R2 mates.
    progress(g, pairs) is called after genome g's pairs are drawn.  Returns (PackedReads, meta)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.maximum((rng.uniform(1e6, 6e6, n_genomes) * scale).astype(np.int64), 2000)
    ab = np.exp(sigma * rng.standard_normal(n_genomes))
    total = max(int(n_pairs * scale), n_genomes)
    w = ab * sizes
    pairs = np.floor(w / w.sum() * total).astype(np.int64)
    pairs[int(np.argmax(w))] += total - int(pairs.sum())
    nb = 38  # bytes of a 150-base read
    r1, r2 = [], []
    for g in range(n_genomes):
        genome = random_genome(rng, int(sizes[g]), 0.5)
        st, _ = uniform_genome_pairs(seed=seed * 1000 + g, genome_size=int(sizes[g]), n_pairs=int(pairs[g]),
                                     p_start=p_err, genome=genome)
        d = st.data.reshape(2 * int(pairs[g]), nb)
        r1.append(d[: pairs[g]])
        r2.append(d[pairs[g]:])
        del st
        if progress:
            progress(g, int(pairs[g]))
    perm = rng.permutation(total)
    data = np.concatenate([np.concatenate(r1)[perm].reshape(-1), np.concatenate(r2)[perm].reshape(-1)])
    del r1, r2
    offsets = np.arange(2 * total, dtype=np.uint64) * np.uint64(nb)
    lengths = np.full(2 * total, 150, dtype=np.uint32)
    store = PackedReads.from_arrays(data, offsets, lengths, pe_count=total)
    return store, {"genome_sizes": sizes, "abundance": ab, "pairs": pairs, "n_pairs": total}


def uneven_pairs(rng, genome, n_pairs, median_len=150.0, sigma=0.25, min_len=50, max_len=250, p_err=0.003,
                 insert_sd=30.0, chunk=1 << 16):
    """Paired reads whose lengths are lognormal(median_len, sigma) truncated to
    [min_len, max_len], from fragments ~ N(2.2 * median_len, insert_sd) (at least max_len),
    uniform substitutions at p_err.  R1 mates then R2 mates, packed contiguously."""
    G = genome.shape[0]
    ar = np.arange(max_len, dtype=np.int64)
    perr = np.full(max_len, p_err)
    nbmax = (max_len + 3) // 4
    parts, lens1, lens2 = [[], []], [], []
    for c0 in range(0, n_pairs, chunk):
        m = min(chunk, n_pairs - c0)
        L = np.clip(np.rint(median_len * np.exp(sigma * rng.standard_normal((2, m)))), min_len, max_len).astype(np.int64)
        frag = np.clip(np.rint(rng.normal(2.2 * median_len, insert_sd, m)).astype(np.int64), max_len, G)
        start = (rng.random(m) * (G - frag + 1)).astype(np.int64)
        r1 = _mutate(rng, genome[start[:, None] + ar[None, :]], perr)
        tail = start + frag - max_len
        r2 = _mutate(rng, 3 - genome[tail[:, None] + ar[None, ::-1]], perr)
        # R2 reads the fragment from its end: keep its first L bases as well; each read's
        # bases past L are zeroed and its packed row cut to ceil(L / 4) bytes
        for side, (rr, LL) in enumerate(((r1, L[0]), (r2, L[1]))):
            rr = np.where(ar[None, :] < LL[:, None], rr, 0).astype(np.uint8)
            packed = pack_fixed_length(rr)[0].reshape(m, nbmax)
            parts[side].append(packed[np.arange(nbmax)[None, :] < ((LL + 3) // 4)[:, None]])
        lens1.append(L[0])
        lens2.append(L[1])
    lengths = np.concatenate([np.concatenate(lens1) if lens1 else np.zeros(0, np.int64),
                              np.concatenate(lens2) if lens2 else np.zeros(0, np.int64)])
    data = np.concatenate(parts[0] + parts[1]) if n_pairs else np.zeros(0, np.uint8)
    nbytes = (lengths + 3) // 4
    offsets = np.zeros(lengths.shape[0], dtype=np.uint64)
    if lengths.shape[0]:
        offsets[1:] = np.cumsum(nbytes)[:-1].astype(np.uint64)
    return PackedReads.from_arrays(data, offsets, lengths.astype(np.uint32), pe_count=n_pairs)


def workload_64_libraries(seed=3, scale=1.0, n_libs=64, progress=None):
    """Many independent libraries: genome sizes log-uniform in [5 kb, 6 Mb], coverage
    uniform in [20x, 200x], read lengths lognormal (median 150) truncated to [50, 250],
    0.3% substitutions; `scale` multiplies every genome size (coverage is kept).
    progress (optional) is called with (index, store) after each library.  Returns a list
    of (PackedReads, meta)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(np.log(5_000), np.log(6_000_000), n_libs))
    covs = rng.uniform(20.0, 200.0, n_libs)
    out = []
    for i in range(n_libs):
        g = max(int(sizes[i] * scale), 600)
        genome = random_genome(rng, g, 0.5)
        n_pairs = max(int(round(covs[i] * g / (2 * 155.0))), 1)
        store = uneven_pairs(rng, genome, n_pairs)
        out.append((store, {"genome_size": g, "coverage": float(covs[i]), "n_pairs": n_pairs}))
        if progress:
            progress(i, store)
    return out


def repeat_genome(rng, size, repeat_frac=0.2, fam_len=(300, 6000), copies=(2, 50), divergence=(0.01, 0.05)):
    """A random genome in which about repeat_frac of the bases are copies of repeat
    families (lengths log-uniform in fam_len, copies uniform in copies, each copy a
    mutated version of its family consensus at a per-copy divergence uniform in
    divergence)."""
    g = random_genome(rng, size, 0.5)
    covered = 0
    target = int(repeat_frac * size)
    while covered < target:
        L = int(np.exp(rng.uniform(np.log(fam_len[0]), np.log(fam_len[1]))))
        L = min(L, size // 4)
        cons = random_genome(rng, L, 0.5)
        for _ in range(int(rng.integers(copies[0], copies[1] + 1))):
            d = rng.uniform(*divergence)
            copy = _mutate(rng, cons[None, :], np.full(L, d))[0]
            pos = int(rng.integers(0, size - L + 1))
            g[pos:pos + L] = copy
            covered += L
            if covered >= target:
                break
    return g


def repeat_pairs_with_quals(seed, genome_size=12_000_000, coverage=60.0, read_len=150, p_err=0.003,
                            q_good=(30, 40), q_bad=(5, 20), chunk=1 << 17):
    """Paired 2x150 reads at `coverage` from a repeat_genome, uniform substitutions at
    p_err, and a Phred quality per base: uniform in q_good for correct bases and in q_bad
    for substituted ones.  Returns (PackedReads, quals, meta): quals holds one byte per
    base, read i's bases at quals[qoff[i] : qoff[i] + len_i] with qoff = meta['qoff']."""
    rng = np.random.Generator(np.random.PCG64(seed))
    genome = repeat_genome(rng, genome_size)
    G = genome.shape[0]
    n_pairs = max(int(round(coverage * G / (2 * read_len))), 1)
    ar = np.arange(read_len, dtype=np.int64)
    parts = [[], []]
    qparts = [[], []]
    for c0 in range(0, n_pairs, chunk):
        m = min(chunk, n_pairs - c0)
        frag = np.clip(np.rint(rng.normal(300.0, 30.0, m)).astype(np.int64), read_len, G)
        start = (rng.random(m) * (G - frag + 1)).astype(np.int64)
        r1 = genome[start[:, None] + ar[None, :]]
        tail = start + frag - read_len
        r2 = 3 - genome[tail[:, None] + ar[None, ::-1]]
        for side, r in enumerate((r1, r2)):
            err = rng.random(r.shape) < p_err
            sub = rng.integers(1, 4, r.shape, dtype=np.uint8)
            rr = np.where(err, (r + sub) & 3, r).astype(np.uint8)
            q = np.where(err, rng.integers(q_bad[0], q_bad[1] + 1, r.shape),
                         rng.integers(q_good[0], q_good[1] + 1, r.shape)).astype(np.uint8)
            parts[side].append(pack_fixed_length(rr)[0])
            qparts[side].append(q.reshape(-1))
    nb = (read_len + 3) // 4
    data = np.concatenate(parts[0] + parts[1])
    quals = np.concatenate(qparts[0] + qparts[1])
    n = 2 * n_pairs
    offsets = np.arange(n, dtype=np.uint64) * np.uint64(nb)
    lengths = np.full(n, read_len, dtype=np.uint32)
    store = PackedReads.from_arrays(data, offsets, lengths, pe_count=n_pairs)
    qoff = np.arange(n, dtype=np.uint64) * np.uint64(read_len)
    return store, quals, {"genome": genome, "n_pairs": n_pairs, "qoff": qoff, "seed": seed}
