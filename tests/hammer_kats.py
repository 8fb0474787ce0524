"""Known-answer cases for ErrorCorrector::correct_read_in_place (hammer.rs:34-120) on
hand-built solid sets.  Each case lists the k-mers made solid (each inserted twice into
an otherwise empty filter), the read, and the result the rule fixes:

  for each window i in order (over the read as corrected so far) that is not solid, for
  each offset o and each other base c in ACGT order: if window i with c at i + o is
  solid, gain = (1 + solid windows s != i covering i + o, with c) - (solid windows s != i
  covering i + o, as read), saturating at 0; the first strictly greater gain wins.

Every expected result below is derived by hand from that rule in the case's comment, and
each case is chosen so that one plausible mistake gives a different answer.  The filter
must show no false positive on the windows a case depends on (checked by `preconditions`).
"""

# name, k, solid k-mers, read, expected (modified, result)
CASES = [
    # One window (L = k): every candidate's scan range is {i}, so each solid candidate has
    # gain 1 and the first in (offset, ACGT) order must win.  Last-max picks AAAAT.
    ("tie_first_offset", 5, ["CAAAA", "AAAAT"], "AAAAA", (True, "CAAAA")),
    # Same offset, two bases: C precedes G in ACGT order.
    ("tie_acgt_order", 5, ["GAAAA", "CAAAA"], "AAAAA", (True, "CAAAA")),
    # Offset before base: (0, T) precedes (4, C) although C < T.
    ("tie_offset_before_base", 5, ["TAAAA", "AAAAC"], "AAAAA", (True, "TAAAA")),
    # L = 6: (0, C) has gain 1; (3, T) makes window 0 AAATA (solid) but window 1 AATAA is
    # not, gain 1; (4, T) makes window 0 AAAAT and window 1 AAATA solid, gain 2 > 1 wins.
    # Then window 1 (AAATA) is solid.  Taking the first solid candidate gives CAAAAA.
    ("larger_gain_later", 5, ["CAAAA", "AAAAT", "AAATA"], "AAAAAA", (True, "AAAATA")),
    # L = 9, k = 7: windows 1 and 2 are solid, window 0 is not.  The only candidate that
    # makes window 0 solid is T at its last base (position 6), which breaks windows 1 and 2:
    # gain = 1 - 2, saturating to 0, so nothing changes.  Wrapping arithmetic would apply it.
    ("saturated_gain", 7, ["GATTCAT", "ATTCAGG", "TTCAGGA"], "GATTCAGGA", (False, "GATTCAGGA")),
]

# Two errors whose windows overlap, with no window holding the second error alone: the
# read (23 bases, k = 11) ends with the window that starts at the first error p1 = 12, so
# every window containing the second error (p2 = 15) also contains the first.
# Sequential correction fixes p1 at window 2 (gain 3: itself and windows 3, 4; windows
# 5..12 still hold p2), then, because windows 5..12 see the corrected p1, p2 at window 5
# (gain 8: windows 5..12).  Windows scanned over the uncorrected read would leave p2.
# The solid set is every k-mer of a fixed clean sequence.
DEPENDENT = {
    "name": "dependent_errors", "k": 11,
    "clean": "TGCATCAGGTACCTGAAGCTTGGATCCCAGTTAGCAAGCTGACTTACGGTACCATTG",
    "read_start": 20, "p1": 12, "p2": 15, "subs": {12: "C", 15: "G"},
}


def dependent_case():
    d = DEPENDENT
    k, p1 = d["k"], d["p1"]
    L = p1 + k
    clean = d["clean"][d["read_start"]:d["read_start"] + L]
    read = list(clean)
    for p, b in d["subs"].items():
        assert read[p] != b
        read[p] = b
    kmers = [d["clean"][i:i + k] for i in range(len(d["clean"]) - k + 1)]
    mid = list(read)
    mid[p1] = clean[p1]
    return (d["name"], k, kmers, "".join(read), (True, clean), ["".join(mid)])


def all_cases():
    """(name, k, solid k-mers, read, (modified, result), intermediate reads)."""
    return [c + ([],) for c in CASES] + [dependent_case()]


def revcomp(s):
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


def neighbourhood(k, read):
    """Every window of the read, and every single-base variant of every window: the
    k-mers whose solidity the rule can query on this read (while at most two
    substitutions separate it from the answer's windows)."""
    out = set()
    for i in range(len(read) - k + 1):
        w = read[i:i + k]
        out.add(w)
        for o in range(k):
            for c in "ACGT":
                out.add(w[:o] + c + w[o + 1:])
    return out


def preconditions(k, kmers, read, expected, states, is_solid):
    """The filter answers solid exactly for the listed k-mers (either strand) among the
    k-mers of the neighbourhoods of the read, the intermediate reads and the result, and
    no reverse complement of a listed k-mer lies there unless listed itself (so the
    hand derivations, which use the listed strings, hold)."""
    want = set(kmers) | {revcomp(x) for x in kmers}
    near = set()
    for seq in [read, expected[1]] + list(states):
        near |= neighbourhood(k, seq)
    assert not (near & ({revcomp(x) for x in kmers} - set(kmers)))
    for w in near:
        assert is_solid(w) == (w in want), (w, w in want)
