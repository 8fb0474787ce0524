"""Python restatement of the learned first pass's features (TEST INFRASTRUCTURE ONLY: the
checker for the device feature kernels; small inputs).

For a count-table entry x of count c (canonical k-mer, uncorrected table):
  q    = (sum over x's occurrences of the k base qualities) / (c k)    (0 without qualities)
  n_y  = raw count of each of the 3k Hamming-1 neighbours y (canonical), 0 when absent
  mx, mx2 = largest and second largest n_y;  cutoff, top = the library's cutoff and
  95th-percentile count (as reported, top capped at 1023 and at least 1)
  f0 = log2(1+c)            f1 = q/40              f2 = log2(1+mx)
  f3 = log2(1+sum n_y)      f4 = #(n_y>0)/(3k)     f5 = c/(c+mx)
  f6 = f0 - log2(1+cutoff)  f7 = [c >= cutoff]     f8 = min(#(n_y >= cutoff), 4)/4
  f9 = f0 - f2              f10 = (q/40)^2         f11 = log2(1+top)/8
  f12 = c/top               f13 = min(#(n_y > c), 4)/4   f14 = log2(1+mx2)   f15 = 1
"""
import math
from collections import Counter

import numpy as np

from . import hash_np as H

_DEC = "ACGT"


def kmer_string(w, k):
    v = int(w[0]) | (int(w[1]) << 64)
    return "".join(_DEC[(v >> (2 * (k - 1 - i))) & 3] for i in range(k))


def window_stats(store, k, quals=None, qoff=None):
    """Raw counts and quality sums of every canonical window of a PackedReads store."""
    cnt, qs = Counter(), Counter()
    for i in range(len(store)):
        r = store.get_read(i)
        q = None if quals is None else quals[int(qoff[i]):int(qoff[i]) + len(r)].astype(np.int64)
        for j in range(len(r) - k + 1):
            w = H.canonical_words_of_string(r[j:j + k])
            cnt[w] += 1
            if q is not None:
                qs[w] += int(q[j:j + k].sum())
    return cnt, qs


def features(keys, counts, k, cutoff, top, cnt, qs):
    top = max(int(top), 1)
    out = np.zeros((len(keys), 16), dtype=np.float64)
    for t, w in enumerate(keys):
        w = tuple(int(x) for x in w)
        c = float(counts[t])
        q = qs.get(w, 0) / (c * k) if c > 0 else 0.0
        s = kmer_string(w, k)
        ns = []
        for pos in range(k):
            for b in "ACGT":
                if b == s[pos]:
                    continue
                y = s[:pos] + b + s[pos + 1:]
                ns.append(cnt.get(tuple(H.canonical_words_of_string(y.encode())), 0))
        srt = sorted(ns, reverse=True)
        mx, mx2 = srt[0], srt[1]
        lc, lm = math.log2(1 + c), math.log2(1 + mx)
        out[t] = [lc, q / 40, lm, math.log2(1 + sum(ns)), sum(n > 0 for n in ns) / (3 * k), c / (c + mx),
                  lc - math.log2(1 + cutoff), float(c >= cutoff), min(sum(n >= cutoff for n in ns), 4) / 4, lc - lm,
                  (q / 40) ** 2, math.log2(1 + top) / 8, c / top, min(sum(n > c for n in ns), 4) / 4,
                  math.log2(1 + mx2), 1.0]
    return out
