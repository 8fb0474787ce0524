"""Compacted unitigs on the GPU: CompactedGraph::build (graph.rs:65-234) over the solid set
an engine context holds after its count (assemble.rs:286-301 passes the solid k-mers and
the count table, which is where the coverage comes from).

Unitigs come ordered by their first k-mer's rank in the sorted count table; a linear
unitig may be the reverse complement of the one the reference's hash-ordered walk emits
(the same unitig of the bidirected graph), and a cycle starts at its smallest k-mer on the
forward strand.  Odd k only.
"""
import ctypes
from dataclasses import dataclass
from typing import List

import numpy as np

from . import _abi
from ._abi import KcError, lib, ptr, u8p, u64p


@dataclass
class Unitig:
    """graph.rs:43-53."""
    id: int
    sequence: bytes
    mean_coverage: float
    kmers_count: int


class CompactedGraph:
    """graph.rs:56-61: k and the maximal unitigs, held as flat arrays (seq bytes, offsets
    with n + 1 entries, k-mer counts, mean coverage)."""

    STAGES = ("decode+index", "adjacency+links", "ranking", "emit")

    def __init__(self, k, seq, offsets, kmers, mean_cov, ms_device=0.0, ms_stage=None):
        self.k = k
        self.seq = seq
        self.offsets = offsets
        self.kmers = kmers
        self.mean_cov = mean_cov
        self.ms_device = ms_device
        self.ms_stage = dict(zip(self.STAGES, ms_stage or [0.0] * 4))

    @classmethod
    def build(cls, engine, library=0):
        """CompactedGraph::build for `library` of a SpectrumEngine after count()/run()."""
        h = ctypes.c_void_p()
        engine._chk(lib().kc_unitigs_build(engine._h, library, ctypes.byref(h)))
        try:
            info = _abi.kc_unitigs_info()
            st = lib().kc_unitigs_info_get(h, ctypes.byref(info))
            if st != _abi.KC_OK:
                raise KcError(st, "kc_unitigs_info_get failed")
            m, nb = int(info.n_unitigs), int(info.n_bases)
            seq = np.zeros(max(nb, 1), np.uint8)
            off = np.zeros(m + 1, np.uint64)
            km = np.zeros(max(m, 1), np.uint64)
            cov = np.zeros(max(m, 1), np.float64)
            st = lib().kc_unitigs_export(h, ptr(seq, u8p), nb, ptr(off, u64p), ptr(km, u64p),
                                         cov.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m)
            if st != _abi.KC_OK:
                raise KcError(st, lib().kc_unitigs_last_error(h).decode())
        finally:
            lib().kc_unitigs_destroy(h)
        return cls(engine.config.k, seq[:nb], off, km[:m], cov[:m], float(info.ms_device), list(info.ms_stage))

    def __len__(self):
        return int(self.kmers.shape[0])

    def sequence(self, i) -> bytes:
        return self.seq[int(self.offsets[i]):int(self.offsets[i + 1])].tobytes()

    @property
    def unitigs(self) -> List[Unitig]:
        return [Unitig(i, self.sequence(i), float(self.mean_cov[i]), int(self.kmers[i])) for i in range(len(self))]
