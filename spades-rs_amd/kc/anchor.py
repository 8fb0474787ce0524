"""Unique-anchor k-mer indexes over contigs / unitigs on the GPU (libkc.so
kc_anchor_*): the maps of Polisher::build_kmer_index (polisher.rs:33-60),
PairedInfoIndex::build_kmer_index (paired_info.rs:34-70) and the splitter's anchor step
(splitter.rs:57-70), as a sorted table of canonical k-mers with their first occurrence."""
import ctypes

import numpy as np

from . import _abi
from ._abi import KcError, lib, ptr, u8p, u32p, u64p

KINDS = {"polisher": _abi.KC_ANCHOR_POLISHER, "paired_info": _abi.KC_ANCHOR_PAIRED_INFO,
         "splitter": _abi.KC_ANCHOR_SPLITTER}


def _concat(seqs):
    bs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    lens = np.array([len(b) for b in bs], dtype=np.uint64)
    starts = np.zeros(len(bs), dtype=np.uint64)
    if len(bs) > 1:
        starts[1:] = np.cumsum(lens)[:-1]
    return np.frombuffer(b"".join(bs) + b"\0", dtype=np.uint8), starts, lens


class AnchorIndex:
    """AnchorIndex(kind, k, seqs): kind is "polisher", "paired_info" or "splitter"; seqs
    are ASCII sequences (str or bytes).  export() -> (keys (n, 4) Kmer256 limbs sorted,
    sequence index, position, is_rc); query(keys) -> anchor index per key, or -1."""

    def __init__(self, kind, k, seqs, device=0):
        data, starts, lens = _concat(seqs)
        self._h = ctypes.c_void_p()
        self.k = k
        st = lib().kc_anchor_build(device, k, KINDS[kind], ptr(data, u8p), int(lens.sum()) if len(lens) else 0,
                                   ptr(starts, u64p), ptr(lens, u64p), len(seqs), ctypes.byref(self._h))
        if st != _abi.KC_OK:
            raise KcError(st, lib().kc_anchor_last_error().decode())
        n = ctypes.c_uint64(0)
        lib().kc_anchor_count(self._h, ctypes.byref(n))
        self.n = n.value

    def export(self):
        n = self.n
        keys = np.zeros((max(n, 1), 4), dtype=np.uint64)
        seq = np.zeros(max(n, 1), dtype=np.uint32)
        pos = np.zeros(max(n, 1), dtype=np.uint32)
        rc = np.zeros(max(n, 1), dtype=np.uint8)
        st = lib().kc_anchor_export(self._h, ptr(keys, u64p), ptr(seq, u32p), ptr(pos, u32p), ptr(rc, u8p), n)
        if st != _abi.KC_OK:
            raise KcError(st, lib().kc_anchor_last_error().decode())
        return keys[:n], seq[:n], pos[:n], rc[:n]

    def query(self, keys4):
        q = np.ascontiguousarray(keys4, dtype=np.uint64).reshape(-1, 4)
        out = np.zeros(max(len(q), 1), dtype=np.int64)
        st = lib().kc_anchor_query(self._h, ptr(q, u64p), len(q), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        if st != _abi.KC_OK:
            raise KcError(st, lib().kc_anchor_last_error().decode())
        return out[:len(q)]

    def close(self):
        if self._h:
            lib().kc_anchor_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
