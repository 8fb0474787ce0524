"""FASTQ ingestion on the GPU (packed_reads.rs from_paired_fastq / append_from_file):
file contents -> the three-block PackedReads layout, through libkc.so's kc_ingest_pack."""
import ctypes
import gzip

import numpy as np

from . import _abi
from ._abi import KcError, lib, ptr, u8p, u32p, u64p
from .packed_reads import PackedReads


def _contents(x):
    if isinstance(x, (bytes, bytearray, memoryview)):
        return bytes(x)
    path = str(x)
    opener = gzip.open if path.lower().endswith(".gz") else open
    with opener(path, "rb") as fh:
        return fh.read()


def _u8(b):
    """A uint8 view of b with one byte after it, so that an empty input has an address."""
    if isinstance(b, np.ndarray):
        return b
    return np.frombuffer(b + b"\0", dtype=np.uint8)


class FastqIngest:
    """A device context for repeated ingestion (its buffers grow to the largest input and
    are reused).  pack(r1, r2=None) takes bytes, uint8 arrays or paths (.gz decompressed
    on the host) and returns (PackedReads, info dict).  With keep_quals, info also holds
    "quals" (the Phred value of every packed base) and "qoff" (read i's at quals[qoff[i]:]),
    the arrays SpectrumEngine.set_quals takes.  FASTQ (with blank lines between records) and
    FASTA (single- or multi-line) are parsed on the device.  Raises KcError:
    KC_EUNSUPPORTED for keep_quals with FASTA input, KC_EINVAL for a malformed record."""

    def __init__(self, device=0, keep_quals=False):
        self._h = ctypes.c_void_p()
        st = lib().kc_ingest_create(device, ctypes.byref(self._h))
        if st != _abi.KC_OK:
            raise KcError(st, lib().kc_fastq_last_error().decode())
        if keep_quals:
            lib().kc_ingest_keep_quals(self._h, 1)
        self.keep_quals = bool(keep_quals)

    def pack(self, r1, r2=None):
        a = r1 if isinstance(r1, np.ndarray) else _contents(r1)
        b = None if r2 is None else (r2 if isinstance(r2, np.ndarray) else _contents(r2))
        aa, bb = _u8(a), None if b is None else _u8(b)
        h = ctypes.c_void_p()
        st = lib().kc_ingest_pack(self._h, ptr(aa, u8p), len(a), None if bb is None else ptr(bb, u8p),
                                  0 if b is None else len(b), ctypes.byref(h))
        if st != _abi.KC_OK:
            raise KcError(st, lib().kc_fastq_last_error().decode())
        return _take(h, self.keep_quals)

    def close(self):
        if self._h:
            lib().kc_ingest_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _take(h, quals=False):
    try:
        info = _abi.kc_packed_info()
        lib().kc_packed_info_get(h, ctypes.byref(info))
        n, dl = info.n_reads, info.data_len
        data = np.zeros(max(dl, 1), dtype=np.uint8)
        offs = np.zeros(max(n, 1), dtype=np.uint64)
        lens = np.zeros(max(n, 1), dtype=np.uint32)
        st = lib().kc_packed_export(h, ptr(data, u8p), dl, ptr(offs, u64p), ptr(lens, u32p), n)
        if st != _abi.KC_OK:
            raise KcError(st, "kc_packed_export failed")
        store = PackedReads.from_arrays(data[:dl], offs[:n], lens[:n], pe_count=info.pe_count)
        out = info.as_dict()
        if quals:
            nq = info.n_quals
            q = np.zeros(max(nq, 1), dtype=np.uint8)
            qo = np.zeros(max(n, 1), dtype=np.uint64)
            st = lib().kc_packed_export_quals(h, ptr(q, u8p), nq, ptr(qo, u64p), n)
            if st != _abi.KC_OK:
                raise KcError(st, "kc_packed_export_quals failed")
            out["quals"], out["qoff"] = q[:nq], qo[:n]
        return store, out
    finally:
        lib().kc_packed_destroy(h)


def fastq_pack(r1, r2=None, device=0):
    """One-shot form of FastqIngest(device).pack(r1, r2) (kc_fastq_pack)."""
    a = _contents(r1)
    b = None if r2 is None else _contents(r2)
    aa, bb = _u8(a), None if b is None else _u8(b)
    h = ctypes.c_void_p()
    st = lib().kc_fastq_pack(device, ptr(aa, u8p), len(a), None if bb is None else ptr(bb, u8p),
                             0 if b is None else len(b), ctypes.byref(h))
    if st != _abi.KC_OK:
        raise KcError(st, lib().kc_fastq_last_error().decode())
    return _take(h)
