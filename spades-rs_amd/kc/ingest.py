"""FASTQ ingestion on the GPU (packed_reads.rs from_paired_fastq / append_from_file):
file contents -> the three-block PackedReads layout, through libkc.so's kc_fastq_pack."""
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


def fastq_pack(r1, r2=None, device=0):
    """Pack FASTQ content (bytes or a path, .gz decompressed on the host) on the device.
    Returns (PackedReads, info dict).  Raises KcError: KC_EUNSUPPORTED for FASTA input
    or blank lines between records (use PackedReads.from_paired_fastq), KC_EINVAL for a
    malformed record."""
    a = _contents(r1)
    b = None if r2 is None else _contents(r2)
    aa = np.frombuffer(a + b"\0", dtype=np.uint8)
    bb = None if b is None else np.frombuffer(b + b"\0", dtype=np.uint8)
    h = ctypes.c_void_p()
    st = lib().kc_fastq_pack(device, ptr(aa, u8p), len(a), ptr(bb, u8p) if bb is not None else None,
                             0 if b is None else len(b), ctypes.byref(h))
    if st != _abi.KC_OK:
        raise KcError(st, lib().kc_fastq_last_error().decode())
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
        return store, info.as_dict()
    finally:
        lib().kc_packed_destroy(h)
