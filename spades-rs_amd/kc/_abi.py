"""ctypes binding of libkc.so (the C ABI declared in include/kc.h).

The library is the only compute path: if it is missing this module raises at
import time instead of falling back to anything else.
"""
import ctypes
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libkc.so")

KC_OK = 0
KC_EINVAL = -1
KC_ENOMEM = -2
KC_EHIP = -3
KC_ESTATE = -4
KC_EUNSUPPORTED = -5
KC_MAX_K = 127
KC_ANCHOR_POLISHER, KC_ANCHOR_PAIRED_INFO, KC_ANCHOR_SPLITTER = 0, 1, 2
KC_FLAG_TIMING = 1
KC_FLAG_EVENTS64 = 2
KC_NOALLOC_MAX_READS = 1 << 22

STATUS_NAMES = {KC_EINVAL: "KC_EINVAL", KC_ENOMEM: "KC_ENOMEM", KC_EHIP: "KC_EHIP",
                KC_ESTATE: "KC_ESTATE", KC_EUNSUPPORTED: "KC_EUNSUPPORTED"}

# Every entry point include/kc.h declares.
EXPORTS = (
    "kc_filter_create", "kc_filter_destroy", "kc_filter_last_error", "kc_filter_memory_usage_bytes",
    "kc_filter_num_words", "kc_filter_insert_kmer256", "kc_filter_is_solid_kmer256", "kc_filter_export",
    "kc_filter_correct_reads", "kc_filter_correct_reads_noalloc", "kc_hamming_distance_2bit",
    "kc_create", "kc_destroy", "kc_last_error", "kc_reads_upload", "kc_bloom_build", "kc_count", "kc_run", "kc_run_many",
    "kc_info_get", "kc_bloom_export", "kc_bloom_query", "kc_counts_export", "kc_reads_export", "kc_kernel_times",
    "kc_library_verify", "kc_library_add", "kc_library_count", "kc_library_info", "kc_library_bloom_export", "kc_library_bloom_query",
    "kc_library_counts_export", "kc_library_reads_export", "kc_library_priors_set", "kc_priors_set",
    "kc_tune_get", "kc_tune_set", "kc_autotune", "kc_graph_run", "kc_graph_release", "kc_set_k",
    "kc_mlp_create", "kc_mlp_destroy", "kc_mlp_last_error", "kc_mlp_num_params", "kc_mlp_get_params",
    "kc_mlp_set_params", "kc_mlp_forward", "kc_mlp_gradient", "kc_mlp_train",
    "kc_library_quals_set", "kc_library_features", "kc_library_learned_select", "kc_library_cutoff_fragile",
    "kc_cutoff_is_fragile",
    "kc_fastq_pack", "kc_fastq_last_error", "kc_packed_info_get", "kc_packed_export", "kc_packed_destroy",
    "kc_ingest_create", "kc_ingest_destroy", "kc_ingest_pack", "kc_ingest_keep_quals", "kc_packed_export_quals",
    "kc_anchor_build", "kc_anchor_count", "kc_anchor_export", "kc_anchor_query", "kc_anchor_last_error",
    "kc_anchor_destroy",
    "kc_unitigs_build", "kc_unitigs_info_get", "kc_unitigs_export", "kc_unitigs_last_error", "kc_unitigs_destroy",
)


class KcError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class kc_opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("k", ctypes.c_uint32), ("bloom_bits", ctypes.c_uint64),
                ("min_coverage", ctypes.c_double), ("error_correct", ctypes.c_int32),
                ("table_load", ctypes.c_float), ("flags", ctypes.c_uint32)]


class kc_verify(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("windows", "solid_windows", "sum_counts", "seen_words_differing",
                                                "unordered_pairs", "mask_mismatches", "checksum")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class kc_info(ctypes.Structure):
    _fields_ = [("n_reads", ctypes.c_uint64), ("n_occ", ctypes.c_uint64), ("n_distinct", ctypes.c_uint64),
                ("n_counts", ctypes.c_uint64), ("n_solid", ctypes.c_uint64), ("corrected_reads", ctypes.c_uint64),
                ("num_words", ctypes.c_uint64), ("table_capacity", ctypes.c_uint64), ("cutoff", ctypes.c_uint32),
                ("top_cov", ctypes.c_uint32), ("ms_stage2", ctypes.c_double), ("ms_stage3", ctypes.c_double),
                ("hammer_queries", ctypes.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class kc_packed_info(ctypes.Structure):
    _fields_ = [("n_reads", ctypes.c_uint64), ("data_len", ctypes.c_uint64), ("pe_count", ctypes.c_uint64),
                ("n_ambiguous", ctypes.c_uint64), ("r1_records", ctypes.c_uint64), ("r2_records", ctypes.c_uint64),
                ("ms_device", ctypes.c_double), ("n_quals", ctypes.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class kc_unitigs_info(ctypes.Structure):
    _fields_ = [("n_unitigs", ctypes.c_uint64), ("n_bases", ctypes.c_uint64), ("n_kmers", ctypes.c_uint64),
                ("ms_device", ctypes.c_double), ("ms_stage", ctypes.c_double * 4)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["ms_stage"] = list(self.ms_stage)
        return d


class kc_tune(ctypes.Structure):
    _fields_ = [("windows_per_thread", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("table_load", ctypes.c_float), ("insert_path", ctypes.c_uint32),
                ("setter_path", ctypes.c_uint32), ("kmers_per_bin", ctypes.c_uint32),
                ("minimizer_len", ctypes.c_uint32), ("ms", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_vp = ctypes.c_void_p
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_dblp = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_sz = ctypes.c_size_t

_SIGS = {
    "kc_filter_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int32, ctypes.c_uint64]),
    "kc_filter_destroy": (None, [_vp]),
    "kc_filter_last_error": (ctypes.c_char_p, [_vp]),
    "kc_filter_memory_usage_bytes": (ctypes.c_uint64, [_vp]),
    "kc_filter_num_words": (ctypes.c_uint64, [_vp]),
    "kc_filter_insert_kmer256": (ctypes.c_int, [_vp, _u64p, _sz, _u8p]),
    "kc_filter_is_solid_kmer256": (ctypes.c_int, [_vp, _u64p, _sz, _u8p]),
    "kc_filter_export": (ctypes.c_int, [_vp, _u64p, _u64p, _sz]),
    "kc_filter_correct_reads": (ctypes.c_int, [_vp, ctypes.c_uint32, _u8p, _sz, _u64p, _u32p, _sz, _u8p, _u64p]),
    "kc_filter_correct_reads_noalloc": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp, _sz, _vp, _vp, _sz, _vp, _u64p]),
    "kc_hamming_distance_2bit": (ctypes.c_int, [ctypes.c_int32, _u64p, _u64p, _sz, _u32p]),
    "kc_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(kc_opts)]),
    "kc_destroy": (None, [_vp]),
    "kc_last_error": (ctypes.c_char_p, [_vp]),
    "kc_reads_upload": (ctypes.c_int, [_vp, _u8p, _sz, _u64p, _u32p, _sz]),
    "kc_bloom_build": (ctypes.c_int, [_vp]),
    "kc_count": (ctypes.c_int, [_vp, ctypes.POINTER(kc_info)]),
    "kc_run": (ctypes.c_int, [_vp, ctypes.POINTER(kc_info)]),
    "kc_run_many": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.POINTER(kc_info)]),
    "kc_info_get": (ctypes.c_int, [_vp, ctypes.POINTER(kc_info)]),
    "kc_bloom_export": (ctypes.c_int, [_vp, _u64p, _u64p, _sz]),
    "kc_bloom_query": (ctypes.c_int, [_vp, _u64p, _sz, _u8p]),
    "kc_counts_export": (ctypes.c_int, [_vp, _u64p, _u32p, _u8p, _sz]),
    "kc_reads_export": (ctypes.c_int, [_vp, _u8p, _sz]),
    "kc_kernel_times": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_char_p), _dblp, _u32p, ctypes.c_int]),
    "kc_library_add": (ctypes.c_int, [_vp, _u8p, _sz, _u64p, _u32p, _sz, ctypes.c_uint64, ctypes.c_double, _u32p]),
    "kc_library_count": (ctypes.c_int, [_vp, _u32p]),
    "kc_library_info": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(kc_info)]),
    "kc_library_verify": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(kc_verify)]),
    "kc_library_cutoff_fragile": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                                 ctypes.POINTER(ctypes.c_int32)]),
    "kc_cutoff_is_fragile": (ctypes.c_int, [_u64p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint32,
                                            ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32)]),
    "kc_library_bloom_export": (ctypes.c_int, [_vp, ctypes.c_uint32, _u64p, _u64p, _sz]),
    "kc_library_bloom_query": (ctypes.c_int, [_vp, ctypes.c_uint32, _u64p, _sz, _u8p]),
    "kc_library_counts_export": (ctypes.c_int, [_vp, ctypes.c_uint32, _u64p, _u32p, _u8p, _sz]),
    "kc_library_reads_export": (ctypes.c_int, [_vp, ctypes.c_uint32, _u8p, _sz]),
    "kc_library_priors_set": (ctypes.c_int, [_vp, ctypes.c_uint32, _u8p, _sz, _u64p, _u64p, _sz]),
    "kc_priors_set": (ctypes.c_int, [_vp, _u8p, _sz, _u64p, _u64p, _sz]),
    "kc_tune_get": (ctypes.c_int, [_vp, ctypes.POINTER(kc_tune)]),
    "kc_tune_set": (ctypes.c_int, [_vp, ctypes.POINTER(kc_tune)]),
    "kc_autotune": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(kc_tune)]),
    "kc_graph_run": (ctypes.c_int, [_vp, ctypes.POINTER(kc_info)]),
    "kc_graph_release": (ctypes.c_int, [_vp]),
    "kc_set_k": (ctypes.c_int, [_vp, ctypes.c_uint32]),
    "kc_mlp_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int32, ctypes.c_uint64]),
    "kc_mlp_destroy": (None, [_vp]),
    "kc_mlp_last_error": (ctypes.c_char_p, [_vp]),
    "kc_mlp_num_params": (ctypes.c_uint32, []),
    "kc_mlp_get_params": (ctypes.c_int, [_vp, _f32p, _sz]),
    "kc_mlp_set_params": (ctypes.c_int, [_vp, _f32p, _sz]),
    "kc_mlp_forward": (ctypes.c_int, [_vp, _f32p, _sz, _f32p]),
    "kc_mlp_gradient": (ctypes.c_int, [_vp, _f32p, _u8p, _f32p, _sz, ctypes.c_float, _f32p, _dblp]),
    "kc_mlp_train": (ctypes.c_int, [_vp, _f32p, _u8p, _f32p, _sz, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_uint64, _dblp]),
    "kc_library_quals_set": (ctypes.c_int, [_vp, ctypes.c_uint32, _u8p, _sz, _u64p]),
    "kc_library_features": (ctypes.c_int, [_vp, ctypes.c_uint32, _f32p, _sz]),
    "kc_library_learned_select": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp, ctypes.c_float, _u8p, _u8p, _sz, _dblp]),
    "kc_fastq_pack": (ctypes.c_int, [ctypes.c_int32, _u8p, _sz, _u8p, _sz, ctypes.POINTER(_vp)]),
    "kc_fastq_last_error": (ctypes.c_char_p, []),
    "kc_packed_info_get": (ctypes.c_int, [_vp, ctypes.POINTER(kc_packed_info)]),
    "kc_packed_export": (ctypes.c_int, [_vp, _u8p, _sz, _u64p, _u32p, _sz]),
    "kc_packed_destroy": (None, [_vp]),
    "kc_ingest_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(_vp)]),
    "kc_ingest_destroy": (None, [_vp]),
    "kc_ingest_pack": (ctypes.c_int, [_vp, _u8p, _sz, _u8p, _sz, ctypes.POINTER(_vp)]),
    "kc_ingest_keep_quals": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "kc_anchor_build": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint32, ctypes.c_int32, _u8p, _sz, _u64p, _u64p, _sz,
                                       ctypes.POINTER(_vp)]),
    "kc_anchor_count": (ctypes.c_int, [_vp, _u64p]),
    "kc_anchor_export": (ctypes.c_int, [_vp, _u64p, _u32p, _u32p, _u8p, _sz]),
    "kc_anchor_query": (ctypes.c_int, [_vp, _u64p, _sz, ctypes.POINTER(ctypes.c_int64)]),
    "kc_anchor_last_error": (ctypes.c_char_p, []),
    "kc_anchor_destroy": (None, [_vp]),
    "kc_unitigs_build": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "kc_unitigs_info_get": (ctypes.c_int, [_vp, ctypes.POINTER(kc_unitigs_info)]),
    "kc_unitigs_export": (ctypes.c_int, [_vp, _u8p, _sz, _u64p, _u64p, ctypes.POINTER(ctypes.c_double), _sz]),
    "kc_unitigs_last_error": (ctypes.c_char_p, [_vp]),
    "kc_unitigs_destroy": (None, [_vp]),
    "kc_packed_export_quals": (ctypes.c_int, [_vp, _u8p, _sz, _u64p, _sz]),
}

_lib = None


def lib():
    """Load libkc.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C spades-rs_amd` "
                              f"(or __graft_entry__.build()); there is no fallback path")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a, t):
    """Pointer to a numpy array's data (None for an empty array)."""
    if a is None:
        return None
    return a.ctypes.data_as(t)


u8p, u32p, u64p, f32p = _u8p, _u32p, _u64p, _f32p
