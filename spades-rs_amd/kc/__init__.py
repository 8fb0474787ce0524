"""kc — host side of the MI355X k-mer spectrum engine (spades-rs Stage 2-3).

The compute path is libkc.so (HIP, gfx950) behind the C ABI in include/kc.h;
this package mirrors the reference's Rust interface for that path on top of it.
"""
from ._abi import KC_MAX_K, KcError, LIB_PATH, lib  # noqa: F401
from .engine import (AssemblerConfig, ErrorCorrector, KmerSpectrum, SpectrumEngine,  # noqa: F401
                     TwoTierFilter, count_kmers, hamming_distance_2bit, run_many)
from .graph import CompactedGraph, Unitig  # noqa: F401
from .packed_reads import MIN_K, PackedReads, pack_codes, pack_fixed_length  # noqa: F401
