"""MI355X-native tree (Merkle-forest, lowest-offset) incremental-checkpoint deduplication.

The compute path is libmtcheck_amd.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/mtcheck.h); this package is its host-side mirror of MTCheck's
TreeDeduplicator interface plus the synthetic workload generator.
"""
from .tree import (BASIC, DONE, FIRST_OCUR, FIXED_DUPL, FULL, LIST, MD5, MURMUR3, SHIFT_DUPL, TREE,
                   TREE_LOW_OFFSET, TREE_LOW_OFFSET_REF, TREE_LOW_ROOT, TREE_LOW_ROOT_REF, UNKNOWN,
                   Batch, TreeDeduplicator, debug_labels, debug_tree, digest_host, digest_str, get_mode, hash_chunks,
                   parse_header, restart_into)
from ._lib import MtcError

__all__ = [
    "TreeDeduplicator", "Batch", "restart_into", "digest_host", "digest_str", "hash_chunks", "parse_header", "get_mode",
    "debug_tree", "debug_labels", "MtcError", "MURMUR3", "MD5", "UNKNOWN", "FULL", "BASIC", "LIST", "TREE",
    "TREE_LOW_OFFSET_REF", "TREE_LOW_OFFSET", "TREE_LOW_ROOT_REF", "TREE_LOW_ROOT", "FIRST_OCUR", "FIXED_DUPL",
    "SHIFT_DUPL", "DONE",
]
