"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test infrastructure."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

SUMMARY_NAMES = [
    "divide_max_group_size", "divide_fallback_splits", "merge_candidate_pairs", "merge_selected_pairs",
    "merge_total_local_gain", "merge_group_count", "merge_group_max_size", "merge_raw_pair_count",
    "merge_candidate_pairs_after_prune", "merge_exact_gain_calls", "merge_positive_gain_pairs",
    "merge_rejected_by_overlap", "merge_rejected_by_threshold", "selected_merges_for_validation",
    "accepted_merges_after_validation", "rejected_nonpositive", "validation_exact_calls",
    "validation_exact_row_entry_work", "isolated_gain_sum",
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: make -C oracle")
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        i32, i64 = ctypes.c_int, ctypes.c_longlong
        sigs = {
            "orc_last_error": (ctypes.c_char_p, []),
            "orc_create": (P, [ctypes.c_char_p, i32, i32, i32, i32, ctypes.c_ulonglong]),
            "orc_free": (None, [P]),
            "orc_n": (i32, [P]),
            "orc_arcs": (i64, [P]),
            "orc_graph_csr": (i32, [P, P, P]),
            "orc_record_commits": (None, [P, i32]),
            "orc_run": (i32, [P, i32]),
            "orc_partition_hash": (ctypes.c_ulonglong, [P]),
            "orc_labels": (i32, [P, P]),
            "orc_exact_partition_cost": (i64, [P]),
            "orc_quotient_cost": (i64, [P]),
            "orc_summary": (i32, [P, P]),
            "orc_rounds": (i32, [P, P, i32]),
            "orc_divide": (i32, [P, i32, P, P, P]),
            "orc_active_count": (i32, [P]),
            "orc_prepare_group": (i32, [P, P, i32, P, P, i32, P]),
            "orc_gain": (i32, [P, i32, i32, P, P]),
            "orc_full_gain": (i32, [P, i32, i32, P]),
            "orc_merge": (i32, [P, i32, i32]),
            "orc_merge_bulk": (i32, [P, P, P, i32]),
            "orc_nnz_cross": (i64, [P]),
            "orc_export_quotient": (i32, [P, P, P, P, P, P, P]),
            "orc_encode_counts": (i32, [P, P]),
            "orc_commit_log": (i32, [P, P, P, P, i32, i32]),
            "orc_write_output": (i32, [P, ctypes.c_char_p]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _ok(rc):
    if rc < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return rc


class Oracle:
    def __init__(self, path: str, seed: int, top_k=16, group_batch=64, hash_dims=16, max_group=512):
        h = lib().orc_create(path.encode(), top_k, group_batch, hash_dims, max_group, seed)
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        self.h = h
        self.n = lib().orc_n(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def csr(self):
        arcs = lib().orc_arcs(self.h)
        row_ptr = np.zeros(self.n + 1, np.int64)
        col = np.zeros(max(arcs, 1), np.int32)
        lib().orc_graph_csr(self.h, _p(row_ptr), _p(col))
        return row_ptr, col[:arcs]

    def run(self, iterations: int):
        _ok(lib().orc_run(self.h, iterations))

    def record_commits(self, on=True):
        lib().orc_record_commits(self.h, 1 if on else 0)

    def partition_hash(self) -> int:
        return int(lib().orc_partition_hash(self.h))

    def labels(self):
        out = np.zeros(self.n, np.int32)
        _ok(lib().orc_labels(self.h, _p(out)))
        return out

    def summary(self) -> dict:
        out = np.zeros(len(SUMMARY_NAMES), np.int64)
        lib().orc_summary(self.h, _p(out))
        return dict(zip(SUMMARY_NAMES, (int(v) for v in out)))

    def rounds(self):
        out = np.zeros(10 * 4096, np.int64)
        k = lib().orc_rounds(self.h, _p(out), 4096)
        return out[: 10 * k].reshape(k, 10)

    def divide(self, it: int):
        order = np.zeros(self.n, np.int32)
        starts = np.zeros(self.n, np.int32)
        lengths = np.zeros(self.n, np.int32)
        g = _ok(lib().orc_divide(self.h, it, _p(order), _p(starts), _p(lengths)))
        act = lib().orc_active_count(self.h)
        return order[:act], starts[:g].copy(), lengths[:g].copy()

    def prepare_group(self, reps):
        reps = np.ascontiguousarray(reps, np.int32)
        cap = max(1, len(reps) * len(reps))
        a = np.zeros(cap, np.int32)
        b = np.zeros(cap, np.int32)
        ex = ctypes.c_ulonglong()
        k = _ok(lib().orc_prepare_group(self.h, _p(reps), len(reps), _p(a), _p(b), cap, ctypes.byref(ex)))
        return np.stack([a[:k], b[:k]], axis=1), int(ex.value)

    def gain(self, a: int, b: int):
        g = ctypes.c_longlong()
        bf = ctypes.c_longlong()
        _ok(lib().orc_gain(self.h, a, b, ctypes.byref(g), ctypes.byref(bf)))
        return g.value, bf.value

    def full_gain(self, a: int, b: int) -> int:
        g = ctypes.c_longlong()
        _ok(lib().orc_full_gain(self.h, a, b, ctypes.byref(g)))
        return g.value

    def merge(self, keep: int, remove: int):
        _ok(lib().orc_merge(self.h, keep, remove))

    def merge_bulk(self, keeps, removes) -> int:
        keeps = np.ascontiguousarray(keeps, np.int32)
        removes = np.ascontiguousarray(removes, np.int32)
        return _ok(lib().orc_merge_bulk(self.h, _p(keeps), _p(removes), len(keeps)))

    def quotient(self):
        nnz = lib().orc_nnz_cross(self.h)
        offsets = np.zeros(self.n + 1, np.int64)
        nbr = np.zeros(max(nnz, 1), np.int32)
        cnt = np.zeros(max(nnz, 1), np.int64)
        sizes = np.zeros(self.n, np.int64)
        internal = np.zeros(self.n, np.int64)
        active = np.zeros(self.n, np.uint8)
        lib().orc_export_quotient(self.h, _p(offsets), _p(nbr), _p(cnt), _p(sizes), _p(internal), _p(active))
        return offsets, nbr[:nnz], cnt[:nnz], sizes, internal, active

    def quotient_cost(self) -> int:
        return int(lib().orc_quotient_cost(self.h))

    def exact_partition_cost(self) -> int:
        return int(lib().orc_exact_partition_cost(self.h))

    def encode_counts(self) -> dict:
        out = np.zeros(5, np.int64)
        _ok(lib().orc_encode_counts(self.h, _p(out)))
        return dict(zip(["P", "Cp", "Cm", "P_loops", "supernode_count"], (int(v) for v in out)))

    def write_output(self, directory: str) -> None:
        _ok(lib().orc_write_output(self.h, directory.encode()))
