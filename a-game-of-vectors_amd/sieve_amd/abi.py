"""ctypes binding of libsieve_gpu.so (include/sieve_gpu.h).

Loading fails loudly when the library is missing: there is no CPU fallback anywhere in the
package. Array arguments may be numpy arrays (host memory; the library copies them) or torch
tensors on the GPU (device memory, used in place).
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsieve_gpu.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "sieve_gpu.h")

PAD = 0xFFFFFFFF
RERANK_TAGS = {"none": 0, "prefix": 1, "full": 2, "tiered": 3, "ssd": 4}
STATUS = {0: "ok", -1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EIO", -5: "ELIMIT"}


class SieveError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sieve {STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("cand", ctypes.c_uint32), ("rerank", ctypes.c_uint32), ("rerank_r", ctypes.c_uint32),
                ("k", ctypes.c_uint32), ("prune_m", ctypes.c_uint32), ("prune_w", ctypes.c_float),
                ("nprobe", ctypes.c_uint32), ("margin", ctypes.c_float), ("budget", ctypes.c_uint64)]


class Timing(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("route", "scan", "rerank", "tier2", "scanned", "reranked", "probed", "fetch", "wait")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class IndexDesc(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("n", ctypes.c_uint64), ("nblocks", ctypes.c_uint64),
                ("centroids", ctypes.c_void_p), ("block_off", ctypes.c_void_p), ("blocks", ctypes.c_void_p),
                ("ids", ctypes.c_void_p), ("factors", ctypes.c_void_p), ("rotation", ctypes.c_void_p),
                ("rerank", ctypes.c_void_p), ("quant", ctypes.c_void_p), ("norm256", ctypes.c_void_p),
                ("tier2_sigma", ctypes.c_void_p), ("tier2_codes", ctypes.c_void_p),
                ("tier2_norm", ctypes.c_void_p), ("borrow_rerank", ctypes.c_uint32)]


class View(ctypes.Structure):
    _fields_ = [("deleted", ctypes.c_void_p), ("deleted_words", ctypes.c_uint64), ("allow", ctypes.c_void_p),
                ("allow_words", ctypes.c_uint64), ("allow_stride", ctypes.c_uint64),
                ("allow_flags", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [("queries", ctypes.c_uint64), ("probed", ctypes.c_uint64), ("scanned", ctypes.c_uint64),
                ("id_reads", ctypes.c_uint64), ("reranked", ctypes.c_uint64), ("route_ns", ctypes.c_double),
                ("scan_ns", ctypes.c_double), ("rerank_ns", ctypes.c_double), ("rescanned", ctypes.c_uint64),
                ("graph", ctypes.c_uint64), ("cm_pairs", ctypes.c_uint64), ("cm_ns", ctypes.c_double),
                ("cm_launches", ctypes.c_uint64), ("cm_staged", ctypes.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class IndexInfo(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("n", ctypes.c_uint64), ("nblocks", ctypes.c_uint64),
                ("rabitq", ctypes.c_uint32), ("tier2", ctypes.c_uint32), ("max_cluster_blocks", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64)]


class L2Params(ctypes.Structure):
    _fields_ = [("nprobe", ctypes.c_uint32), ("cand", ctypes.c_uint32), ("k", ctypes.c_uint32),
                ("refine", ctypes.c_uint32), ("margin", ctypes.c_float)]


class L2Desc(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("n", ctypes.c_uint64), ("nblocks", ctypes.c_uint64),
                ("leaves", ctypes.c_void_p), ("block_off", ctypes.c_void_p), ("blocks", ctypes.c_void_p),
                ("ids", ctypes.c_void_p), ("factors", ctypes.c_void_p), ("rotation", ctypes.c_void_p),
                ("base", ctypes.c_void_p), ("borrow_base", ctypes.c_uint32), ("refine_bits", ctypes.c_uint32),
                ("refine", ctypes.c_void_p)]


class LiveOpts(ctypes.Structure):
    _fields_ = [("max_rows", ctypes.c_uint64), ("soar_lambda", ctypes.c_float)]


class LiveInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("rows", "slots", "arena_blocks", "deleted", "max_rows",
                                               "device_bytes")]


class ReplayStats(ctypes.Structure):
    _fields_ = [("inserts", ctypes.c_uint64), ("deletes", ctypes.c_uint64), ("torn", ctypes.c_uint64),
                ("seconds", ctypes.c_double)]


class CompactStats(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint64), ("dropped", ctypes.c_uint64), ("blocks", ctypes.c_uint64),
                ("seconds", ctypes.c_double)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_uint32), ("max_inflight", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64),
                ("live", ctypes.c_void_p)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("queue_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("batch_size", ctypes.c_uint32),
                ("batch_id", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint64), ("queries", ctypes.c_uint64), ("pending", ctypes.c_uint64),
                ("inflight", ctypes.c_uint64), ("slots", ctypes.c_uint32), ("slot_bytes", ctypes.c_uint64),
                ("writes", ctypes.c_uint64)]


_lib = None


# filter::Plan tags (SIEVE_PLAN_*).
PLANS = {"post": 0, "expand": 1, "brute": 2, "auto": 3}
# Scan kernel variants the autotuner times (SIEVE_SCAN_VARIANTS).
SCAN_VARIANTS = 12
# The cluster-major scan variants (sieve_index_set_cluster_scan).
SCAN_VARIANTS_CM = (10, 11)

def header_symbols():
    """Every function the public header declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(sieve_\w+)\s*\(", text, re.M)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libsieve_gpu.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    # torch ships its own HIP runtime with the same soname (libamdhip64.so.7). Loading torch
    # first makes the dynamic loader bind this library to that runtime, so a process that uses
    # both has one runtime; loaded the other way round, torch would start a second runtime
    # that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32
    P, T, V = ctypes.POINTER(Params), ctypes.POINTER(Timing), ctypes.POINTER(View)
    sig = {
        "sieve_last_error": (ctypes.c_char_p, []),
        "sieve_stream_create": (i32, [ctypes.POINTER(vp)]),
        "sieve_stream_destroy": (None, [vp]),
        "sieve_stream_sync": (i32, [vp]),
        "sieve_stream_set_graphs": (i32, [vp, ctypes.c_int]),
        "sieve_stream_handle": (vp, [vp]),
        "sieve_stream_last_stats": (i32, [vp, ctypes.POINTER(Stats)]),
        "sieve_gemm_abt": (i32, [vp, vp, sz, sz, sz, vp]),
        "sieve_lut_scan": (i32, [vp, u32, vp, sz, vp]),
        "sieve_index_open": (i32, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]),
        "sieve_index_create": (i32, [ctypes.POINTER(IndexDesc), ctypes.POINTER(vp)]),
        "sieve_index_free": (None, [vp]),
        "sieve_index_get_info": (i32, [vp, ctypes.POINTER(IndexInfo)]),
        "sieve_route": (i32, [vp, vp, vp, sz, P, vp, vp, vp, T]),
        "sieve_candidates": (i32, [vp, vp, vp, sz, vp, vp, vp, sz, P, V, vp, vp, vp, vp, T]),
        "sieve_rerank": (i32, [vp, vp, vp, vp, sz, vp, vp, vp, vp, sz, P, V, vp, vp, vp, T]),
        "sieve_search_batch": (i32, [vp, vp, vp, vp, sz, P, V, vp, vp, vp, T]),
        "sieve_search_batch_noalloc": (i32, [vp, vp, vp, vp, sz, P, V, vp, vp, vp]),
        "sieve_index_autotune": (i32, [vp, vp, vp, vp, sz, P, ctypes.POINTER(ctypes.c_uint32), vp]),
        "sieve_index_set_scan_variant": (i32, [vp, u32]),
        "sieve_index_set_scan_chunk": (i32, [vp, u32]),
        "sieve_index_set_cluster_scan": (i32, [vp, u32, u32, u32, u32, u32]),
        "sieve_index_set_small_batch": (i32, [vp, u32]),
        "sieve_l2_index_set_scan_chunk": (i32, [vp, u32]),
        "sieve_l2_index_set_scan_variant": (i32, [vp, u32]),
        "sieve_l2_index_set_small_batch": (i32, [vp, u32]),
        "sieve_index_set_route_fuse": (i32, [vp, i32]),
        "sieve_l2_index_set_route_fuse": (i32, [vp, i32]),
        "sieve_server_create": (i32, [vp, P, ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)]),
        "sieve_server_submit": (i32, [vp, vp, vp, ctypes.POINTER(ctypes.c_uint64)]),
        "sieve_server_submit_filtered": (i32, [vp, vp, vp, vp, u32, ctypes.POINTER(ctypes.c_uint64)]),
        "sieve_server_insert": (i32, [vp, vp, vp, sz, ctypes.POINTER(ctypes.c_uint64)]),
        "sieve_server_delete": (i32, [vp, vp, sz, ctypes.POINTER(ctypes.c_uint64)]),
        "sieve_server_wait": (i32, [vp, ctypes.c_uint64, vp, vp, ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.POINTER(RequestInfo)]),
        "sieve_server_get_stats": (i32, [vp, ctypes.POINTER(ServerStats)]),
        "sieve_server_destroy": (None, [vp]),
        "sieve_l2_index_create": (i32, [ctypes.POINTER(L2Desc), ctypes.POINTER(vp)]),
        "sieve_l2_index_free": (None, [vp]),
        "sieve_l2_refine_encode": (i32, [vp, sz, u32, vp, vp]),
        "sieve_l2_index_open": (i32, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]),
        "sieve_l2_hot_leaves": (i32, [vp, vp, vp, sz, u32, ctypes.c_float, ctypes.c_uint64, vp,
                                      ctypes.POINTER(u32)]),
        "sieve_l2_search_batch": (i32, [vp, vp, vp, sz, ctypes.POINTER(L2Params), vp, vp, vp, vp, vp, vp, vp, vp,
                                        T]),
        "sieve_filter_create": (i32, [vp, vp, ctypes.c_uint64, ctypes.POINTER(vp)]),
        "sieve_filter_free": (None, [vp]),
        "sieve_filter_get_info": (i32, [vp, ctypes.POINTER(ctypes.c_uint64), vp]),
        "sieve_plan_resolve": (u32, [u32, ctypes.c_uint64, ctypes.c_uint64]),
        "sieve_filter_set_gather": (i32, [vp, i32]),
        "sieve_filter_search_batch": (i32, [vp, vp, u32, ctypes.c_uint64, vp, vp, vp, sz, P, vp, vp, vp,
                                            ctypes.POINTER(u32), T]),
        "sieve_live_open": (i32, [vp, ctypes.POINTER(LiveOpts), ctypes.POINTER(vp)]),
        "sieve_live_open_dir": (i32, [vp, ctypes.POINTER(LiveOpts), ctypes.c_char_p, ctypes.c_char_p,
                                      ctypes.POINTER(ReplayStats), ctypes.POINTER(vp)]),
        "sieve_live_compact": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(CompactStats)]),
        "sieve_live_truncate_wal": (i32, [vp]),
        "sieve_live_close": (None, [vp]),
        "sieve_live_insert": (i32, [vp, vp, vp, sz]),
        "sieve_live_delete": (i32, [vp, vp, sz]),
        "sieve_live_current": (u32, [vp, u32]),
        "sieve_live_get_info": (i32, [vp, ctypes.POINTER(LiveInfo)]),
        "sieve_live_export": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "sieve_live_search_batch": (i32, [vp, vp, vp, vp, sz, P, V, vp, vp, vp, T]),
        "sieve_build_kmeans": (i32, [vp, sz, u32, u32, ctypes.c_uint64, vp]),
        "sieve_build_assign": (i32, [vp, sz, vp, u32, ctypes.c_float, vp]),
        "sieve_build_tier2_fit": (i32, [vp, sz, vp]),
        "sieve_build_l2_train": (i32, [vp, sz, u32, u32, sz, u32, vp, vp]),
        "sieve_build_l2_assign": (i32, [vp, sz, vp, u32, vp, u32, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise SieveError(rc, lib().sieve_last_error().decode())


def make_params(p):
    """A sieve_params from a sieve_amd.Params."""
    tag, r = p.rerank
    prune = p.prune if p.prune is not None else (0, 0.0)
    return Params(cand=p.cand, rerank=RERANK_TAGS[tag], rerank_r=r, k=p.k, prune_m=prune[0], prune_w=prune[1],
                  nprobe=p.nprobe, margin=float("nan") if p.margin is None else p.margin, budget=p.budget or 0)


def ptr(a):
    """Address of a numpy array or a torch tensor (contiguous)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
        return ctypes.c_void_p(a.ctypes.data)
    assert a.is_contiguous(), "tensor must be contiguous"
    return ctypes.c_void_p(a.data_ptr())
