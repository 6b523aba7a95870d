"""The device index and its query path, mirroring the reference's Index API (src/ivf.rs:552-815):
open / route / candidates / rerank / search, here over batches of queries.

Inputs may be numpy arrays (results come back as numpy) or torch tensors on the GPU (results
are torch tensors on the same device, and nothing crosses PCIe).
"""
import ctypes

import numpy as np

from . import abi
from .params import Params

_HOST_DTYPES = dict(centroids=np.float32, block_off=np.uint64, blocks=np.uint8, ids=np.uint32, factors=np.float32,
                    rotation=np.float32, rerank=np.int8, quant=np.float32, norm256=np.float32,
                    tier2_sigma=np.float32, tier2_codes=np.uint8, tier2_norm=np.float32)


def _is_torch(x):
    return hasattr(x, "data_ptr") and hasattr(x, "is_cuda")


class Stream:
    """A sieve_stream: one HIP stream plus scratch, for one host thread."""

    def __init__(self):
        h = ctypes.c_void_p()
        abi.check(abi.lib().sieve_stream_create(ctypes.byref(h)))
        self._h = h

    def sync(self):
        abi.check(abi.lib().sieve_stream_sync(self._h))

    def set_graphs(self, enable=True):
        """Run batched searches on this stream as captured HIP graphs (sieve_stream_set_graphs)."""
        abi.check(abi.lib().sieve_stream_set_graphs(self._h, int(bool(enable))))

    def last_stats(self):
        """Counters of the last call on this stream (sieve_stats) as a dict."""
        st = abi.Stats()
        abi.check(abi.lib().sieve_stream_last_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    @property
    def handle(self):
        return abi.lib().sieve_stream_handle(self._h)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                abi.lib().sieve_stream_destroy(self._h)
        except Exception:
            pass
        self._h = None


class Index:
    """A RaBitQ/SOAR IVF index resident in HBM (sieve_index)."""

    def __init__(self, handle, keep=None):
        self._h = handle
        self._keep = keep  # arrays the index borrows
        info = abi.IndexInfo()
        abi.check(abi.lib().sieve_index_get_info(self._h, ctypes.byref(info)))
        self.k, self.n, self.nblocks = info.k, info.n, info.nblocks
        self.tier2 = bool(info.tier2)
        self.device_bytes = info.device_bytes
        self.max_cluster_blocks = info.max_cluster_blocks
        self._stream = None

    @classmethod
    def open(cls, index_dir, ingest_dir):
        """Index::open + Corpus::open over the reference's files."""
        h = ctypes.c_void_p()
        abi.check(abi.lib().sieve_index_open(index_dir.encode(), ingest_dir.encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def create(cls, arrays, borrow_rerank=False):
        """From a dict of arrays in the on-disk layout (numpy, or torch tensors on the GPU)."""
        keep = {}
        d = abi.IndexDesc()
        for name, dt in _HOST_DTYPES.items():
            a = arrays.get(name)
            if a is None:
                continue
            if not _is_torch(a):
                a = np.ascontiguousarray(a, dtype=dt)
            else:
                a = a.contiguous()
            keep[name] = a
            setattr(d, name, abi.ptr(a).value)
        d.k = _numel(keep["centroids"]) // 256
        d.n = _numel(keep["rerank"]) // 768
        d.nblocks = _numel(keep["ids"]) // 32
        d.borrow_rerank = 1 if borrow_rerank else 0
        h = ctypes.c_void_p()
        abi.check(abi.lib().sieve_index_create(ctypes.byref(d), ctypes.byref(h)))
        return cls(h, keep if borrow_rerank else None)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                abi.lib().sieve_index_free(self._h)
        except Exception:
            pass
        self._h = None

    def stream(self):
        if self._stream is None:
            self._stream = Stream()
        return self._stream

    # ------------------------------------------------------------------ helpers -----
    @staticmethod
    def _out(like, shape, dtype):
        if _is_torch(like):
            import torch
            tdt = {np.float32: torch.float32, np.uint32: torch.int32, np.int8: torch.int8}[dtype]
            return torch.empty(shape, dtype=tdt, device=like.device)
        return np.empty(shape, dtype)

    @staticmethod
    def _in(x, dtype):
        if x is None or _is_torch(x):
            return x
        return np.ascontiguousarray(x, dtype=dtype)

    # ------------------------------------------------------------------ path --------
    def route(self, q256, p: Params, timing=None, stream=None):
        """Index::route for a batch: (scores nq x nprobe, clusters, counts)."""
        q256 = self._in(q256, np.float32)
        nq = q256.shape[0]
        s = self._out(q256, (nq, p.nprobe), np.float32)
        c = self._out(q256, (nq, p.nprobe), np.uint32)
        n = self._out(q256, (nq,), np.uint32)
        t = timing if timing is not None else abi.Timing()
        st = stream or self.stream()
        abi.check(abi.lib().sieve_route(self._h, st._h, abi.ptr(q256), nq, ctypes.byref(abi.make_params(p)),
                                        abi.ptr(s), abi.ptr(c), abi.ptr(n), ctypes.byref(t)))
        return s, c, n

    def candidates(self, q256, probe_s, probe_c, probe_n, p: Params, view=None, timing=None, stream=None):
        """Index::candidates for a batch: (est, rows, errs) nq x cand and counts."""
        q256 = self._in(q256, np.float32)
        probe_s, probe_c, probe_n = self._in(probe_s, np.float32), self._in(probe_c, np.uint32), \
            self._in(probe_n, np.uint32)
        nq = q256.shape[0]
        est = self._out(q256, (nq, p.cand), np.float32)
        rows = self._out(q256, (nq, p.cand), np.uint32)
        errs = self._out(q256, (nq, p.cand), np.float32)
        n = self._out(q256, (nq,), np.uint32)
        t = timing if timing is not None else abi.Timing()
        st = stream or self.stream()
        v = view.c if view is not None else None
        abi.check(abi.lib().sieve_candidates(self._h, st._h, abi.ptr(q256), nq, abi.ptr(probe_s), abi.ptr(probe_c),
                                             abi.ptr(probe_n), probe_s.shape[1], ctypes.byref(abi.make_params(p)),
                                             ctypes.byref(v) if v is not None else None, abi.ptr(est),
                                             abi.ptr(rows), abi.ptr(errs), abi.ptr(n), ctypes.byref(t)))
        return est, rows, errs, n

    def rerank(self, q768, q256, est, rows, errs, counts, p: Params, timing=None, stream=None):
        """Index::rerank for a batch: (scores nq x k, rows, counts)."""
        q768, q256 = self._in(q768, np.float32), self._in(q256, np.float32)
        est, rows, errs, counts = (self._in(est, np.float32), self._in(rows, np.uint32), self._in(errs, np.float32),
                                   self._in(counts, np.uint32))
        nq = q256.shape[0]
        s = self._out(q256, (nq, p.k), np.float32)
        r = self._out(q256, (nq, p.k), np.uint32)
        n = self._out(q256, (nq,), np.uint32)
        t = timing if timing is not None else abi.Timing()
        st = stream or self.stream()
        abi.check(abi.lib().sieve_rerank(self._h, st._h, abi.ptr(q768), abi.ptr(q256), nq, abi.ptr(est),
                                         abi.ptr(rows), abi.ptr(errs), abi.ptr(counts), est.shape[1],
                                         ctypes.byref(abi.make_params(p)), None, abi.ptr(s), abi.ptr(r), abi.ptr(n),
                                         ctypes.byref(t)))
        return s, r, n

    def autotune(self, q768, q256, p: Params, stream=None):
        """Times every scan-kernel variant on this batch and keeps the fastest:
        (chosen variant, ms per variant)."""
        q768, q256 = self._in(q768, np.float32), self._in(q256, np.float32)
        ch = ctypes.c_uint32()
        ms = np.zeros(abi.SCAN_VARIANTS, np.float32)
        st = stream or self.stream()
        abi.check(abi.lib().sieve_index_autotune(self._h, st._h, abi.ptr(q768), abi.ptr(q256), q256.shape[0],
                                                 ctypes.byref(abi.make_params(p)), ctypes.byref(ch), abi.ptr(ms)))
        return ch.value, ms.tolist()

    def set_scan_variant(self, v):
        abi.check(abi.lib().sieve_index_set_scan_variant(self._h, v))

    def set_scan_chunk(self, blocks):
        """Chunk size (32-slot blocks) of the scan's work split; 0 = chosen per batch."""
        abi.check(abi.lib().sieve_index_set_scan_chunk(self._h, blocks))

    def set_cluster_scan(self, min_queries=0, seed_slots=0, ext_cap=0, chunk_blocks=0, phases=0):
        """Settings of the cluster-major scan (scan variants 10, 11); 0 keeps a default."""
        abi.check(abi.lib().sieve_index_set_cluster_scan(self._h, min_queries, seed_slots, ext_cap, chunk_blocks,
                                                         phases))

    def set_small_batch(self, max_queries=12):
        """Largest unfiltered batch for the small-batch scan of the cluster-major variants
        (sieve_index_set_small_batch); 0 disables it."""
        abi.check(abi.lib().sieve_index_set_small_batch(self._h, max_queries))

    def set_route_fuse(self, mode=-1):
        """Routing GEMM fused with its select: -1 auto, 0 never, 1 whenever the select runs over
        tile maxima (sieve_index_set_route_fuse)."""
        abi.check(abi.lib().sieve_index_set_route_fuse(self._h, mode))

    def search(self, q768, q256, p: Params, view=None, timing=None, stream=None, out=None):
        """Index::search / search_view for a batch: (scores nq x k, rows, counts). `out`: the
        (scores, rows, counts) arrays to write, else new ones."""
        q768, q256 = self._in(q768, np.float32), self._in(q256, np.float32)
        nq = q256.shape[0]
        if out is not None:
            s, r, n = out
            assert _numel(s) == nq * p.k and _numel(r) == nq * p.k and _numel(n) == nq
        else:
            s = self._out(q256, (nq, p.k), np.float32)
            r = self._out(q256, (nq, p.k), np.uint32)
            n = self._out(q256, (nq,), np.uint32)
        t = timing if timing is not None else abi.Timing()
        st = stream or self.stream()
        v = view.c if view is not None else None
        abi.check(abi.lib().sieve_search_batch(self._h, st._h, abi.ptr(q768), abi.ptr(q256), nq,
                                               ctypes.byref(abi.make_params(p)),
                                               ctypes.byref(v) if v is not None else None, abi.ptr(s), abi.ptr(r),
                                               abi.ptr(n), ctypes.byref(t)))
        return s, r, n


class Filter:
    """filter::Filter over an Index (src/filter.rs:18-39) and filter::search with a Plan."""

    def __init__(self, index, allow):
        self._ix = index
        allow = index._in(allow, np.uint64)
        h = ctypes.c_void_p()
        abi.check(abi.lib().sieve_filter_create(index._h, abi.ptr(allow), _numel(allow), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            abi.lib().sieve_filter_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rows(self):
        n = ctypes.c_uint64()
        abi.check(abi.lib().sieve_filter_get_info(self._h, ctypes.byref(n), None))
        return n.value

    def per_cluster(self):
        out = np.empty(self._ix.k, np.uint32)
        abi.check(abi.lib().sieve_filter_get_info(self._h, None, abi.ptr(out)))
        return out

    def set_gather(self, enable=True):
        """Expand scans read only the filter's listed allowed slots (default), or whole clusters
        (sieve_filter_set_gather); the results are the same."""
        abi.check(abi.lib().sieve_filter_set_gather(self._h, 1 if enable else 0))

    def search(self, q768, q256, p: Params, plan="auto", target=200_000, timing=None, stream=None):
        """filter::search for a batch sharing this filter: (scores, rows, counts, plan used)."""
        ix = self._ix
        q768, q256 = ix._in(q768, np.float32), ix._in(q256, np.float32)
        nq = q256.shape[0]
        s = ix._out(q256, (nq, p.k), np.float32)
        r = ix._out(q256, (nq, p.k), np.uint32)
        n = ix._out(q256, (nq,), np.uint32)
        used = ctypes.c_uint32()
        t = timing if timing is not None else abi.Timing()
        st = stream or ix.stream()
        abi.check(abi.lib().sieve_filter_search_batch(ix._h, self._h, abi.PLANS[plan], target, st._h, abi.ptr(q768),
                                                      abi.ptr(q256), nq, ctypes.byref(abi.make_params(p)), abi.ptr(s),
                                                      abi.ptr(r), abi.ptr(n), ctypes.byref(used), ctypes.byref(t)))
        return s, r, n, used.value


class Live:
    """Live updates over an Index (src/live.rs): insert / delete / upsert by external id, and
    searches over the base index, the delta lists and the tombstones."""

    def __init__(self, index, max_rows, soar_lambda=1.0, index_dir=None, wal=None):
        """Live::open (src/live.rs:166-207). Without index_dir and wal: an index from arrays with
        the given soar_lambda. With index_dir (the directory the index was opened from):
        soar_lambda from its meta.json and a compacted index's row segment; with wal: the log
        replayed (self.replay holds its ReplayStats) and appended to from then on."""
        self._ix = index
        o = abi.LiveOpts(max_rows=max_rows, soar_lambda=float("nan") if soar_lambda is None else soar_lambda)
        h = ctypes.c_void_p()
        self.replay = None
        if index_dir is None and wal is None:
            abi.check(abi.lib().sieve_live_open(index._h, ctypes.byref(o), ctypes.byref(h)))
        else:
            st = abi.ReplayStats()
            abi.check(abi.lib().sieve_live_open_dir(index._h, ctypes.byref(o),
                                                    None if index_dir is None else str(index_dir).encode(),
                                                    None if wal is None else str(wal).encode(), ctypes.byref(st),
                                                    ctypes.byref(h)))
            self.replay = {f: getattr(st, f) for f, _ in st._fields_}
        self._h = h

    def compact(self, out_dir):
        """Live::compact (src/live.rs:395-468) into out_dir: its CompactStats."""
        st = abi.CompactStats()
        abi.check(abi.lib().sieve_live_compact(self._h, str(out_dir).encode(), ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def truncate_wal(self):
        abi.check(abi.lib().sieve_live_truncate_wal(self._h))

    def close(self):
        if getattr(self, "_h", None):
            abi.lib().sieve_live_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def insert(self, ext, vecs):
        ext = self._ix._in(ext, np.uint32)
        vecs = self._ix._in(vecs, np.float32)
        abi.check(abi.lib().sieve_live_insert(self._h, abi.ptr(ext), abi.ptr(vecs), _numel(ext)))

    def delete(self, ext):
        ext = self._ix._in(ext, np.uint32)
        abi.check(abi.lib().sieve_live_delete(self._h, abi.ptr(ext), _numel(ext)))

    def current(self, ext):
        r = abi.lib().sieve_live_current(self._h, int(ext))
        return None if r == 0xFFFFFFFF else r

    def info(self):
        st = abi.LiveInfo()
        abi.check(abi.lib().sieve_live_get_info(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def export(self):
        """The delta segment on the host: rows, norm256, t2, t2n, ext, dblock_off, codes
        (slots x 32), mul, err, ids."""
        inf = self.info()
        r, slots, k = inf["rows"], inf["arena_blocks"] * 32, self._ix.k
        out = dict(rows=np.empty((r, 768), np.int8), norm256=np.empty(r, np.float32), t2=np.empty((r, 64), np.uint8),
                   t2n=np.empty(r, np.float32), ext=np.empty(r, np.uint32), dblock_off=np.empty(k + 1, np.uint32),
                   codes=np.empty((slots, 32), np.uint8), mul=np.empty(slots, np.float32),
                   err=np.empty(slots, np.float32), ids=np.empty(slots, np.uint32))
        abi.check(abi.lib().sieve_live_export(self._h, *[abi.ptr(out[f]) for f in
                                                          ("rows", "norm256", "t2", "t2n", "ext", "dblock_off",
                                                           "codes", "mul", "err", "ids")]))
        return out

    def search(self, q768, q256, p: Params, view=None, timing=None, stream=None):
        """Live::search for a batch: (scores nq x k, external ids, counts)."""
        ix = self._ix
        q768, q256 = ix._in(q768, np.float32), ix._in(q256, np.float32)
        nq = q256.shape[0]
        s = ix._out(q256, (nq, p.k), np.float32)
        r = ix._out(q256, (nq, p.k), np.uint32)
        n = ix._out(q256, (nq,), np.uint32)
        t = timing if timing is not None else abi.Timing()
        st = stream or ix.stream()
        v = view.c if view is not None else None
        abi.check(abi.lib().sieve_live_search_batch(self._h, st._h, abi.ptr(q768), abi.ptr(q256), nq,
                                                    ctypes.byref(abi.make_params(p)),
                                                    ctypes.byref(v) if v is not None else None, abi.ptr(s),
                                                    abi.ptr(r), abi.ptr(n), ctypes.byref(t)))
        return s, r, n


class L2Index:
    """The L2 index over 100-d int8 vectors (bigann::Index, src/bigann.rs:604-747)."""

    def __init__(self, arrays, borrow_base=False):
        """arrays: leaves, block_off, blocks, ids, factors, rotation, base; optionally refine_bits and
        refine (the per-slot extra section of a multi-bit index, src/bigann.rs:430-436)."""
        keep = {}
        d = abi.L2Desc()
        dts = dict(leaves=np.float32, block_off=np.uint64, blocks=np.uint8, ids=np.uint32, factors=np.float32,
                   rotation=np.float32, base=np.int8)
        if arrays.get("refine") is not None and int(arrays.get("refine_bits", 1)) > 1:
            dts["refine"] = np.uint8
            d.refine_bits = int(arrays["refine_bits"])
        for name, dt in dts.items():
            a = arrays[name]
            a = a.contiguous() if _is_torch(a) else np.ascontiguousarray(a, dtype=dt)
            keep[name] = a
            setattr(d, name, abi.ptr(a).value)
        self.bits = max(1, int(d.refine_bits))
        d.k = _numel(keep["leaves"]) // 100
        d.n = _numel(keep["base"]) // 100
        d.nblocks = _numel(keep["ids"]) // 32
        d.borrow_base = 1 if borrow_base else 0
        h = ctypes.c_void_p()
        abi.check(abi.lib().sieve_l2_index_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self._keep = keep if borrow_base else None
        self.k = d.k
        self._stream = None

    @classmethod
    def open(cls, index_dir, base_dir):
        """bigann::Index::open over the reference's files (sieve_l2_index_open)."""
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        abi.check(abi.lib().sieve_l2_index_open(index_dir.encode(), base_dir.encode(), ctypes.byref(h)))
        self._h = h
        self._keep = None
        self._stream = None
        import os
        self.k = os.path.getsize(os.path.join(index_dir, "centroids.f32")) // 400
        bits_path = os.path.join(index_dir, "bits.u64")
        self.bits = int(np.fromfile(bits_path, np.uint64)[0]) if os.path.exists(bits_path) else 1
        return self

    def hot_leaves(self, q, nprobe, margin, budget, stream=None):
        """hot_leaves (src/bigann.rs:537-563): leaf ids to keep resident under `budget` record
        bytes, hottest per byte first, for these queries."""
        q = np.ascontiguousarray(q, dtype=np.int8)
        out = np.zeros(self.k, np.uint32)
        n = ctypes.c_uint32()
        if self._stream is None:
            self._stream = Stream()
        st = stream or self._stream
        abi.check(abi.lib().sieve_l2_hot_leaves(self._h, st._h, abi.ptr(q), q.shape[0], nprobe, margin, budget,
                                                abi.ptr(out), ctypes.byref(n)))
        return out[:n.value]

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                abi.lib().sieve_l2_index_free(self._h)
        except Exception:
            pass
        self._h = None

    def set_scan_chunk(self, blocks):
        """Chunk size (32-slot blocks) of the scan's work split; 0 = chosen per batch."""
        abi.check(abi.lib().sieve_l2_index_set_scan_chunk(self._h, blocks))

    def set_scan_variant(self, v):
        """One of the per-query scan variants (0 .. 9)."""
        abi.check(abi.lib().sieve_l2_index_set_scan_variant(self._h, v))

    def set_small_batch(self, max_queries=8):
        """Largest batch for the small-batch scan (sieve_l2_index_set_small_batch); 0 disables it."""
        abi.check(abi.lib().sieve_l2_index_set_small_batch(self._h, max_queries))

    def set_route_fuse(self, mode=-1):
        """Routing GEMM fused with its select: -1 auto, 0 never, 1 whenever the select runs over
        tile maxima (sieve_l2_index_set_route_fuse)."""
        abi.check(abi.lib().sieve_l2_index_set_route_fuse(self._h, mode))

    def search(self, q, nprobe, cand, k, margin=0.0, timing=None, stream=None, details=False, refine=0):
        """bigann::Index::search for a batch of int8 queries (nq x 100): (scores, rows, counts)
        and, with details, (probes, probe_counts, cand_rows, cand_est, cand_counts). refine > 0 on
        a multi-bit index re-estimates the 1-bit shortlist of max(refine, cand) slots."""
        q = q if _is_torch(q) else np.ascontiguousarray(q, dtype=np.int8)
        nq = q.shape[0]
        out = Index._out
        s, r, n = out(q, (nq, k), np.float32), out(q, (nq, k), np.uint32), out(q, (nq,), np.uint32)
        npb = min(nprobe, self.k)
        extra = [out(q, (nq, npb), np.uint32), out(q, (nq,), np.uint32), out(q, (nq, cand), np.uint32),
                 out(q, (nq, cand), np.float32), out(q, (nq,), np.uint32)] if details else [None] * 5
        t = timing if timing is not None else abi.Timing()
        if self._stream is None:
            self._stream = Stream()
        st = stream or self._stream
        p = abi.L2Params(nprobe=nprobe, cand=cand, k=k, refine=refine, margin=margin)
        abi.check(abi.lib().sieve_l2_search_batch(self._h, st._h, abi.ptr(q), nq, ctypes.byref(p), abi.ptr(s),
                                                  abi.ptr(r), abi.ptr(n), *[abi.ptr(x) for x in extra],
                                                  ctypes.byref(t)))
        return (s, r, n, *extra) if details else (s, r, n)


class ViewSpec:
    """ivf::View (src/ivf.rs:462-496): tombstones and allowlists as u64 bitsets over rows."""

    def __init__(self, deleted=None, allow=None, allow_flags=None):
        self._keep = [deleted, allow, allow_flags]
        v = abi.View()
        if deleted is not None:
            v.deleted = abi.ptr(deleted).value
            v.deleted_words = deleted.shape[0]
        if allow is not None:
            v.allow = abi.ptr(allow).value
            if allow.ndim == 2:
                v.allow_words = allow.shape[1]
                v.allow_stride = allow.shape[1]
            else:
                v.allow_words = allow.shape[0]
            if allow_flags is not None:
                v.allow_flags = abi.ptr(allow_flags).value
        self.c = v


def bitset(rows, n):
    """A u64 bitset over n rows with the given rows set (src/live.rs:123-129)."""
    words = np.zeros((n + 63) // 64, np.uint64)
    rows = np.asarray(rows, np.uint64)
    np.bitwise_or.at(words, (rows >> np.uint64(6)).astype(np.int64), np.uint64(1) << (rows & np.uint64(63)))
    return words


def _numel(a):
    return int(a.numel()) if _is_torch(a) else int(a.size)
