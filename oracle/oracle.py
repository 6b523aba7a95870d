"""CPU oracle for the sieve search path — TEST INFRASTRUCTURE, NOT PRODUCT CODE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only to check results (or, in bench.py, as the timed CPU baseline). It wraps
oracle/libsieve_oracle.so, a C++ restatement of the reference crate's x86 (non-NEON) path
(sieve_oracle.cpp; every function there cites the Rust it follows).

The reference is a Rust crate and no Rust toolchain exists in this image, so the reference
itself cannot be built or run (there is no oracle/_ref). The reference ships no golden vector
files; its own unit tests hold the known answers and invariants the oracle is pinned to.

Parity status of each oracle function and what pins it (tests/test_oracle_*.py):

  gemm_abt        (blas.rs:43-54)    pinned: blas.rs:58-65 known answer; float64 bound;
                                      bit-exact against a sequential float32 numpy sum
  Rng             (ivf.rs:120-135)   pinned: published splitmix64 output for seed 0/1234567
  lut_new         (lut.rs:33-56)     pinned: lut.rs:198-209 float bound; table invariants
                                      (min entry 0 per group, 255 in the widest group)
  score_code/scan (lut.rs:59-127)    pinned: lut.rs:179-196, :212-241 (kernel == score_code,
                                      G=32 fixture, partial block); brute-force bit sum
  pack/unpack     (lut.rs:73-94)    pinned: round trip (lut.rs:221-227); bit-level brute force
  hamming         (lut.rs:157-166)   pinned: lut.rs:243-248 known answer; numpy popcount
  fwht/rotation   (rabitq.rs:34-84)  pinned: explicit Hadamard matrix product (float64),
                                      orthogonality rabitq.rs:147-158
  rabitq_encode   (rabitq.rs:88-110) pinned: closed form of mul/err in float64;
                                      estimate bound rabitq.rs:159-188
  query_new/estimate (rabitq.rs:112-131) pinned: estimator identity
                                      a*lut+b == sum(sign_i * (Pq)_i) within LUT rounding
  normalize/sign_code/int8 (quant.rs) pinned: quant.rs:147-172 known answers
  query_i8/dot_i8i8 (quant.rs:104-135) pinned: integer identity vs numpy int64
  TopK            (topk.rs:4-62)     pinned: topk.rs:69-82 (== full sort)
  tier2           (tier2.rs)         pinned: tier2.rs:81-93 (table score == reconstruction)
  prune_threshold (ivf.rs:542-551)   pinned: ivf.rs:856-868 known answer
  route           (ivf.rs:589-614)   pinned: float64 brute-force probe set (up to f32 ties)
  candidates      (ivf.rs:647-733)   pinned: brute-force estimate over all probed slots,
                                      top-cand by full sort, dedup by row
  rerank          (ivf.rs:734-818)   pinned: brute-force int64 dots + full sort per mode
  l2 search       (bigann.rs:653-747) pinned: brute-force L2 estimate/rerank by full sort
  assign_top/soar (ivf.rs:138-192)   pinned: float64 brute force (up to near-ties)
  Filter/Plan     (filter.rs)        pinned: filter.rs:157-166 (auto plan); per-cluster
                                      counts and row lists by numpy; each plan's result
                                      against brute force over the allowed rows
  Live            (live.rs)          pinned: live.rs:493-520 (delta list layout, bitset);
                                      an inserted corpus vector encodes to the build's slots
                                      of that row (live.rs:2-6, bit for bit); upsert/delete
                                      semantics; live search == search over a rebuilt index
                                      holding the same rows (brute force, up to ties)
  Live log        (live.rs:131-135, :209-258, :470-479) pinned: record layout against bytes
                                      written by hand; replay == the writes applied directly
                                      (state and searches); torn tails cut, bad ops refused
  Live compact    (live.rs:395-468)  pinned: the merged lists against a rebuild of the same
                                      live rows' slots by numpy (order, padding, drops);
                                      searches of the compacted index == live searches
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsieve_oracle.so")

PDIM = 256
DIM = 768
BLOCK = 32
PAD = 0xFFFFFFFF

RERANK_TAGS = {"none": 0, "prefix": 1, "full": 2, "tiered": 3, "ssd": 4}


class _Params(ctypes.Structure):
    _fields_ = [("cand", ctypes.c_uint32), ("rerank", ctypes.c_uint32), ("rerank_r", ctypes.c_uint32),
                ("k", ctypes.c_uint32), ("prune_m", ctypes.c_uint32), ("prune_w", ctypes.c_float),
                ("nprobe", ctypes.c_uint32), ("margin", ctypes.c_float), ("budget", ctypes.c_uint64)]


class _Timing(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("route", "scan", "rerank", "tier2", "scanned", "reranked", "probed", "fetch", "wait")]


class _Desc(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("n", ctypes.c_uint64), ("nblocks", ctypes.c_uint64),
                ("centroids", ctypes.c_void_p), ("block_off", ctypes.c_void_p), ("blocks", ctypes.c_void_p),
                ("ids", ctypes.c_void_p), ("factors", ctypes.c_void_p), ("rotation", ctypes.c_void_p),
                ("rerank", ctypes.c_void_p), ("quant", ctypes.c_void_p), ("norm256", ctypes.c_void_p),
                ("tier2_sigma", ctypes.c_void_p), ("tier2_codes", ctypes.c_void_p),
                ("tier2_norm", ctypes.c_void_p), ("borrow_rerank", ctypes.c_uint32)]


class _View(ctypes.Structure):
    _fields_ = [("deleted", ctypes.c_void_p), ("deleted_words", ctypes.c_uint64), ("allow", ctypes.c_void_p),
                ("allow_words", ctypes.c_uint64), ("allow_stride", ctypes.c_uint64),
                ("allow_flags", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i32, u32, u64, f32 = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32,
                                      ctypes.c_uint64, ctypes.c_float)
        sig = {
            "or_rng_next": (u64, [vp]),
            "or_gemm_abt": (None, [vp, vp, sz, sz, sz, vp, i32]),
            "or_lut_new": (None, [vp, i32, vp, vp, vp]),
            "or_score_code": (ctypes.c_uint16, [vp, i32, vp]),
            "or_float_score": (f32, [vp, i32, vp]),
            "or_pack_block": (None, [vp, i32, i32, vp]),
            "or_unpack_slot": (None, [vp, i32, i32, vp]),
            "or_scan": (None, [vp, i32, vp, sz, vp]),
            "or_hamming": (u32, [vp, vp, i32]),
            "or_rotation_signs": (None, [u64, i32, vp]),
            "or_fwht": (None, [vp, i32]),
            "or_rotation_apply": (None, [vp, i32, vp]),
            "or_rabitq_encode": (None, [vp, vp, i32, vp, vp]),
            "or_query_new": (None, [vp, i32, vp, vp, vp]),
            "or_estimate": (f32, [f32, f32, ctypes.c_uint16, f32, f32]),
            "or_normalize": (f32, [vp, i32]),
            "or_sign_code": (None, [vp, i32, vp]),
            "or_int8_fit": (None, [vp, sz, vp]),
            "or_int8_encode": (None, [vp, vp, i32, vp]),
            "or_dot_f32": (f32, [vp, vp, i32]),
            "or_dot_i8": (f32, [vp, vp, i32]),
            "or_query_i8": (f32, [vp, i32, vp]),
            "or_dot_i8i8": (ctypes.c_int32, [vp, vp, i32]),
            "or_topk_f32": (sz, [vp, vp, sz, sz, vp, vp]),
            "or_topk_u16": (sz, [vp, vp, sz, sz, vp, vp]),
            "or_tier2_fit": (None, [vp, sz, vp]),
            "or_tier2_encode": (f32, [vp, vp, vp]),
            "or_tier2_tables": (None, [vp, vp, vp]),
            "or_tier2_score": (f32, [vp, vp, f32]),
            "or_prune_threshold": (f32, [vp, vp, sz, sz, f32]),
            "or_index_create": (vp, [ctypes.POINTER(_Desc)]),
            "or_index_free": (None, [vp]),
            "or_route": (sz, [vp, vp, ctypes.POINTER(_Params), vp, vp]),
            "or_candidates": (sz, [vp, vp, vp, vp, sz, ctypes.POINTER(_Params), ctypes.POINTER(_View), sz,
                                   vp, vp, vp, vp]),
            "or_rerank": (sz, [vp, vp, vp, vp, vp, vp, sz, ctypes.POINTER(_Params), vp, vp, vp]),
            "or_search_batch": (None, [vp, vp, vp, sz, ctypes.POINTER(_Params), ctypes.POINTER(_View), i32,
                                       vp, vp, vp, ctypes.POINTER(_Timing)]),
            "or_search_latency": (None, [vp, vp, vp, sz, ctypes.POINTER(_Params), vp]),
            "or_assign_top": (None, [vp, sz, vp, sz, sz, vp, i32]),
            "or_soar_secondary": (u32, [vp, vp, vp, sz, f32]),
            "or_kmeans": (None, [vp, sz, sz, sz, ctypes.c_uint64, vp, i32]),
            "or_l2_norms": (None, [vp, sz, vp]),
            "or_top_l2": (None, [vp, sz, vp, vp, sz, sz, vp, vp, i32]),
            "or_kmeans_l2": (None, [vp, sz, sz, sz, ctypes.c_uint64, vp, i32]),
            "or_l2_train": (None, [vp, sz, sz, sz, sz, sz, vp, vp, i32]),
            "or_l2_assign": (None, [vp, sz, vp, sz, vp, sz, vp, i32]),
            "or_build_pairs": (None, [vp, sz, vp, sz, f32, vp, i32]),
            "or_l2_leaf_norms": (None, [vp, sz, vp]),
            "or_l2_create": (vp, [u32, vp, vp, vp, vp, vp, vp, vp, vp]),
            "or_l2_free": (None, [vp]),
            "or_l2_search": (sz, [vp, vp, u32, u32, u32, f32, vp, vp, vp, vp, vp, vp, vp]),
            "or_l2_search_batch": (None, [vp, vp, sz, u32, u32, u32, f32, u32, i32, vp, vp, vp]),
            "or_l2_search_latency": (None, [vp, vp, sz, u32, u32, u32, f32, u32, vp]),
            "or_l2_search_refine": (sz, [vp, vp, u32, u32, u32, f32, u32, vp, vp, vp, vp, vp, vp, vp]),
            "or_l2_set_refine": (None, [vp, i32, vp]),
            "or_refine_encode": (None, [vp, i32, vp, vp]),
            "or_refine_estimate": (f32, [vp, i32, vp, vp, f32]),
            "or_l2_hot_leaves": (sz, [vp, vp, sz, u32, f32, vp, ctypes.c_uint64, vp]),
            "or_live_create": (vp, [vp, i32, f32]),
            "or_delta_list_fixture": (sz, [sz, vp, vp, vp, sz, vp, vp, vp]),
            "or_filter_create": (vp, [vp, vp, sz]),
            "or_filter_free": (None, [vp]),
            "or_filter_rows": (sz, [vp]),
            "or_filter_per_cluster": (None, [vp, vp]),
            "or_plan_resolve": (u32, [u32, u64, u64]),
            "or_filter_search": (u32, [vp, vp, u32, u64, vp, vp, ctypes.POINTER(_Params), vp, vp, vp]),
            "or_live_free": (None, [vp]),
            "or_live_insert": (i32, [vp, vp, vp, sz]),
            "or_live_delete": (i32, [vp, vp, sz]),
            "or_live_open_wal": (i32, [vp, ctypes.c_char_p, vp]),
            "or_live_truncate_wal": (i32, [vp]),
            "or_live_load_segment": (None, [vp, vp, vp, vp, vp, vp, sz, vp, sz]),
            "or_live_compact": (sz, [vp, sz, vp, vp, vp, vp, vp]),
            "or_live_current": (u32, [vp, u32]),
            "or_live_rows": (sz, [vp]),
            "or_live_row": (None, [vp, sz, vp, vp, vp, vp, vp]),
            "or_live_list": (sz, [vp, u32, sz, vp, vp, vp]),
            "or_live_deleted": (sz, [vp, vp, sz]),
            "or_live_search_batch": (None, [vp, vp, vp, sz, ctypes.POINTER(_Params), ctypes.POINTER(_View), i32,
                                            vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------------ small kernels ----
def rng_stream(seed, n):
    st = np.array([seed], dtype=np.uint64)
    return np.array([lib().or_rng_next(_p(st)) for _ in range(n)], dtype=np.uint64)


def gemm_abt(a, b, threads=1):
    a, b = _c(a, np.float32), _c(b, np.float32)
    m, k = a.shape
    n = b.shape[0]
    c = np.empty((m, n), np.float32)
    lib().or_gemm_abt(_p(a), _p(b), m, n, k, _p(c), threads)
    return c


def lut_new(q):
    q = _c(q, np.float32)
    G = q.shape[0] // 4
    t = np.empty((G, 16), np.uint8)
    s, b = ctypes.c_float(), ctypes.c_float()
    lib().or_lut_new(_p(q), G, _p(t), ctypes.byref(s), ctypes.byref(b))
    return t, s.value, b.value


def score_code(tables, code):
    tables, code = _c(tables, np.uint8), _c(code, np.uint8)
    return lib().or_score_code(_p(tables), tables.shape[0], _p(code))


def float_score(q, code):
    q, code = _c(q, np.float32), _c(code, np.uint8)
    return lib().or_float_score(_p(q), q.shape[0], _p(code))


def pack_block(codes, G):
    codes = _c(codes, np.uint8).reshape(-1, G // 2)
    out = np.empty(G * 16, np.uint8)
    lib().or_pack_block(_p(codes), codes.shape[0], G, _p(out))
    return out


def unpack_slot(block, k):
    block = _c(block, np.uint8)
    G = block.shape[0] // 16
    code = np.empty(G // 2, np.uint8)
    lib().or_unpack_slot(_p(block), G, k, _p(code))
    return code


def scan(tables, blocks):
    tables, blocks = _c(tables, np.uint8), _c(blocks, np.uint8)
    G = tables.shape[0]
    nb = blocks.size // (G * 16)
    out = np.empty(nb * BLOCK, np.uint16)
    lib().or_scan(_p(tables), G, _p(blocks), nb, _p(out))
    return out


def hamming(a, b):
    a, b = _c(a, np.uint8), _c(b, np.uint8)
    return lib().or_hamming(_p(a), _p(b), a.shape[0])


def rotation_signs(seed, d=PDIM):
    s = np.empty(3 * d, np.float32)
    lib().or_rotation_signs(seed, d, _p(s))
    return s


def fwht(v):
    v = _c(v, np.float32).copy()
    lib().or_fwht(_p(v), v.shape[0])
    return v


def rotation_apply(signs, v):
    v = _c(v, np.float32).copy()
    signs = _c(signs, np.float32)
    lib().or_rotation_apply(_p(signs), v.shape[0], _p(v))
    return v


def rabitq_encode(xr, cr):
    xr, cr = _c(xr, np.float32), _c(cr, np.float32)
    d = xr.shape[0]
    code = np.empty(d // 8, np.uint8)
    me = np.empty(2, np.float32)
    lib().or_rabitq_encode(_p(xr), _p(cr), d, _p(code), _p(me))
    return code, me


def query_new(signs, q):
    """Returns (tables G x 16 u8, scale, bias, a, b) of rabitq::Query::new."""
    signs, q = _c(signs, np.float32), _c(q, np.float32)
    d = q.shape[0]
    t = np.empty((d // 4, 16), np.uint8)
    o = np.empty(4, np.float32)
    lib().or_query_new(_p(signs), d, _p(q), _p(t), _p(o))
    return t, o[0], o[1], o[2], o[3]


def estimate(cent, mul, sc, a, b):
    return lib().or_estimate(cent, mul, int(sc), a, b)


def normalize(v):
    v = _c(v, np.float32).copy()
    n = lib().or_normalize(_p(v), v.shape[0])
    return v, n


def sign_code(p):
    p = _c(p, np.float32)
    out = np.empty(p.shape[0] // 8, np.uint8)
    lib().or_sign_code(_p(p), p.shape[0], _p(out))
    return out


def int8_fit(rows):
    rows = _c(rows, np.float32).reshape(-1, DIM)
    s = np.empty(DIM, np.float32)
    lib().or_int8_fit(_p(rows), rows.shape[0], _p(s))
    return s


def int8_encode(v, scale):
    v, scale = _c(v, np.float32), _c(scale, np.float32)
    out = np.empty(v.shape[0], np.int8)
    lib().or_int8_encode(_p(v), _p(scale), v.shape[0], _p(out))
    return out


def dot_f32(a, b):
    a, b = _c(a, np.float32), _c(b, np.float32)
    return lib().or_dot_f32(_p(a), _p(b), a.shape[0])


def dot_i8(w, x):
    w, x = _c(w, np.float32), _c(x, np.int8)
    return lib().or_dot_i8(_p(w), _p(x), w.shape[0])


def query_i8(w):
    w = _c(w, np.float32)
    out = np.empty(w.shape[0], np.int8)
    s = lib().or_query_i8(_p(w), w.shape[0], _p(out))
    return out, s


def dot_i8i8(a, b):
    a, b = _c(a, np.int8), _c(b, np.int8)
    return lib().or_dot_i8i8(_p(a), _p(b), a.shape[0])


def topk(scores, k, ids=None):
    scores = np.ascontiguousarray(scores)
    n = scores.shape[0]
    ids = _c(np.arange(n) if ids is None else ids, np.uint32)
    if scores.dtype == np.uint16:
        os_ = np.empty(max(k, 1), np.uint16)
        oi = np.empty(max(k, 1), np.uint32)
        c = lib().or_topk_u16(_p(scores), _p(ids), n, k, _p(os_), _p(oi))
    else:
        scores = _c(scores, np.float32)
        os_ = np.empty(max(k, 1), np.float32)
        oi = np.empty(max(k, 1), np.uint32)
        c = lib().or_topk_f32(_p(scores), _p(ids), n, k, _p(os_), _p(oi))
    return os_[:c], oi[:c]


def tier2_fit(rows):
    rows = _c(rows, np.float32).reshape(-1, PDIM)
    s = np.empty(PDIM, np.float32)
    lib().or_tier2_fit(_p(rows), rows.shape[0], _p(s))
    return s


def tier2_encode(sigma, prefix):
    sigma, prefix = _c(sigma, np.float32), _c(prefix, np.float32)
    out = np.empty(PDIM // 4, np.uint8)
    n = lib().or_tier2_encode(_p(sigma), _p(prefix), _p(out))
    return out, n


def tier2_tables(sigma, q):
    sigma, q = _c(sigma, np.float32), _c(q, np.float32)
    t = np.empty((PDIM // 4, 256), np.float32)
    lib().or_tier2_tables(_p(sigma), _p(q), _p(t))
    return t


def tier2_score(tables, code, norm):
    tables, code = _c(tables, np.float32), _c(code, np.uint8)
    return lib().or_tier2_score(_p(tables), _p(code), norm)


def prune_threshold(est, err, m, w):
    est, err = _c(est, np.float32), _c(err, np.float32)
    return lib().or_prune_threshold(_p(est), _p(err), est.shape[0], m, w)


def assign_top(rows, cents, m, threads=8):
    rows, cents = _c(rows, np.float32), _c(cents, np.float32)
    out = np.empty((rows.shape[0], m), np.uint32)
    lib().or_assign_top(_p(rows), rows.shape[0], _p(cents), cents.shape[0], m, _p(out), threads)
    return out


def soar_secondary(x, cents, cand, lam):
    x, cents, cand = _c(x, np.float32), _c(cents, np.float32), _c(cand, np.uint32)
    return lib().or_soar_secondary(_p(x), _p(cents), _p(cand), cand.shape[0], lam)


def kmeans(sample, k, iters, seed=7, threads=8):
    """kmeans (src/ivf.rs:188-225, balance None): k x 256 centroids."""
    sample = _c(sample, np.float32)
    out = np.empty((k, PDIM), np.float32)
    lib().or_kmeans(_p(sample), sample.shape[0], k, iters, seed, _p(out), threads)
    return out


def l2_norms(cents):
    """norms (src/bigann.rs:135-137) of k x 100 centroids."""
    cents = _c(cents, np.float32)
    out = np.empty(cents.shape[0], np.float32)
    lib().or_l2_norms(_p(cents), cents.shape[0], _p(out))
    return out


def top_l2(rows, cents, m, threads=8):
    """top_l2 (src/bigann.rs:117-133): (ids, scores) n x m, best first."""
    rows, cents = _c(rows, np.float32), _c(cents, np.float32)
    cn = l2_norms(cents)
    ids = np.empty((rows.shape[0], m), np.uint32)
    sc = np.empty((rows.shape[0], m), np.float32)
    lib().or_top_l2(_p(rows), rows.shape[0], _p(cents), _p(cn), cents.shape[0], m, _p(ids), _p(sc), threads)
    return ids, sc


def kmeans_l2(sample, k, iters, seed, threads=8):
    """kmeans_l2 (src/bigann.rs:139-161): k x 100 centroids."""
    sample = _c(sample, np.float32)
    out = np.empty((k, 100), np.float32)
    lib().or_kmeans_l2(_p(sample), sample.shape[0], k, iters, seed, _p(out), threads)
    return out


def l2_train(base, k1, k2, sample_n, iters, threads=8):
    """train (src/bigann.rs:175-227) without the files: (coarse k1 x 100, leaves k1 k2 x 100)."""
    base = _c(base, np.int8)
    coarse = np.empty((k1, 100), np.float32)
    leaves = np.empty((k1 * k2, 100), np.float32)
    lib().or_l2_train(_p(base), base.shape[0], k1, k2, sample_n, iters, _p(coarse), _p(leaves), threads)
    return coarse, leaves


def l2_assign(rows, coarse, leaves, k2, threads=8):
    """Model::assign (src/bigann.rs:277-302): each row's leaf."""
    rows, coarse, leaves = _c(rows, np.float32), _c(coarse, np.float32), _c(leaves, np.float32)
    out = np.empty(rows.shape[0], np.uint32)
    lib().or_l2_assign(_p(rows), rows.shape[0], _p(coarse), coarse.shape[0], _p(leaves), k2, _p(out), threads)
    return out


def build_pairs(rows, cents, lam, threads=8):
    """(primary, secondary) cluster of each row as ivf::build assigns them (src/ivf.rs:313-321);
    lam None: no SOAR (secondary 0xFFFFFFFF)."""
    rows, cents = _c(rows, np.float32), _c(cents, np.float32)
    out = np.empty((rows.shape[0], 2), np.uint32)
    lib().or_build_pairs(_p(rows), rows.shape[0], _p(cents), cents.shape[0],
                         float("nan") if lam is None else lam, _p(out), threads)
    return out


# ------------------------------------------------------------------ params / view ----
def make_params(p):
    """p: any object with nprobe, cand, k, rerank (str tag or (tag, r)), prune, margin, budget."""
    rr = p.rerank
    tag, r = (rr, 0) if isinstance(rr, str) else rr
    prune = p.prune if p.prune is not None else (0, 0.0)
    margin = float("nan") if p.margin is None else p.margin
    return _Params(cand=p.cand, rerank=RERANK_TAGS[tag], rerank_r=r, k=p.k, prune_m=prune[0], prune_w=prune[1],
                   nprobe=p.nprobe, margin=margin, budget=p.budget or 0)


def make_view(deleted=None, allow=None, allow_stride=0, allow_flags=None):
    v = _View()
    if deleted is not None:
        v.deleted = deleted.ctypes.data
        v.deleted_words = deleted.shape[0]
    if allow is not None:
        v.allow = allow.ctypes.data
        v.allow_words = allow.shape[-1] if allow_stride else allow.shape[0]
        v.allow_stride = allow_stride
        if allow_flags is not None:
            v.allow_flags = allow_flags.ctypes.data
    return v


# ------------------------------------------------------------------ the index --------
class Index:
    """The reference's ivf::Index + Corpus over arrays in the on-disk layout (host numpy)."""

    def __init__(self, arrays):
        keep = {}
        for name, dt in (("centroids", np.float32), ("block_off", np.uint64), ("blocks", np.uint8),
                         ("ids", np.uint32), ("factors", np.float32), ("rotation", np.float32),
                         ("rerank", np.int8), ("quant", np.float32), ("norm256", np.float32),
                         ("tier2_sigma", np.float32), ("tier2_codes", np.uint8), ("tier2_norm", np.float32)):
            a = arrays.get(name)
            keep[name] = None if a is None else np.ascontiguousarray(a, dtype=dt)
        self._keep = keep
        d = _Desc()
        d.k = keep["centroids"].size // PDIM
        d.n = keep["rerank"].size // DIM
        d.nblocks = keep["ids"].size // BLOCK
        for name, a in keep.items():
            setattr(d, name, None if a is None else a.ctypes.data)
        self.k, self.n, self.nblocks = d.k, d.n, d.nblocks
        self._h = lib().or_index_create(ctypes.byref(d))

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().or_index_free(self._h)
        except Exception:
            pass
        self._h = None

    def route(self, q256, p):
        q = _c(q256, np.float32)
        s = np.empty(max(p.nprobe, 1), np.float32)
        c = np.empty(max(p.nprobe, 1), np.uint32)
        n = lib().or_route(self._h, _p(q), ctypes.byref(make_params(p)), _p(s), _p(c))
        return s[:n], c[:n]

    def candidates(self, q256, probe_scores, probe_clusters, p, view=None, qi=0):
        q = _c(q256, np.float32)
        ps, pc = _c(probe_scores, np.float32), _c(probe_clusters, np.uint32)
        est = np.empty(max(p.cand, 1), np.float32)
        rows = np.empty(max(p.cand, 1), np.uint32)
        errs = np.empty(max(p.cand, 1), np.float32)
        sc = np.zeros(1, np.uint64)
        v = view if view is not None else _View()
        n = lib().or_candidates(self._h, _p(q), _p(ps), _p(pc), ps.shape[0], ctypes.byref(make_params(p)),
                                ctypes.byref(v), qi, _p(est), _p(rows), _p(errs), _p(sc))
        return est[:n], rows[:n], errs[:n], int(sc[0])

    def rerank(self, q768, q256, est, rows, errs, p):
        q7, q2 = _c(q768, np.float32), _c(q256, np.float32)
        est, rows, errs = _c(est, np.float32), _c(rows, np.uint32), _c(errs, np.float32)
        os_ = np.empty(max(p.k, 1), np.float32)
        orr = np.empty(max(p.k, 1), np.uint32)
        rr = np.zeros(1, np.uint64)
        n = lib().or_rerank(self._h, _p(q7), _p(q2), _p(est), _p(rows), _p(errs), est.shape[0],
                            ctypes.byref(make_params(p)), _p(os_), _p(orr), _p(rr))
        return os_[:n], orr[:n], int(rr[0])

    def search_batch(self, q768, q256, p, view=None, threads=1):
        q7, q2 = _c(q768, np.float32), _c(q256, np.float32)
        nq = q7.shape[0]
        s = np.zeros((nq, p.k), np.float32)
        r = np.full((nq, p.k), PAD, np.uint32)
        cnt = np.zeros(nq, np.uint32)
        t = _Timing()
        v = view if view is not None else _View()
        lib().or_search_batch(self._h, _p(q7), _p(q2), nq, ctypes.byref(make_params(p)), ctypes.byref(v), threads,
                              _p(s), _p(r), _p(cnt), ctypes.byref(t))
        return s, r, cnt, {f: getattr(t, f) for f, _ in _Timing._fields_}

    def search_latency(self, q768, q256, p):
        q7, q2 = _c(q768, np.float32), _c(q256, np.float32)
        lat = np.zeros(q7.shape[0], np.uint64)
        lib().or_search_latency(self._h, _p(q7), _p(q2), q7.shape[0], ctypes.byref(make_params(p)), _p(lat))
        return lat


class L2Index:
    """bigann::Index over arrays (leaves k x 100, blocks of 512 B, [mul, |x|^2] factors)."""

    def __init__(self, leaves, block_off, blocks, factors, ids, rotation, base):
        self._keep = [np.ascontiguousarray(leaves, np.float32), np.ascontiguousarray(block_off, np.uint64),
                      np.ascontiguousarray(blocks, np.uint8), np.ascontiguousarray(factors, np.float32),
                      np.ascontiguousarray(ids, np.uint32), np.ascontiguousarray(rotation, np.float32),
                      np.ascontiguousarray(base, np.int8)]
        lv = self._keep[0]
        self.k = lv.shape[0]
        self.leaf_norm = np.empty(self.k, np.float32)
        lib().or_l2_leaf_norms(_p(lv), self.k, _p(self.leaf_norm))
        lv, bo, bl, fa, ids, rot, base = self._keep
        self._h = lib().or_l2_create(self.k, _p(lv), _p(self.leaf_norm), _p(bo), _p(bl), _p(fa), _p(ids), _p(rot),
                                     _p(base))

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().or_l2_free(self._h)
        except Exception:
            pass
        self._h = None

    def set_refine(self, bits, extra):
        """The multi-bit refinement section (src/bigann.rs:38-113): `bits` and, per slot in slot
        order, refine_bytes(bits) bytes (the low bits - 1 planes, then f as LE f32)."""
        self._extra = _c(extra, np.uint8)
        lib().or_l2_set_refine(self._h, int(bits), _p(self._extra))

    def search_batch(self, q, nprobe, cand, k, margin=0.0, threads=1, refine=0):
        """or_l2_search over a batch (nq x 100), whole queries per worker: (scores, rows, counts)."""
        q = _c(q, np.int8)
        nq = q.shape[0]
        s = np.zeros((nq, k), np.float32)
        r = np.full((nq, k), PAD, np.uint32)
        cnt = np.zeros(nq, np.uint32)
        lib().or_l2_search_batch(self._h, _p(q), nq, nprobe, cand, k, margin, refine, threads, _p(s), _p(r), _p(cnt))
        return s, r, cnt

    def search_latency(self, q, nprobe, cand, k, margin=0.0, refine=0):
        """Single-thread latency of each query, ns."""
        q = _c(q, np.int8)
        lat = np.zeros(q.shape[0], np.uint64)
        lib().or_l2_search_latency(self._h, _p(q), q.shape[0], nprobe, cand, k, margin, refine, _p(lat))
        return lat

    def hot_leaves(self, q, nprobe, margin, record_bytes, budget):
        """hot_leaves (src/bigann.rs:537-563): leaf ids, hottest per byte first, within budget."""
        q = _c(q, np.int8)
        rb = _c(record_bytes, np.uint64)
        out = np.zeros(self.k, np.uint32)
        n = lib().or_l2_hot_leaves(self._h, _p(q), q.shape[0], nprobe, margin, _p(rb), budget, _p(out))
        return out[:n]

    def search(self, q, nprobe, cand, k, margin=0.0, refine=0):
        q = _c(q, np.int8)
        probe = np.empty(max(nprobe, 1), np.uint32)
        npo = np.zeros(1, np.uint32)
        ci = np.empty(max(cand, 1), np.uint32)
        ce = np.empty(max(cand, 1), np.float32)
        nc = np.zeros(1, np.uint32)
        os_ = np.empty(max(k, 1), np.float32)
        orr = np.empty(max(k, 1), np.uint32)
        n = lib().or_l2_search_refine(self._h, _p(q), nprobe, cand, k, margin, refine, _p(probe), _p(npo), _p(ci),
                                      _p(ce), _p(nc), _p(os_), _p(orr))
        return (probe[:npo[0]], ci[:nc[0]], ce[:nc[0]], os_[:n], orr[:n])


def refine_bytes(bits):
    """Bytes per slot of the refinement section (src/bigann.rs:110-113)."""
    return (bits - 1) * 16 + 4 if bits > 1 else 0


def refine_encode(r, bits):
    """refine_encode (src/bigann.rs:70-108) of one rotated 128-d residual: (planes 32 B, f)."""
    r = _c(r, np.float32)
    planes = np.zeros(32, np.uint8)
    f = np.zeros(1, np.float32)
    lib().or_refine_encode(_p(r), int(bits), _p(planes), _p(f))
    return planes, float(f[0])


def refine_estimate(qr, bits, sign, planes, f):
    """Refiner::new(qr, bits).estimate(sign, planes, f) (src/bigann.rs:565-602)."""
    qr, sign, planes = _c(qr, np.float32), _c(sign, np.uint8), _c(planes, np.uint8)
    return float(lib().or_refine_estimate(_p(qr), int(bits), _p(sign), _p(planes), np.float32(f)))


def delta_list_fixture(codes, factors, ids):
    """DeltaList::push of each (code, factors, id) into a fresh list: (blocks, factors, ids)."""
    codes, factors, ids = _c(codes, np.uint8), _c(factors, np.float32), _c(ids, np.uint32)
    n = codes.shape[0]
    nb = lib().or_delta_list_fixture(n, _p(codes), _p(factors), _p(ids), 0, None, None, None)
    b = np.empty((nb, 1024), np.uint8)
    f = np.empty((nb * BLOCK, 2), np.float32)
    i = np.empty(nb * BLOCK, np.uint32)
    lib().or_delta_list_fixture(n, _p(codes), _p(factors), _p(ids), nb, _p(b), _p(f), _p(i))
    return b, f, i


class Live:
    """Live over an oracle Index (src/live.rs): inserts into per-cluster delta lists, deletes as
    tombstones, external ids."""

    def __init__(self, index, soar_lambda=1.0):
        self._ix = index  # keeps the index alive
        soar = soar_lambda is not None
        self._h = lib().or_live_create(index._h, 1 if soar else 0, float(soar_lambda) if soar else 0.0)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().or_live_free(self._h)
        except Exception:
            pass
        self._h = None

    def insert(self, ext, vecs):
        ext, vecs = _c(ext, np.uint32), _c(vecs, np.float32)
        if lib().or_live_insert(self._h, _p(ext), _p(vecs), ext.shape[0]) != 0:
            raise OSError("oracle Live.insert: log write failed")

    def delete(self, ext):
        ext = _c(ext, np.uint32)
        if lib().or_live_delete(self._h, _p(ext), ext.shape[0]) != 0:
            raise OSError("oracle Live.delete: log write failed")

    def open_wal(self, path):
        """Live::replay of `path`, then appends every later write to it: {inserts, deletes, torn}.
        Raises ValueError (with the byte offset) on an unknown op."""
        st = np.zeros(4, np.uint64)
        rc = lib().or_live_open_wal(self._h, str(path).encode(), _p(st))
        if rc == -1:
            raise ValueError(f"{path}: bad op at byte {int(st[3])}")
        if rc != 0:
            raise OSError(f"{path}: replay failed")
        return dict(inserts=int(st[0]), deletes=int(st[1]), torn=int(st[2]))

    def truncate_wal(self):
        if lib().or_live_truncate_wal(self._h) != 0:
            raise OSError("truncate_wal failed")

    def load_segment(self, rows, norm256, t2, t2n, ext, deleted):
        """The row segment of a compacted index (live.* files) as Live::open reads it."""
        rows, norm256, ext = _c(rows, np.int8), _c(norm256, np.float32), _c(ext, np.uint32)
        deleted = _c(deleted, np.uint64)
        has_t2 = t2 is not None and len(t2) > 0
        t2 = _c(t2, np.uint8) if has_t2 else None
        t2n = _c(t2n, np.float32) if has_t2 else None
        self._seg = (rows, norm256, t2, t2n, ext, deleted)
        lib().or_live_load_segment(self._h, _p(rows), _p(norm256), _p(t2) if has_t2 else None,
                                   _p(t2n) if has_t2 else None, _p(ext), ext.shape[0], _p(deleted),
                                   deleted.shape[0])

    def compact(self):
        """Live::compact's merged lists: dict(block_off u64 k+1, blocks nb x 1024, factors
        nb*32 x 2, ids nb*32, slots, dropped)."""
        st = np.zeros(2, np.uint64)
        nb = lib().or_live_compact(self._h, 0, None, None, None, None, _p(st))
        k = self._ix.k
        off = np.zeros(k + 1, np.uint64)
        b = np.zeros((nb, 1024), np.uint8)
        f = np.zeros((nb * BLOCK, 2), np.float32)
        i = np.zeros(nb * BLOCK, np.uint32)
        lib().or_live_compact(self._h, nb, _p(off), _p(b), _p(f), _p(i), _p(st))
        return dict(block_off=off, blocks=b, factors=f, ids=i, slots=int(st[0]), dropped=int(st[1]))

    def current(self, ext):
        r = lib().or_live_current(self._h, int(ext))
        return None if r == PAD else r

    def rows(self):
        return lib().or_live_rows(self._h)

    def row(self, j):
        row = np.empty(DIM, np.int8)
        t2 = np.zeros(64, np.uint8)
        norm, t2n, ext = ctypes.c_float(), ctypes.c_float(0.0), ctypes.c_uint32()
        lib().or_live_row(self._h, j, _p(row), ctypes.byref(norm), _p(t2), ctypes.byref(t2n), ctypes.byref(ext))
        return row, norm.value, t2, t2n.value, ext.value

    def delta_list(self, c):
        """(raw codes n x 32, factors n x 2, row ids n) of cluster c's delta list."""
        n = lib().or_live_list(self._h, c, 0, None, None, None)
        codes = np.empty((n, 32), np.uint8)
        f = np.empty((n, 2), np.float32)
        ids = np.empty(n, np.uint32)
        lib().or_live_list(self._h, c, n, _p(codes), _p(f), _p(ids))
        return codes, f, ids

    def deleted(self):
        n = lib().or_live_deleted(self._h, None, 0)
        w = np.zeros(n, np.uint64)
        lib().or_live_deleted(self._h, _p(w), n)
        return w

    def search_batch(self, q768, q256, p, view=None, threads=1):
        q7, q2 = _c(q768, np.float32), _c(q256, np.float32)
        nq = q7.shape[0]
        s = np.zeros((nq, p.k), np.float32)
        r = np.full((nq, p.k), PAD, np.uint32)
        cnt = np.zeros(nq, np.uint32)
        v = view if view is not None else _View()
        lib().or_live_search_batch(self._h, _p(q7), _p(q2), nq, ctypes.byref(make_params(p)), ctypes.byref(v),
                                   threads, _p(s), _p(r), _p(cnt))
        return s, r, cnt


PLANS = {"post": 0, "expand": 1, "brute": 2, "auto": 3}


def plan_resolve(plan, rows, n):
    return lib().or_plan_resolve(PLANS[plan] if isinstance(plan, str) else plan, rows, n)


class Filter:
    """filter::Filter over an oracle Index (src/filter.rs:18-39)."""

    def __init__(self, index, allow):
        self._ix = index
        self._allow = _c(allow, np.uint64)
        self._h = lib().or_filter_create(index._h, _p(self._allow), self._allow.shape[0])

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().or_filter_free(self._h)
        except Exception:
            pass
        self._h = None

    def rows(self):
        return lib().or_filter_rows(self._h)

    def per_cluster(self):
        out = np.empty(self._ix.k, np.uint32)
        lib().or_filter_per_cluster(self._h, _p(out))
        return out

    def search(self, q768, q256, p, plan="auto", target=200_000):
        """filter::search for each query: (scores nq x k, rows, counts, resolved plans)."""
        q7, q2 = _c(q768, np.float32), _c(q256, np.float32)
        nq = q7.shape[0]
        s = np.zeros((nq, p.k), np.float32)
        r = np.full((nq, p.k), PAD, np.uint32)
        cnt = np.zeros(nq, np.uint32)
        plans = np.zeros(nq, np.uint32)
        pr = make_params(p)
        c = ctypes.c_uint32()
        for i in range(nq):
            plans[i] = lib().or_filter_search(self._ix._h, self._h, PLANS[plan], target, _p(q7[i]), _p(q2[i]),
                                              ctypes.byref(pr), _p(s[i]), _p(r[i]), ctypes.byref(c))
            cnt[i] = c.value
        return s, r, cnt, plans
