"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same seeded index.

Integer work (LUT sums, int8 dots) and the reference's f32 steps are bit-exact, so probe
scores, estimates, error half-widths and rerank scores are compared bit for bit; id sets are
compared up to ties at the selection boundary (the reference breaks those arbitrarily).
"""
import numpy as np
import pytest

from oracle import oracle as O
from sieve_amd import Params, Rerank
from sieve_amd.index import Index, ViewSpec, bitset

pytestmark = pytest.mark.gpu

PAD = 0xFFFFFFFF


@pytest.fixture(scope="module")
def gpu_index(small):
    print(f"[gpu] creating index k={small.k} n={small.n}", flush=True)
    return Index.create(small.host)


def _flt_eq(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def assert_same_up_to_ties(g_rows, g_sc, w_rows, w_sc, boundary_inclusive=True):
    """Both lists best first. Every entry scoring above the lowest kept score of either list
    must be in both with a bit-identical score."""
    g_rows, w_rows = list(map(int, g_rows)), list(map(int, w_rows))
    g_sc, w_sc = np.asarray(g_sc, np.float32), np.asarray(w_sc, np.float32)
    if len(g_rows) == 0 or len(w_rows) == 0:
        assert len(g_rows) == len(w_rows)
        return
    b = max(g_sc.min(), w_sc.min())
    g = {r: s for r, s in zip(g_rows, g_sc) if s > b}
    w = {r: s for r, s in zip(w_rows, w_sc) if s > b}
    assert g.keys() == w.keys()
    for r in g:
        assert _flt_eq(g[r], w[r]), (r, g[r], w[r])


# ---------------------------------------------------------------------- seams -----------
def test_gemm_abt_known_answer_and_bound(rng):
    from sieve_amd import abi
    a = np.array([[1, 2, 3], [4, 5, 6]], np.float32)
    b = np.array([[1, 0, 1], [0, 1, 0]], np.float32)
    c = np.zeros((2, 2), np.float32)
    abi.check(abi.lib().sieve_gemm_abt(abi.ptr(a), abi.ptr(b), 2, 2, 3, abi.ptr(c)))
    assert c.ravel().tolist() == [4.0, 2.0, 10.0, 5.0]
    for m, n, k in ((37, 53, 256), (5, 1000, 100), (128, 64, 7)):
        a = rng.standard_normal((m, k)).astype(np.float32)
        b = rng.standard_normal((n, k)).astype(np.float32)
        c = np.zeros((m, n), np.float32)
        abi.check(abi.lib().sieve_gemm_abt(abi.ptr(a), abi.ptr(b), m, n, k, abi.ptr(c)))
        c64 = a.astype(np.float64) @ b.astype(np.float64).T
        bound = k * 2.0 ** -23 * (np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64).T) + 1e-30
        assert np.all(np.abs(c - c64) <= bound)


@pytest.mark.parametrize("G", [64, 32])
def test_lut_scan_bit_exact(rng, G):
    from sieve_amd import abi
    q = rng.uniform(-1, 1, 4 * G).astype(np.float32)
    tables, _, _ = O.lut_new(q)
    nb = 1237
    blocks = rng.integers(0, 256, nb * G * 16, dtype=np.uint8)
    out = np.zeros(nb * 32, np.uint16)
    abi.check(abi.lib().sieve_lut_scan(abi.ptr(tables), G, abi.ptr(blocks), nb, abi.ptr(out)))
    assert np.array_equal(out, O.scan(tables, blocks))


def test_lut_scan_narrow_fixture():
    # lut.rs:212-228 fixture through the GPU seam.
    from sieve_amd import abi
    q = np.cos(np.arange(128, dtype=np.float32) * np.float32(0.91)).astype(np.float32)
    tables, _, _ = O.lut_new(q)
    codes = np.array([[(k * 37 + i * 11) % 256 for i in range(16)] for k in range(32)], np.uint8)
    block = O.pack_block(codes, 32)
    out = np.zeros(32, np.uint16)
    abi.check(abi.lib().sieve_lut_scan(abi.ptr(tables), 32, abi.ptr(block), 1, abi.ptr(out)))
    assert [int(x) for x in out] == [O.score_code(tables, c) for c in codes]


# ---------------------------------------------------------------------- route -----------
@pytest.mark.parametrize("cfg", [dict(nprobe=7), dict(nprobe=16), dict(nprobe=48), dict(nprobe=20, margin=0.05),
                                 dict(nprobe=20, budget=3000), dict(nprobe=1), dict(nprobe=1, fuse=1)])
def test_route_matches_oracle(small, oracle_index, gpu_index, cfg):
    cfg = dict(cfg)
    gpu_index.set_route_fuse(cfg.pop("fuse", -1))
    p = Params(cand=100, **cfg)
    s, c, n = gpu_index.route(small.q256, p)
    gpu_index.set_route_fuse(-1)
    for i in range(small.nq):
        ws, wc = oracle_index.route(small.q256[i], p)
        gs, gc = s[i, :n[i]], c[i, :n[i]]
        assert n[i] == len(wc), (i, n[i], len(wc))
        assert_same_up_to_ties(gc, gs, wc, ws)
        if len(set(ws.tolist())) == len(ws):
            assert np.array_equal(gc, wc) and _flt_eq(gs, ws)


@pytest.fixture(scope="module")
def wide():
    """An index of 6000 clusters over the small corpus (4 rows per cluster on average): routing
    selects over more than 4096 centroids, the register-held select with 64 keys per thread."""
    from conftest import SmallWorkload
    from oracle import oracle as O
    print("[gpu] building a 6000-cluster index", flush=True)
    w = SmallWorkload(k=6000, nq=24, tier2=False)
    return w, O.Index(w.host), Index.create(w.host)


@pytest.mark.parametrize("fuse", [-1, 1])
@pytest.mark.parametrize("cfg", [dict(nprobe=64), dict(nprobe=256, margin=0.05), dict(nprobe=9)])
def test_route_over_many_centroids_matches_oracle(wide, cfg, fuse):
    """fuse 1: the GEMM writes only tile maxima and the select recomputes its tiles' scores."""
    w, oix, gix = wide
    p = Params(cand=100, **cfg)
    gix.set_route_fuse(fuse)
    s, c, n = gix.route(w.q256, p)
    gix.set_route_fuse(-1)
    for i in range(w.nq):
        ws, wc = oix.route(w.q256[i], p)
        assert n[i] == len(wc)
        assert_same_up_to_ties(c[i, :n[i]], s[i, :n[i]], wc, ws)


@pytest.mark.parametrize("frac,target", [(0.3, 400), (0.02, 200), (0.3, 1), (0.01, 10**9)])
def test_expand_cut_over_many_centroids_matches_oracle(wide, frac, target):
    """Expand's probe list over 6000 centroids (24 per thread of the cut): the weighted
    selection finds the centroid, in (score desc, id asc) order, at which the allowed slots
    covered reach the target (one centroid for a target of 1; every centroid with an allowed
    slot for a target above the total). Results equal the oracle's up to ties."""
    from sieve_amd import parse_config
    from sieve_amd.index import Filter
    w, oix, gix = wide
    rng = np.random.default_rng(41)
    rows = np.sort(rng.choice(w.n, max(1, int(w.n * frac)), replace=False))
    allow = bitset(rows, w.n)
    f, of = Filter(gix, allow), O.Filter(oix, allow)
    _, p = parse_config("12:400:full", w.k, k=10)
    s, r, n, used = f.search(w.q768, w.q256, p, plan="expand", target=target)
    ws, wr, wn, wused = of.search(w.q768, w.q256, p, plan="expand", target=target)
    assert used == 1 and (wused == 1).all()  # Expand
    for i in range(w.nq):
        assert np.isin(r[i, :n[i]], rows).all()
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])


@pytest.fixture(scope="module")
def huge_k(small):
    """The small index with 20000 more, empty clusters (random unit centroids, some of them
    near the queries): routing selects over 20048 centroids, from the 1253 maxima of 16
    consecutive scores (the select over tile maxima with 16 maxima per thread)."""
    from oracle import oracle as O
    rng = np.random.default_rng(21)
    extra = 20000
    h = dict(small.host)
    c = rng.standard_normal((extra, 256)).astype(np.float32)
    c[:200] += 3.0 * small.q256[rng.integers(0, small.nq, 200)]  # some close to the queries
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    h["centroids"] = np.concatenate([h["centroids"], c.astype(np.float32)])
    h["block_off"] = np.concatenate([h["block_off"], np.full(extra, h["block_off"][-1], np.uint64)])
    return h, O.Index(h), Index.create(h)


@pytest.mark.parametrize("fuse", [-1, 1, 0])
@pytest.mark.parametrize("cfg", [dict(nprobe=32), dict(nprobe=200, margin=0.1), dict(nprobe=1)])
def test_route_over_20k_centroids_matches_oracle(small, huge_k, cfg, fuse):
    h, oix, gix = huge_k
    p = Params(cand=100, **cfg)
    gix.set_route_fuse(fuse)
    s, c, n = gix.route(small.q256, p)
    for i in range(small.nq):
        ws, wc = oix.route(small.q256[i], p)
        assert n[i] == len(wc)
        assert_same_up_to_ties(c[i, :n[i]], s[i, :n[i]], wc, ws)
    # the searches over it still equal the oracle's (empty clusters scan nothing)
    p = Params(nprobe=24, cand=300, k=10, rerank=Rerank.FULL)
    gs, gr, gn = gix.search(small.q768, small.q256, p)
    ws, wr, wn, _ = oix.search_batch(small.q768, small.q256, p, threads=4)
    gix.set_route_fuse(-1)
    for i in range(small.nq):
        assert_same_up_to_ties(gr[i, :gn[i]], gs[i, :gn[i]], wr[i, :wn[i]], ws[i, :wn[i]])


# ---------------------------------------------------------------------- candidates ------
def _oracle_probes(oracle_index, q256, p, nq):
    stride = p.nprobe
    ps = np.zeros((nq, stride), np.float32)
    pc = np.zeros((nq, stride), np.uint32)
    pn = np.zeros(nq, np.uint32)
    for i in range(nq):
        s, c = oracle_index.route(q256[i], p)
        ps[i, :len(s)], pc[i, :len(c)], pn[i] = s, c, len(c)
    return ps, pc, pn


def _check_candidates(got, want):
    (ge, gr, gx), (we, wr, wx) = got, want
    assert np.all(np.diff(gr.astype(np.int64)) > 0), "row order, deduplicated"
    if len(gr) == 0 or len(wr) == 0:
        assert len(gr) == len(wr)
        return
    b = max(ge.min(), we.min())
    g = {int(r): (e, x) for r, e, x in zip(gr, ge, gx) if e > b}
    w = {int(r): (e, x) for r, e, x in zip(wr, we, wx) if e > b}
    assert g.keys() == w.keys()
    for r in g:
        assert _flt_eq(g[r][0], w[r][0]) and _flt_eq(g[r][1], w[r][1]), r


@pytest.mark.parametrize("cand,nprobe", [(150, 6), (1000, 16), (37, 3), (2048, 48)])
def test_candidates_match_oracle(small, oracle_index, gpu_index, cand, nprobe):
    p = Params(nprobe=nprobe, cand=cand)
    ps, pc, pn = _oracle_probes(oracle_index, small.q256, p, small.nq)
    est, rows, errs, n = gpu_index.candidates(small.q256, ps, pc, pn, p)
    for i in range(small.nq):
        we, wr, wx, _ = oracle_index.candidates(small.q256[i], ps[i, :pn[i]], pc[i, :pn[i]], p)
        _check_candidates((est[i, :n[i]], rows[i, :n[i]], errs[i, :n[i]]), (we, wr, wx))


def test_candidates_tombstones_and_allowlists(small, oracle_index, gpu_index):
    p = Params(nprobe=8, cand=200)
    ps, pc, pn = _oracle_probes(oracle_index, small.q256, p, small.nq)
    rs = np.random.default_rng(3)
    dead = rs.choice(small.n, small.n // 5, replace=False)
    deleted = bitset(dead, small.n)
    allow = np.stack([bitset(rs.choice(small.n, small.n // 3, replace=False), small.n) for _ in range(small.nq)])
    flags = (np.arange(small.nq) % 3 != 0).astype(np.uint8)
    v = ViewSpec(deleted=deleted, allow=allow, allow_flags=flags)
    est, rows, errs, n = gpu_index.candidates(small.q256, ps, pc, pn, p, view=v)
    for i in range(small.nq):
        ov = O.make_view(deleted=deleted, allow=allow, allow_stride=allow.shape[1], allow_flags=flags)
        we, wr, wx, _ = oracle_index.candidates(small.q256[i], ps[i, :pn[i]], pc[i, :pn[i]], p, view=ov, qi=i)
        _check_candidates((est[i, :n[i]], rows[i, :n[i]], errs[i, :n[i]]), (we, wr, wx))
        assert not set(rows[i, :n[i]].tolist()) & set(dead.tolist())


# ---------------------------------------------------------------------- rerank ----------
@pytest.mark.parametrize("rr", [Rerank.FULL, Rerank.PREFIX, Rerank.NONE, Rerank.tiered(40), Rerank.ssd(40)])
@pytest.mark.parametrize("prune", [None, (10, 0.8)])
def test_rerank_matches_oracle(small, oracle_index, gpu_index, rr, prune):
    p = Params(nprobe=8, cand=300, k=10, rerank=rr, prune=prune)
    stride = p.cand
    ce = np.zeros((small.nq, stride), np.float32)
    cr = np.zeros((small.nq, stride), np.uint32)
    cx = np.zeros((small.nq, stride), np.float32)
    cn = np.zeros(small.nq, np.uint32)
    for i in range(small.nq):
        s, c = oracle_index.route(small.q256[i], p)
        e, r, x, _ = oracle_index.candidates(small.q256[i], s, c, p)
        ce[i, :len(e)], cr[i, :len(r)], cx[i, :len(x)], cn[i] = e, r, x, len(e)
    s, r, n = gpu_index.rerank(small.q768, small.q256, ce, cr, cx, cn, p)
    for i in range(small.nq):
        ws, wr, _ = oracle_index.rerank(small.q768[i], small.q256[i], ce[i, :cn[i]], cr[i, :cn[i]], cx[i, :cn[i]], p)
        assert n[i] == len(wr)
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr, ws)


def test_rerank_tiered_without_tier2(small):
    h = dict(small.host)
    for key in ("tier2_sigma", "tier2_codes", "tier2_norm"):
        h.pop(key)
    oix, gix = O.Index(h), Index.create(h)
    p = Params(nprobe=8, cand=300, k=10, rerank=Rerank.tiered(40))
    s, r, n = gix.search(small.q768, small.q256, p)
    ws, wr, wn, _ = oix.search_batch(small.q768, small.q256, p)
    for i in range(small.nq):
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])


# ---------------------------------------------------------------------- end to end ------
@pytest.mark.parametrize("cfg", ["8:300:full", "16:1000:full:b10x0.8", "32:500:tiered100", "48:200:prefix:p0.05",
                                 "24:400:ssd50:s4000", "4:50:none"])
def test_search_matches_oracle(small, oracle_index, gpu_index, cfg):
    from sieve_amd import parse_config
    _, p = parse_config(cfg, small.k, k=10)
    t = {}
    from sieve_amd import abi
    tm = abi.Timing()
    s, r, n = gpu_index.search(small.q768, small.q256, p, timing=tm)
    ws, wr, wn, wt = oracle_index.search_batch(small.q768, small.q256, p, threads=4)
    for i in range(small.nq):
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])
    assert tm.probed == wt["probed"] and tm.scanned == wt["scanned"]


@pytest.mark.parametrize("variant", list(range(10)))
def test_scan_variants_match_oracle(small, oracle_index, gpu_index, variant):
    from sieve_amd import abi
    p = Params(nprobe=16, cand=700, k=10, rerank=Rerank.FULL)
    gpu_index.set_scan_variant(variant)
    try:
        tm = abi.Timing()
        s, r, n = gpu_index.search(small.q768, small.q256, p, timing=tm)
    finally:
        gpu_index.set_scan_variant(0)
    ws, wr, wn, wt = oracle_index.search_batch(small.q768, small.q256, p, threads=4)
    for i in range(small.nq):
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])
    assert tm.scanned == wt["scanned"]


def test_autotune_picks_a_variant(small, gpu_index):
    p = Params(nprobe=16, cand=700, k=10)
    v, ms = gpu_index.autotune(small.q768, small.q256, p)
    from sieve_amd import abi
    n = abi.SCAN_VARIANTS
    assert 0 <= v < n and len(ms) == n and all(x > 0 for x in ms) and ms[v] == min(ms)
    gpu_index.set_scan_variant(0)


def test_device_arrays_match_host_arrays(small, gpu_index):
    import torch
    p = Params(nprobe=12, cand=400, k=10)
    s, r, n = gpu_index.search(small.q768, small.q256, p)
    dq7 = torch.tensor(small.q768, device="cuda")
    dq2 = torch.tensor(small.q256, device="cuda")
    ds, dr, dn = gpu_index.search(dq7, dq2, p)
    torch.cuda.synchronize()
    assert np.array_equal(s, ds.cpu().numpy()) and np.array_equal(r, dr.cpu().numpy().view(np.uint32))
    assert np.array_equal(n, dn.cpu().numpy().view(np.uint32))
    def dev(v):
        v = np.ascontiguousarray(v)
        if v.dtype == np.uint32:
            v = v.view(np.int32)
        elif v.dtype == np.uint64:
            v = v.view(np.int64)
        return torch.from_numpy(v).cuda()
    dix = Index.create({k: dev(v) for k, v in small.host.items()}, borrow_rerank=True)
    s2, r2, n2 = dix.search(small.q768, small.q256, p)
    assert np.array_equal(s, s2) and np.array_equal(r, r2)


def test_index_open_from_files(small, gpu_index, tmp_path):
    from sieve_amd import build
    build.write_index(small.arrays, str(tmp_path / "index"), str(tmp_path / "ingest"))
    fix = Index.open(str(tmp_path / "index"), str(tmp_path / "ingest"))
    assert fix.k == small.k and fix.n == small.n and fix.tier2
    p = Params(nprobe=10, cand=300, k=10, rerank=Rerank.tiered(60))
    a = gpu_index.search(small.q768, small.q256, p)
    b = fix.search(small.q768, small.q256, p)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_batch_composition_does_not_change_results(small, gpu_index):
    p = Params(nprobe=12, cand=500, k=10, rerank=Rerank.tiered(80), prune=(10, 1.0))
    s, r, n = gpu_index.search(small.q768, small.q256, p)
    for i in (0, 7, small.nq - 1):
        s1, r1, n1 = gpu_index.search(small.q768[i:i + 1], small.q256[i:i + 1], p)
        assert np.array_equal(r1[0], r[i]) and np.array_equal(s1[0], s[i]) and n1[0] == n[i]


@pytest.mark.parametrize("chunk", [1, 3, 40])
def test_scan_split_does_not_change_candidates(small, gpu_index, chunk):
    """Cutting each query's blocks into chunks scanned by separate workgroups and merged on the
    device gives the same candidates, bit for bit, as one workgroup per query."""
    p = Params(nprobe=12, cand=300, k=10, rerank=Rerank.FULL)
    ps, pc, pn = gpu_index.route(small.q256, p)
    try:
        gpu_index.set_scan_chunk(1 << 30)
        want = gpu_index.candidates(small.q256, ps, pc, pn, p)
        gpu_index.set_scan_chunk(chunk)
        got = gpu_index.candidates(small.q256, ps, pc, pn, p)
        got1 = gpu_index.candidates(small.q256[5:6], ps[5:6], pc[5:6], pn[5:6], p)
    finally:
        gpu_index.set_scan_chunk(0)
    n = want[3]
    assert np.array_equal(got[3], n) and got1[3][0] == n[5]
    for i in range(small.nq):
        for w, g in zip(want[:3], got[:3]):
            assert np.array_equal(w[i, :n[i]], g[i, :n[i]])
    for w, g in zip(want[:3], got1[:3]):
        assert np.array_equal(w[5, :n[5]], g[0, :n[5]])


def test_scan_split_with_growing_batches(small, gpu_index):
    """Split scans on one stream while the batch grows (its per-query counters are reallocated
    each time): every batch still gives the unsplit candidates."""
    p = Params(nprobe=12, cand=300, k=10, rerank=Rerank.FULL)
    ps, pc, pn = gpu_index.route(small.q256, p)
    try:
        gpu_index.set_scan_chunk(1 << 30)
        want = gpu_index.candidates(small.q256, ps, pc, pn, p)
        gpu_index.set_scan_chunk(3)
        for m in (1, 3, 9, 27, small.nq):
            got = gpu_index.candidates(small.q256[:m], ps[:m], pc[:m], pn[:m], p)
            assert np.array_equal(got[3], want[3][:m])
            for i in range(m):
                n = want[3][i]
                for w, g in zip(want[:3], got[:3]):
                    assert np.array_equal(w[i, :n], g[i, :n])
    finally:
        gpu_index.set_scan_chunk(0)


def test_errors_are_reported(small, gpu_index):
    from sieve_amd.abi import SieveError
    with pytest.raises(SieveError, match="cand"):
        gpu_index.search(small.q768, small.q256, Params(nprobe=4, cand=5000))
    with pytest.raises(SieveError, match="nprobe"):
        gpu_index.search(small.q768, small.q256, Params(nprobe=0, cand=50))
    bad = np.full((small.nq, 4), small.k + 5, np.uint32)
    with pytest.raises(SieveError, match="out of range"):
        gpu_index.candidates(small.q256, np.zeros((small.nq, 4), np.float32), bad, np.full(small.nq, 4, np.uint32),
                             Params(nprobe=4, cand=50))


# ---------------------------------------------------------------------- degenerate queries ---
@pytest.mark.parametrize("fuse", [-1, 1])
def test_zero_query_routes_exactly_and_leaves_the_batch_alone(small, oracle_index, gpu_index, fuse):
    """A zero q256 ties every centroid score, so its BF16 error band holds every centroid of a
    wide index: that query alone takes the exact route (all scores in f32, then the top n by
    score desc, id asc), and the other queries of its batch get their usual results. fuse 1:
    the same with the select recomputing scores from the BF16 operands."""
    from conftest import SmallWorkload
    w = SmallWorkload(k=6000, nq=12, tier2=False)
    gix = Index.create(w.host)
    gix.set_route_fuse(fuse)
    oix = O.Index(w.host)
    q256 = w.q256.copy()
    q768 = w.q768.copy()
    q256[3] = 0.0
    q768[3] = 0.0
    p = Params(nprobe=40, cand=200, k=10, rerank=Rerank.FULL)
    s, c, n = gix.route(q256, p)
    assert n[3] == 40 and np.all(s[3, :40] == 0.0)
    assert np.array_equal(c[3, :40], np.arange(40))  # ties broken by id asc
    for i in range(w.nq):
        ws, wc = oix.route(q256[i], p)
        assert n[i] == len(wc)
        assert_same_up_to_ties(c[i, :n[i]], s[i, :n[i]], wc, ws)
    gs, gr, gn = gix.search(q768, q256, p)
    alone = [gix.search(q768[i:i + 1], q256[i:i + 1], p) for i in range(w.nq)]
    for i in range(w.nq):
        assert np.array_equal(gr[i, :gn[i]], alone[i][1][0, :alone[i][2][0]])
        assert _flt_eq(gs[i, :gn[i]], alone[i][0][0, :alone[i][2][0]])
    ws, wr, wn, _ = oix.search_batch(q768, q256, p, threads=4)
    for i in range(w.nq):
        assert_same_up_to_ties(gr[i, :gn[i]], gs[i, :gn[i]], wr[i, :wn[i]], ws[i, :wn[i]])
