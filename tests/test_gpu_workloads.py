"""Parity of the HIP path with the oracle at the benchmark workloads' own shapes.

* 1M x 768-d, 1024 clusters, SOAR, RaBitQ, nprobe 32, 1000 candidates, 1000 queries: every
  query's probe list, candidates and top 10 against the oracle, through the per-query, the
  cluster-major and the small-batch scans.
* 18.4M x 768-d, 8192 clusters, 10,000-query batches with margin probing (uneven probe counts)
  and a bound-cut rerank: the whole batch on the device, 500 of its queries against the oracle
  (probe lists, candidates, top 10), and the cluster-major scan against the per-query scan on
  all 10,000 bit for bit.
* The 1B config's routing shape: 131,072 leaves (512 coarse cells x 256) over 16M x 100-d int8
  rows, L2: probe lists, candidates and top 10 against the oracle (the hi|lo BF16 L2 GEMM and the
  select over 8192 tile maxima per query).
* The serving config's semantics at 1M rows: margin probing, allowlist filters of log-uniform
  selectivity, interleaved live inserts, deletes and upserts through the server, every result
  against the oracle's Live searched in submission order.

Probe and candidate lists are compared up to ties at the selection boundary with bit-identical
scores, top-10 lists the same way (the reference breaks ties with select_nth_unstable).
Indexes are built on the GPU with the package's builder; the oracle reads the same bytes.
"""
import gc
import os

import numpy as np
import pytest

from oracle import oracle as O
from sieve_amd import Params, Rerank, parse_config

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def say(msg):
    print(f"[gpu-workloads] {msg}", flush=True)


def _flt_eq(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def _same_up_to_ties(g_ids, g_sc, w_ids, w_sc):
    """Every entry above the lowest kept score of either list is in both, with a bit-identical
    score (lists best first)."""
    g_sc, w_sc = np.asarray(g_sc, np.float32), np.asarray(w_sc, np.float32)
    if len(g_ids) == 0 or len(w_ids) == 0:
        return len(g_ids) == len(w_ids)
    b = max(g_sc.min(), w_sc.min())
    g = {int(i): s for i, s in zip(g_ids, g_sc) if s > b}
    w = {int(i): s for i, s in zip(w_ids, w_sc) if s > b}
    return g.keys() == w.keys() and all(_flt_eq(g[i], w[i]) for i in g)


def _cands_up_to_ties(got, want):
    (ge, gr, gx), (we, wr, wx) = got, want
    if not np.all(np.diff(gr.astype(np.int64)) > 0):
        return False
    if len(gr) == 0 or len(wr) == 0:
        return len(gr) == len(wr)
    b = max(ge.min(), we.min())
    g = {int(r): (e, x) for r, e, x in zip(gr, ge, gx) if e > b}
    w = {int(r): (e, x) for r, e, x in zip(wr, we, wx) if e > b}
    return g.keys() == w.keys() and all(_flt_eq(g[r][0], w[r][0]) and _flt_eq(g[r][1], w[r][1]) for r in g)


def _build_ip(rows, components, clusters, nq, tier2=False, spread=0.1, seed=20241203):
    import torch
    from sieve_amd import build as B, synth
    from sieve_amd.index import Index
    say(f"building {rows} x 768 over {clusters} clusters on the GPU")
    mix = synth.Mixture(synth.MixtureSpec(components=components, seed=seed, spread=spread), device="cuda")
    corpus = B.ingest(mix, rows, batch=1 << 20)
    arrays = B.build(corpus, clusters, sample_n=max(400_000, 64 * clusters), iters=12, chunk=1 << 19, tier2=tier2)
    q768, q256 = mix.queries(nq)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    gix = Index.create(arrays.device(), borrow_rerank=True)
    say("copying the index to host memory for the oracle")
    host = arrays.numpy()
    oix = O.Index(host)
    return dict(arrays=arrays, mix=mix, gix=gix, oix=oix, host=host, q768=q768, q256=q256,
                q7h=q768.cpu().numpy(), q2h=q256.cpu().numpy())


def _release(w):
    import torch
    w.clear()
    gc.collect()
    torch.cuda.empty_cache()


def _check_route(gix, oix, q2h, p, which):
    s, c, n = gix.route(q2h[which], p)
    bad = 0
    for j, i in enumerate(which):
        ws, wc = oix.route(q2h[i], p)
        if n[j] != len(wc) or not _same_up_to_ties(c[j, :n[j]], s[j, :n[j]], wc, ws):
            bad += 1
    return bad, n


def _check_candidates(gix, oix, q2h, p, which):
    # the oracle's probe lists, so that candidates are compared on the same probes
    stride = p.nprobe
    ps = np.zeros((len(which), stride), np.float32)
    pc = np.zeros((len(which), stride), np.uint32)
    pn = np.zeros(len(which), np.uint32)
    for j, i in enumerate(which):
        s, c = oix.route(q2h[i], p)
        ps[j, :len(s)], pc[j, :len(c)], pn[j] = s, c, len(c)
    est, rows, errs, n = gix.candidates(q2h[which], ps, pc, pn, p)
    bad = 0
    for j, i in enumerate(which):
        we, wr, wx, _ = oix.candidates(q2h[i], ps[j, :pn[j]], pc[j, :pn[j]], p)
        if not _cands_up_to_ties((est[j, :n[j]], rows[j, :n[j]], errs[j, :n[j]]), (we, wr, wx)):
            bad += 1
    return bad


def _check_topk(g, w, which):
    (gs, gr, gn), (ws, wr, wn) = g, w
    bad = 0
    for j, i in enumerate(which):
        if not _same_up_to_ties(gr[i, :gn[i]], gs[i, :gn[i]], wr[j, :wn[j]], ws[j, :wn[j]]):
            bad += 1
    return bad


@pytest.fixture(scope="module")
def c1():
    w = _build_ip(1_000_000, 1024, 1024, 1000)
    yield w
    _release(w)


@pytest.mark.timeout(600)
def test_c1_full_size_matches_oracle(c1):
    """1M x 768, 1024 clusters, 32:1000:full, 1000 queries."""
    gix, oix, q7h, q2h = c1["gix"], c1["oix"], c1["q7h"], c1["q2h"]
    _, p = parse_config("32:1000:full", gix.k, k=10)
    allq = np.arange(len(q2h))
    bad, n = _check_route(gix, oix, q2h, p, allq)
    say(f"c1 route: {bad} of {len(allq)} probe lists differ")
    assert bad == 0 and np.all(n == 32)
    bad = _check_candidates(gix, oix, q2h, p, allq[:300])
    say(f"c1 candidates: {bad} of 300 differ")
    assert bad == 0
    want = oix.search_batch(q7h, q2h, p, threads=THREADS)[:3]
    for variant in (2, 10):  # per-query scan, cluster-major scan
        gix.set_scan_variant(variant)
        got = gix.search(q7h, q2h, p)
        bad = _check_topk(got, want, allq)
        say(f"c1 search, scan variant {variant}: {bad} of {len(allq)} top-10 lists differ")
        assert bad == 0
    # the small-batch scan (batches of up to 8 queries under the cluster-major variants)
    for i0 in (0, 500, 993):
        got = gix.search(q7h[i0:i0 + 7], q2h[i0:i0 + 7], p)
        ws, wr, wn = want
        assert _check_topk(got, (ws[i0:i0 + 7], wr[i0:i0 + 7], wn[i0:i0 + 7]), np.arange(7)) == 0
    gix.set_scan_variant(0)


def test_c1_shape_graphs_and_device_buffers(c1):
    """The same batch from device buffers through captured HIP graphs equals the eager call."""
    import torch
    from sieve_amd.index import Stream
    gix = c1["gix"]
    _, p = parse_config("32:1000:full", gix.k, k=10)
    gix.set_scan_variant(10)
    want = [x.cpu().numpy() for x in gix.search(c1["q768"], c1["q256"], p)]
    st = Stream()
    st.set_graphs(True)
    out = (torch.empty(len(c1["q7h"]), 10, dtype=torch.float32, device="cuda"),
           torch.empty(len(c1["q7h"]), 10, dtype=torch.int32, device="cuda"),
           torch.empty(len(c1["q7h"]), dtype=torch.int32, device="cuda"))
    for rep in range(4):
        gix.search(c1["q768"], c1["q256"], p, stream=st, out=out)
        st.sync()
        got = [x.cpu().numpy() for x in out]
        for g, w in zip(got, want):
            assert np.array_equal(g.view(np.uint32), w.view(np.uint32)), rep
    assert st.last_stats()["graph"] == 2
    gix.set_scan_variant(0)


@pytest.mark.timeout(600)
def test_c5_semantics_at_1m_rows(c1):
    """The serving config's request mix at 1M rows through the server: margin probing, Post
    filters of selectivity 0.1% .. 100% (log-uniform), live inserts, deletes and upserts
    interleaved with the searches. A write is a barrier in submission order, so every result
    equals the oracle's Live searched after the same writes."""
    import torch
    from sieve_amd.build import normalize_rows
    from sieve_amd.index import Filter, Live, bitset
    from sieve_amd.server import Server
    gix, oix, q7h, q2h, mix = c1["gix"], c1["oix"], c1["q7h"], c1["q2h"], c1["mix"]
    n = gix.n
    _, p = parse_config("64:500:full:b10x0.8:p0.1", gix.k, k=10)
    rng = np.random.default_rng(5)
    sel = [0.001, 0.01, 0.1, 1.0]
    allows = [bitset(np.sort(rng.choice(n, max(1, int(x * n)), replace=False)), n) for x in sel]
    filts = [Filter(gix, a) for a in allows]
    oviews = [O.make_view(allow=a) for a in allows]
    live = Live(gix, max_rows=20_000)
    olive = O.Live(oix, soar_lambda=1.0)
    fresh, _ = normalize_rows(mix.rows(n + 10_000, 3000).to("cuda"))
    fresh = fresh.cpu().numpy().astype(np.float32)
    pre = normalize_rows(torch.from_numpy(fresh[:, :256].copy()))[0].numpy()
    # queries: 60 from the query set, 20 at fresh rows (the inserts change their results)
    q7 = np.concatenate([q7h[:60], fresh[:20]])
    q2 = np.concatenate([q2h[:60], pre[:20]])
    sv = Server(gix, p, max_batch=64, max_inflight=3, live=live)
    phases = []  # (tickets, expected)

    def searches():
        tk = []
        for i in range(len(q7)):
            tk.append((None, i, sv.submit(q7[i], q2[i])))
            if i % 2 == 0:
                f = (i // 2) % len(sel)
                tk.append((f, i, sv.submit_filtered(q7[i], q2[i], filts[f], plan="post")))
        exp = {None: olive.search_batch(q7, q2, p, threads=THREADS)}
        for f in range(len(sel)):
            exp[f] = olive.search_batch(q7, q2, p, view=oviews[f], threads=THREADS)
        phases.append((tk, exp))

    writes = []
    ext = (3_000_000_000 + np.arange(3000)).astype(np.uint32)
    searches()
    writes.append(sv.insert(ext[:1000], fresh[:1000]))
    olive.insert(ext[:1000], fresh[:1000])
    searches()
    dels = np.concatenate([ext[:100], rng.choice(n, 100, replace=False)]).astype(np.uint32)
    writes.append(sv.delete(dels))
    olive.delete(dels)
    searches()
    writes.append(sv.insert(ext[500:2500], fresh[500:2500]))  # 500 upserts, 1500 new rows
    olive.insert(ext[500:2500], fresh[500:2500])
    searches()
    bad, probes = 0, []
    for tk, exp in phases:
        for f, i, t in tk:
            gs, gr, _ = sv.wait(t)
            ws, wr, wn = exp[f][:3]
            if not _same_up_to_ties(gr, gs, wr[i, :wn[i]], ws[i, :wn[i]]):
                bad += 1
    for w in writes:
        sv.wait(w)
    total = sum(len(tk) for tk, _ in phases)
    say(f"c5 at 1M rows: {bad} of {total} results differ from the oracle's Live")
    _, _, pn = gix.route(q2h[:200], p)
    say(f"c5 margin probing: probes per query {int(pn.min())} .. {int(pn.max())}")
    assert bad == 0
    assert pn.min() < pn.max()
    sv.close()
    live.close()


@pytest.fixture(scope="module")
def c2():
    w = _build_ip(18_400_000, 8192, 8192, 10_000)
    yield w
    _release(w)


@pytest.mark.timeout(600)
def test_c2_full_size_matches_oracle_on_500_queries(c2):
    """18.4M x 768, 8192 clusters, 10,000 queries, margin probing up to 128 probes, 500
    candidates and a bound-cut rerank."""
    gix, oix, q7h, q2h = c2["gix"], c2["oix"], c2["q7h"], c2["q2h"]
    _, p = parse_config("128:500:full:b10x0.8:p0.1", gix.k, k=10)
    sub = np.arange(500)
    bad, n = _check_route(gix, oix, q2h, p, sub)
    say(f"c2 route: {bad} of 500 probe lists differ; probes per query {int(n.min())} .. {int(n.max())}")
    assert bad == 0 and n.min() < n.max()
    bad = _check_candidates(gix, oix, q2h, p, sub[:100])
    say(f"c2 candidates: {bad} of 100 differ")
    assert bad == 0
    want = oix.search_batch(q7h[sub], q2h[sub], p, threads=THREADS)[:3]
    gix.set_scan_variant(10)
    got_cm = [x.cpu().numpy() for x in gix.search(c2["q768"], c2["q256"], p)]
    gix.set_scan_variant(2)
    got_pq = [x.cpu().numpy() for x in gix.search(c2["q768"], c2["q256"], p)]
    gix.set_scan_variant(0)
    for a, b in zip(got_cm, got_pq):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    bad = _check_topk((got_cm[0], got_cm[1].view(np.uint32), got_cm[2].view(np.uint32)), want, sub)
    say(f"c2 search of the 10,000-query batch: {bad} of 500 top-10 lists differ from the oracle")
    assert bad == 0


@pytest.fixture(scope="module")
def c4shape():
    import torch
    from sieve_amd.index import L2Index
    from sieve_amd.l2build import L2Mixture, build_l2_large
    n = 16_000_000
    say(f"building the L2 index: {n} x 100 int8, 512 x 256 leaves")
    mix = L2Mixture(131072, seed=20241204, topics=16)
    base = torch.empty(n, 100, dtype=torch.int8, device="cuda")
    mix.fill(base)
    q = mix.queries(2000)
    arr = build_l2_large(base, coarse=512, fine=256)
    gix = L2Index(dict(leaves=arr.leaves, block_off=arr.block_off, blocks=arr.blocks, ids=arr.ids,
                       factors=arr.factors, rotation=arr.rotation, base=base), borrow_base=True)
    h = arr.numpy()
    oix = O.L2Index(h["leaves"], h["block_off"], h["blocks"], h["factors"], h["ids"], h["rotation"], h["base"])
    w = dict(gix=gix, oix=oix, q=q, qh=q.cpu().numpy(), arr=arr, base=base)
    yield w
    _release(w)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("nprobe,cand,fuse", [(8, 1000, -1), (64, 2000, -1), (64, 2000, 1), (8, 1000, 0)])
def test_c4_routing_shape_matches_oracle(c4shape, nprobe, cand, fuse):
    """fuse -1 takes the fused route at 8 probes and the unfused one at 64; 1 and 0 the others."""
    gix, oix, qh = c4shape["gix"], c4shape["oix"], c4shape["qh"]
    assert gix.k == 131072
    gix.set_route_fuse(fuse)
    s, r, n, probes, pn, cr, ce, cn = gix.search(c4shape["q"], nprobe, cand, 10, details=True)
    s, r, n = s.cpu().numpy(), r.cpu().numpy().view(np.uint32), n.cpu().numpy().view(np.uint32)
    probes, pn = probes.cpu().numpy().view(np.uint32), pn.cpu().numpy().view(np.uint32)
    cr, ce, cn = cr.cpu().numpy().view(np.uint32), ce.cpu().numpy(), cn.cpu().numpy().view(np.uint32)
    bad_p = bad_c = bad_k = 0
    m = 120
    for i in range(m):
        wp, wci, wce, ws, wr = oix.search(qh[i], nprobe, cand, 10)
        if not np.array_equal(np.sort(probes[i, :pn[i]]), np.sort(wp)):
            bad_p += 1
        if not _same_up_to_ties(cr[i, :cn[i]], ce[i, :cn[i]], wci, wce):
            bad_c += 1
        if not _same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr, ws):
            bad_k += 1
    say(f"c4 shape {nprobe}:{cand}: of {m} queries, {bad_p} probe sets, {bad_c} candidate lists and {bad_k} "
        f"top-10 lists differ")
    assert bad_p == 0 and bad_c == 0 and bad_k == 0
    # single queries through the small-batch scan give the batch's results
    for i in (0, 77):
        s1, r1, n1 = gix.search(qh[i:i + 1], nprobe, cand, 10)
        assert np.array_equal(r1[0, :n1[0]], r[i, :n[i]]) and _flt_eq(s1[0, :n1[0]], s[i, :n[i]])
    gix.set_route_fuse(-1)


@pytest.mark.timeout(600)
def test_c4_shape_refined_search_matches_oracle(c4shape):
    """The reference's best 1B operating point, 2-bit refinement with a 1000-slot shortlist, 100
    candidates and margin probing (128:100:1000:m0.4), on the 131,072-leaf index."""
    from sieve_amd.index import L2Index
    from sieve_amd.l2build import refine_section
    arr, qh, oix = c4shape["arr"], c4shape["qh"], c4shape["oix"]
    ext = refine_section(arr, 2)
    gix = L2Index(dict(leaves=arr.leaves, block_off=arr.block_off, blocks=arr.blocks, ids=arr.ids,
                       factors=arr.factors, rotation=arr.rotation, base=arr.base, refine_bits=2, refine=ext),
                  borrow_base=True)
    oix.set_refine(2, ext.cpu().numpy())
    s, r, n, probes, pn, cr, ce, cn = gix.search(c4shape["q"], 128, 100, 10, margin=0.4, details=True, refine=1000)
    s, r, n = s.cpu().numpy(), r.cpu().numpy().view(np.uint32), n.cpu().numpy().view(np.uint32)
    cr, ce, cn = cr.cpu().numpy().view(np.uint32), ce.cpu().numpy(), cn.cpu().numpy().view(np.uint32)
    pn = pn.cpu().numpy().view(np.uint32)
    bad_c = bad_k = 0
    m = 60
    for i in range(m):
        wp, wci, wce, ws, wr = oix.search(qh[i], 128, 100, 10, margin=0.4, refine=1000)
        assert pn[i] == len(wp)
        if not _same_up_to_ties(cr[i, :cn[i]], ce[i, :cn[i]], wci, wce):
            bad_c += 1
        if not _same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr, ws):
            bad_k += 1
    say(f"c4 shape refined 128:100:1000:m0.4: probes per query {int(pn.min())} .. {int(pn.max())}; of {m} "
        f"queries, {bad_c} candidate lists and {bad_k} top-10 lists differ")
    assert bad_c == 0 and bad_k == 0
