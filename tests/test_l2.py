"""The L2 path over 100-d int8 vectors (bigann::Index::search, src/bigann.rs:652-747): the
oracle against brute force (CPU), and the HIP path against the oracle (GPU)."""
import numpy as np
import pytest

from oracle import oracle as O

_cache = {}


@pytest.fixture(scope="module")
def l2small():
    if "w" not in _cache:
        import torch
        from sieve_amd.l2build import build_l2, int8_mixture
        base = int8_mixture(30000, 80, seed=5)
        arr = build_l2(base, 64, iters=6)
        h = arr.numpy()
        q = int8_mixture(48, 80, seed=5)  # same mixture, other draws: first rows differ by seed stream
        g = torch.Generator().manual_seed(99)
        q = torch.clamp(q.to(torch.float32) + torch.randint(-6, 7, q.shape, generator=g), -127, 127).to(torch.int8)
        _cache["w"] = (h, q.numpy())
    return _cache["w"]


@pytest.fixture(scope="module")
def l2oracle(l2small):
    h, _ = l2small
    return O.L2Index(h["leaves"], h["block_off"], h["blocks"], h["factors"], h["ids"], h["rotation"], h["base"])


def _ties_equal(g_ids, g_sc, w_ids, w_sc):
    g_sc, w_sc = np.asarray(g_sc, np.float32), np.asarray(w_sc, np.float32)
    if len(g_ids) == 0 or len(w_ids) == 0:
        return len(g_ids) == len(w_ids)
    b = max(g_sc.min(), w_sc.min())
    g = {int(i): s for i, s in zip(g_ids, g_sc) if s > b}
    w = {int(i): s for i, s in zip(w_ids, w_sc) if s > b}
    return g.keys() == w.keys() and all(np.float32(g[i]).view(np.uint32) == np.float32(w[i]).view(np.uint32)
                                        for i in g)


def test_oracle_l2_matches_brute_force(l2small, l2oracle):
    h, qs = l2small
    leaves = h["leaves"].reshape(-1, 100)
    bo = h["block_off"].astype(np.int64)
    fac = h["factors"].reshape(-1, 2)
    base = h["base"].reshape(-1, 100).astype(np.int64)
    for q in qs[:24]:
        probe, cids, cest, s, r = l2oracle.search(q, 8, 300, 10)
        qf = q.astype(np.float64)
        route = 2 * leaves.astype(np.float64) @ qf - (leaves.astype(np.float64) ** 2).sum(1)
        order = np.argsort(-route, kind="stable")
        if route[order[7]] - route[order[8]] > 1e-2:
            assert set(probe.tolist()) == set(order[:8].tolist())
        # candidates: every slot of the probed leaves, full sort of the estimate
        dots = O.gemm_abt(q.astype(np.float32)[None, :], leaves)[0]
        qp = np.zeros(128, np.float32)
        qp[:100] = q
        tables, _, _, a, b = O.query_new(h["rotation"], qp)
        ests, ids = [], []
        for c in probe:
            sc = O.scan(tables, h["blocks"][bo[c]:bo[c + 1]].ravel()).astype(np.float32)
            f = fac[bo[c] * 32:bo[c + 1] * 32]
            e = (np.float32(2.0) * (dots[c] + f[:, 0] * (np.float32(a) * sc + np.float32(b))) - f[:, 1])
            keep = h["ids"][bo[c] * 32:bo[c + 1] * 32] != 0xFFFFFFFF
            ests.append(e.astype(np.float32)[keep])
            ids.append(h["ids"][bo[c] * 32:bo[c + 1] * 32][keep])
        ests, ids = np.concatenate(ests), np.concatenate(ids)
        o = np.argsort(-ests, kind="stable")[:300]
        assert _ties_equal(cids, cest, ids[o], ests[o])
        # rerank: exact 2 <q, x> - |x|^2 over the candidates
        nx = {int(i): fac_nx for i, fac_nx in zip(h["ids"], fac[:, 1])}
        sc = np.array([np.float32(2.0) * np.float32(base[i] @ q.astype(np.int64)) - nx[int(i)] for i in cids],
                      np.float32)
        o2 = sorted(range(len(cids)), key=lambda j: (-sc[j], cids[j]))[:10]
        assert r.tolist() == [int(cids[j]) for j in o2]


def test_oracle_l2_batch_equals_single_queries(l2small, l2oracle):
    """The threaded batch form (the CPU baseline's) returns each query's single search exactly."""
    h, qs = l2small
    s, r, n = l2oracle.search_batch(qs, 8, 300, 10, threads=4)
    for i, q in enumerate(qs):
        _, _, _, ws, wr = l2oracle.search(q, 8, 300, 10)
        assert n[i] == len(wr) and np.array_equal(r[i, :n[i]], wr)
        assert np.array_equal(s[i, :n[i]].view(np.uint32), ws.view(np.uint32))
    lat = l2oracle.search_latency(qs[:5], 8, 300, 10)
    assert lat.shape == (5,) and np.all(lat > 0)


@pytest.mark.gpu
@pytest.mark.parametrize("nprobe,cand,margin,chunk", [(8, 300, 0.0, 0), (16, 1000, 0.0, 0), (32, 500, 0.3, 0),
                                                      (1, 50, 0.0, 0), (16, 300, 0.0, 2)])
def test_gpu_l2_matches_oracle(l2small, l2oracle, nprobe, cand, margin, chunk):
    from sieve_amd.index import L2Index
    h, qs = l2small
    ix = L2Index(h)
    ix.set_scan_chunk(chunk)
    s, r, n, probes, pn, cr, ce, cn = ix.search(qs, nprobe, cand, 10, margin=margin, details=True)
    for i, q in enumerate(qs):
        wp, wc, we, ws, wr = l2oracle.search(q, nprobe, cand, 10, margin=margin)
        assert pn[i] == len(wp)
        assert set(probes[i, :pn[i]].tolist()) == set(wp.tolist())
        assert _ties_equal(cr[i, :cn[i]], ce[i, :cn[i]], wc, we)
        assert _ties_equal(r[i, :n[i]], s[i, :n[i]], wr, ws)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", list(range(10)))
def test_gpu_l2_scan_variants_match_oracle(l2small, l2oracle, variant):
    from sieve_amd.index import L2Index
    h, qs = l2small
    ix = L2Index(h)
    ix.set_scan_variant(variant)
    s, r, n, probes, pn, cr, ce, cn = ix.search(qs, 16, 700, 10, details=True)
    for i, q in enumerate(qs):
        wp, wc, we, ws, wr = l2oracle.search(q, 16, 700, 10)
        assert _ties_equal(cr[i, :cn[i]], ce[i, :cn[i]], wc, we)
        assert _ties_equal(r[i, :n[i]], s[i, :n[i]], wr, ws)


@pytest.mark.gpu
@pytest.mark.parametrize("nq", [1, 3, 8])
@pytest.mark.parametrize("nprobe,cand", [(8, 300), (16, 1000), (1, 50), (32, 12)])
def test_gpu_l2_small_batches_equal_the_per_query_scan(l2small, l2oracle, nq, nprobe, cand):
    """Batches of at most 8 queries take the small-batch scan (per-chunk bounds, a threshold per
    query, appended lists, the merge): candidates and results equal the per-query scan's bit for
    bit, and the oracle's up to ties."""
    from sieve_amd.index import L2Index
    h, qs = l2small
    ix = L2Index(h)
    got = ix.search(qs[:nq], nprobe, cand, 10, details=True)
    ix.set_small_batch(0)
    want = ix.search(qs[:nq], nprobe, cand, 10, details=True)
    ws_, wr_, wn_, _, _, wcr, wce, wcn = want
    s, r, n, probes, pn, cr, ce, cn = got
    assert np.array_equal(wn_, n) and np.array_equal(wcn, cn)
    for i in range(nq):
        assert np.array_equal(wcr[i, :cn[i]], cr[i, :cn[i]])
        assert np.array_equal(wce[i, :cn[i]].view(np.uint32), ce[i, :cn[i]].view(np.uint32))
        assert np.array_equal(wr_[i, :n[i]], r[i, :n[i]])
        assert np.array_equal(ws_[i, :n[i]].view(np.uint32), s[i, :n[i]].view(np.uint32))
        wp, wc, we, ws, wr = l2oracle.search(qs[i], nprobe, cand, 10)
        assert _ties_equal(cr[i, :cn[i]], ce[i, :cn[i]], wc, we)
        assert _ties_equal(r[i, :n[i]], s[i, :n[i]], wr, ws)
