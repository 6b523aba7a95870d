"""The continuous-batching server: every request's result is the same alone, mixed into
batches with others, and from the plain batched entry point (bit for bit), and an open-loop
run completes every request."""
import threading

import numpy as np
import pytest

from sieve_amd import Params, Rerank
from sieve_amd.index import Index
from sieve_amd.server import Server, open_loop

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(small):
    ix = Index.create(small.host)
    p = Params(nprobe=12, cand=400, k=10, rerank=Rerank.tiered(60), prune=(10, 1.0), margin=0.08)
    want = ix.search(small.q768, small.q256, p)
    return ix, p, want


def test_alone_and_mixed_results_match_batch(small, setup):
    ix, p, (ws, wr, wn) = setup
    sv = Server(ix, p, max_batch=8, max_inflight=3)
    for i in range(small.nq):  # alone: one request in flight at a time
        s, r, info = sv.wait(sv.submit(small.q768[i], small.q256[i]))
        assert info.batch_size == 1
        assert np.array_equal(r, wr[i, :wn[i]]) and np.array_equal(s, ws[i, :wn[i]])
    tickets = [sv.submit(small.q768[i], small.q256[i]) for i in range(small.nq)]  # mixed
    sizes = []
    for i, t in enumerate(tickets):
        s, r, info = sv.wait(t)
        sizes.append(info.batch_size)
        assert np.array_equal(r, wr[i, :wn[i]]) and np.array_equal(s, ws[i, :wn[i]])
    assert max(sizes) > 1
    st = sv.stats()
    assert st["queries"] == 2 * small.nq and st["slots"] == 3
    sv.close()


def test_concurrent_submitters(small, setup):
    ix, p, (ws, wr, wn) = setup
    sv = Server(ix, p, max_batch=16, max_inflight=2)
    errors = []

    def worker(off):
        try:
            for rep in range(3):
                for i in range(off, small.nq, 4):
                    s, r, _ = sv.wait(sv.submit(small.q768[i], small.q256[i]))
                    if not np.array_equal(r, wr[i, :wn[i]]):
                        errors.append(i)
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(o,)) for o in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    sv.close()
    assert not errors


def test_open_loop_completes(small, setup):
    ix, p, _ = setup
    sv = Server(ix, p, max_batch=32)
    res = open_loop(sv, small.q768, small.q256, rate=2000, seconds=1.0)
    sv.close()
    assert res["requests"] > 500 and res["p99_ms"] >= res["p50_ms"] > 0


def test_filters_and_live_writes_in_the_serving_loop(small, oracle_index):
    """Searches, filtered searches, inserts and deletes through one server: a write is a barrier
    in submission order, so every result equals the oracle's Live searched in that order."""
    import numpy as np
    import torch
    from oracle import oracle as O
    from sieve_amd.build import normalize_rows
    from sieve_amd.index import Filter, Live, bitset
    from test_gpu_parity import assert_same_up_to_ties

    ix = Index.create(small.host)
    live = Live(ix, max_rows=1000)
    olive = O.Live(oracle_index, soar_lambda=1.0)
    p = Params(nprobe=12, cand=400, k=10, rerank=Rerank.FULL)
    rng = np.random.default_rng(21)
    allow_rows = np.sort(rng.choice(small.n, small.n // 4, replace=False))
    allow = bitset(allow_rows, small.n)
    filt = Filter(ix, allow)
    oview = O.make_view(allow=allow)
    new, _ = normalize_rows(small.mix.rows(small.n + 500, 200))
    new = new.numpy().astype(np.float32)
    # queries near the new rows, so the writes change results
    q768 = np.concatenate([small.q768[:12], new[:8]])
    q256 = np.concatenate([small.q256[:12], normalize_rows(torch.from_numpy(new[:8, :256].copy()))[0].numpy()])
    sv = Server(ix, p, max_batch=8, max_inflight=3, live=live)
    expected, tickets = [], []

    def searches():
        for i in range(len(q768)):
            tickets.append(("plain", i, sv.submit(q768[i], q256[i])))
            if i % 3 == 0:
                tickets.append(("post", i, sv.submit_filtered(q768[i], q256[i], filt, plan="post")))
        s, r, n = olive.search_batch(q768, q256, p, threads=4)
        fs, fr, fn = olive.search_batch(q768, q256, p, view=oview, threads=4)
        expected.append(((s, r, n), (fs, fr, fn)))

    ext1 = (80_000_000 + np.arange(150)).astype(np.uint32)
    searches()
    n0 = len(tickets)
    w1 = sv.insert(ext1, new[:150])
    olive.insert(ext1, new[:150])
    searches()
    n1 = len(tickets)
    dels = np.concatenate([ext1[:40], allow_rows[:30]]).astype(np.uint32)
    w2 = sv.delete(dels)
    olive.delete(dels)
    searches()
    w3 = sv.insert(ext1[:5], new[150:155])  # upserts
    olive.insert(ext1[:5], new[150:155])
    searches()
    # phases: tickets [0, n0) before any write, [n0, n1) after w1, then after w2, after w3
    per_phase = (len(tickets) - n1) // 2
    check_ranges = [(0, 0, n0), (1, n0, n1), (2, n1, n1 + per_phase), (3, n1 + per_phase, len(tickets))]
    for phase, a, b in check_ranges:
        (s, r, n), (fs, fr, fn) = expected[phase]
        for kind, i, t in tickets[a:b]:
            gs, gr, _ = sv.wait(t)
            ws, wr, wn = (s, r, n) if kind == "plain" else (fs, fr, fn)
            assert_same_up_to_ties(gr, gs, wr[i, :wn[i]], ws[i, :wn[i]])
    for w in (w1, w2, w3):
        sv.wait(w)
    assert sv.stats()["writes"] == 3
    sv.close()
    live.close()


def test_writes_overlap_the_batches_in_flight(small):
    """Inserts, deletes and upserts submitted back to back with searches, nothing waited for in
    between: each write is encoded while earlier batches run and published after them on the
    device, and later batches wait for the publish there. Every search returns, bit for bit,
    what the same operations done one at a time through the synchronous live API return."""
    import numpy as np
    import torch
    from sieve_amd.build import normalize_rows
    from sieve_amd.index import Live

    p = Params(nprobe=12, cand=400, k=10, rerank=Rerank.FULL)
    new, _ = normalize_rows(small.mix.rows(small.n + 900, 400))
    new = new.numpy().astype(np.float32)
    q768 = np.concatenate([small.q768[:16], new[:16]])
    q256 = np.concatenate([small.q256[:16], normalize_rows(torch.from_numpy(new[:16, :256].copy()))[0].numpy()])
    ops = []
    for r in range(8):
        ops.append(("insert", (70_000_000 + 40 * r + np.arange(40)).astype(np.uint32), new[40 * r:40 * r + 40]))
        if r % 2 == 1:
            ops.append(("delete", (70_000_000 + 40 * (r - 1) + np.arange(12)).astype(np.uint32), None))
        if r == 5:  # upserts of live rows
            ops.append(("insert", (70_000_000 + 40 * r + np.arange(6)).astype(np.uint32), new[:6]))

    ix_a = Index.create(small.host)
    live_a = Live(ix_a, max_rows=1000)
    sv = Server(ix_a, p, max_batch=8, max_inflight=3, live=live_a)
    phases, writes = [[sv.submit(q768[i], q256[i]) for i in range(len(q768))]], []
    for kind, ext, v in ops:
        writes.append(sv.insert(ext, v) if kind == "insert" else sv.delete(ext))
        phases.append([sv.submit(q768[i], q256[i]) for i in range(len(q768))])
    got = [[sv.wait(t)[:2] for t in ts] for ts in phases]
    for w in writes:
        sv.wait(w)
    assert sv.stats()["writes"] == len(ops)
    sv.close()
    info_a = live_a.info()

    ix_b = Index.create(small.host)
    live_b = Live(ix_b, max_rows=1000)
    want = [live_b.search(q768, q256, p)]
    for kind, ext, v in ops:
        if kind == "insert":
            live_b.insert(ext, v)
        else:
            live_b.delete(ext)
        want.append(live_b.search(q768, q256, p))
    info_b = live_b.info()
    for key in ("rows", "slots", "arena_blocks", "deleted"):  # exact block count once settled
        assert info_a[key] == info_b[key], key
    for phase, (g, (ws, wr, wn)) in enumerate(zip(got, want)):
        for i, (gs, gr) in enumerate(g):
            assert np.array_equal(gr, wr[i, :wn[i]]), (phase, i)
            assert np.array_equal(gs.view(np.uint32), ws[i, :wn[i]].view(np.uint32)), (phase, i)
    for x in (live_a, live_b):
        x.close()


def test_expand_over_listed_slots_with_live_writes(small):
    """Expand-filtered searches through the server with a live state: the base clusters' listed
    allowed slots plus the delta lists (tombstones applied) give the same results bit for bit as
    scanning the probed clusters whole, submitted between the same writes."""
    import numpy as np
    import torch
    from sieve_amd.build import normalize_rows
    from sieve_amd.index import Filter, Live, bitset

    ix = Index.create(small.host)
    live = Live(ix, max_rows=1000)
    p = Params(nprobe=12, cand=400, k=10, rerank=Rerank.FULL)
    rng = np.random.default_rng(23)
    allow_rows = np.sort(rng.choice(small.n + 1000, (small.n + 1000) // 20, replace=False))
    allow = bitset(allow_rows, small.n + 1000)
    listed, whole = Filter(ix, allow), Filter(ix, allow)
    whole.set_gather(False)
    new, _ = normalize_rows(small.mix.rows(small.n + 700, 200))
    new = new.numpy().astype(np.float32)
    q768 = np.concatenate([small.q768[:10], new[:6]])
    q256 = np.concatenate([small.q256[:10], normalize_rows(torch.from_numpy(new[:6, :256].copy()))[0].numpy()])
    sv = Server(ix, p, max_batch=8, max_inflight=3, live=live)
    pairs = []

    def searches():
        for i in range(len(q768)):
            pairs.append((sv.submit_filtered(q768[i], q256[i], listed, plan="expand"),
                          sv.submit_filtered(q768[i], q256[i], whole, plan="expand")))

    searches()
    sv.insert((90_000_000 + np.arange(150)).astype(np.uint32), new[:150])
    searches()
    sv.delete(np.concatenate([np.arange(90_000_000, 90_000_040), allow_rows[:30]]).astype(np.uint32))
    searches()
    nonempty = 0
    for a, b in pairs:
        ga, gb = sv.wait(a), sv.wait(b)
        assert np.array_equal(ga[1], gb[1]) and np.array_equal(np.asarray(ga[0]).view(np.uint32),
                                                               np.asarray(gb[0]).view(np.uint32))
        nonempty += len(ga[1]) > 0
    assert nonempty == len(pairs)
    sv.close()
    live.close()


def test_zero_query_in_a_batch_does_not_fail_the_others(small):
    """A degenerate request (zero vectors: every centroid ties, so its routing takes the exact
    path) batched with normal requests: every request succeeds, and each normal one gets the
    result it gets alone."""
    from conftest import SmallWorkload
    w = SmallWorkload(k=6000, nq=10, tier2=False)
    ix = Index.create(w.host)
    p = Params(nprobe=40, cand=200, k=10, rerank=Rerank.FULL)
    q768, q256 = w.q768.copy(), w.q256.copy()
    q768[4] = 0.0
    q256[4] = 0.0
    alone = [ix.search(q768[i:i + 1], q256[i:i + 1], p) for i in range(w.nq)]
    sv = Server(ix, p, max_batch=16, max_inflight=1)
    tickets = [sv.submit(q768[i], q256[i]) for i in range(w.nq)]
    for i, t in enumerate(tickets):
        s, r, info = sv.wait(t)
        ws, wr, wn = alone[i]
        assert np.array_equal(r, wr[0, :wn[0]]) and np.array_equal(s, ws[0, :wn[0]])
    sv.close()


def test_server_reserves_scratch_for_every_path(small, setup):
    """The slot footprint counts the worst-case scan scratch: larger than the inputs, outputs,
    routing scores and candidates alone, and a mixed run (batch sizes from 1 up, so every scan
    path) allocates nothing more on the device than the slots hold after creation."""
    import torch
    _, p, _ = setup
    ix = Index.create(small.host)
    ix.set_scan_variant(10)  # batches of 1-8: small-batch scan, 9-63: per-query, 64+: cluster-major
    sv = Server(ix, p, max_batch=96, max_inflight=2)
    st = sv.stats()
    base = 96 * ((768 + 256) * 4 + p.k * 8 + 4) + 96 * ix.k * 4 + 96 * p.cand * 12
    assert st["slot_bytes"] > base
    free0 = torch.cuda.mem_get_info()[0]
    for m in (1, 5, 9, 40, 96):
        ts = [sv.submit(small.q768[i % small.nq], small.q256[i % small.nq]) for i in range(m)]
        for t in ts:
            sv.wait(t)
    free1 = torch.cuda.mem_get_info()[0]
    sv.close()
    assert free0 - free1 < 2 << 20, (free0, free1)
