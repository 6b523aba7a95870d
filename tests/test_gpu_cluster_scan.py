"""The cluster-major scan (scan variants 10 and 11: a seed pass per query, then each cluster's
blocks read once for a tile of the queries probing it and scored with int8 MFMA, then a
per-query merge) gives the per-query scan's candidates bit for bit, and the oracle's results,
for every seed size, list capacity (overflowing lists are rescanned), work-item size, view and
live delta list."""
import numpy as np
import pytest

from oracle import oracle as O
from sieve_amd import Params, Rerank, abi
from sieve_amd.build import normalize_rows
from sieve_amd.index import Index, Live, ViewSpec, bitset

from test_gpu_parity import assert_same_up_to_ties

pytestmark = pytest.mark.gpu

CM = abi.SCAN_VARIANTS_CM


@pytest.fixture(scope="module")
def cm_index(small):
    print(f"[gpu] cluster-scan index k={small.k} n={small.n}", flush=True)
    return Index.create(small.host)


def _queries(small, reps=3):
    """The fixture's queries repeated (so clusters are probed by more than 16 queries)."""
    return np.tile(small.q768, (reps, 1)), np.tile(small.q256, (reps, 1))


def _cands(ix, q256, probes, p, variant, view=None, timing=None, stream=None):
    ix.set_scan_variant(variant)
    try:
        return ix.candidates(q256, *probes, p, view=view, timing=timing, stream=stream)
    finally:
        ix.set_scan_variant(0)


def _same_candidates(want, got):
    n = want[3]
    assert np.array_equal(got[3], n)
    for i in range(len(n)):
        for w, g in zip(want[:3], got[:3]):
            assert np.array_equal(w[i, :n[i]], g[i, :n[i]]), i


@pytest.mark.parametrize("variant", CM)
@pytest.mark.parametrize("cand,nprobe,seed,cap,chunk,phases", [
    (300, 16, 0, 0, 32, 0),      # defaults
    (300, 32, 1, 0, 1, 1),       # seed = the nearest probe alone; one block per work item; one phase
    (1000, 48, 0, 0, 7, 3),      # every cluster probed
    (37, 8, 0, 0, 32, 8),
    (150, 24, 1, 40, 3, 1),      # tiny appended lists: overflowing queries are rescanned
    (150, 24, 1, 60, 5, 5),      # ... also when a list overflows in an early phase
    (300, 16, 1 << 30, 0, 32, 0),  # the seed pass covers every probe: nothing left for the MFMA pass
])
def test_cluster_scan_equals_per_query_scan(small, cm_index, variant, cand, nprobe, seed, cap, chunk, phases):
    q768, q256 = _queries(small)
    p = Params(nprobe=nprobe, cand=cand, k=10, rerank=Rerank.FULL)
    probes = cm_index.route(q256, p)
    t0, t1 = abi.Timing(), abi.Timing()
    want = _cands(cm_index, q256, probes, p, 0, timing=t0)
    cm_index.set_cluster_scan(min_queries=1, seed_slots=seed, ext_cap=cap, chunk_blocks=chunk, phases=phases)
    s = cm_index.stream()
    try:
        got = _cands(cm_index, q256, probes, p, variant, timing=t1, stream=s)
    finally:
        cm_index.set_cluster_scan()
    _same_candidates(want, got)
    # Every stored slot of every probe is scored once, plus full rescans of overflowed queries.
    if s.last_stats()["rescanned"] == 0:
        assert t1.scanned == t0.scanned
    else:
        assert t1.scanned > t0.scanned


def test_cluster_scan_reports_rescans(small, cm_index):
    q768, q256 = _queries(small)
    p = Params(nprobe=24, cand=150, k=10, rerank=Rerank.FULL)
    probes = cm_index.route(q256, p)
    s = cm_index.stream()
    cm_index.set_cluster_scan(min_queries=1, seed_slots=1, ext_cap=8)
    cm_index.set_scan_variant(CM[0])
    try:
        cm_index.candidates(q256, *probes, p, stream=s)
        st = s.last_stats()
        cm_index.set_cluster_scan(min_queries=1)
        cm_index.candidates(q256, *probes, p, stream=s)
        st2 = s.last_stats()
    finally:
        cm_index.set_cluster_scan()
        cm_index.set_scan_variant(0)
    assert 0 < st["rescanned"] <= len(q256)
    assert st2["rescanned"] == 0


@pytest.mark.parametrize("variant", CM)
def test_cluster_scan_search_matches_oracle(small, oracle_index, cm_index, variant):
    p = Params(nprobe=16, cand=700, k=10, rerank=Rerank.FULL)
    cm_index.set_cluster_scan(min_queries=1)
    cm_index.set_scan_variant(variant)
    try:
        tm = abi.Timing()
        s, r, n = cm_index.search(small.q768, small.q256, p, timing=tm)
    finally:
        cm_index.set_scan_variant(0)
        cm_index.set_cluster_scan()
    ws, wr, wn, wt = oracle_index.search_batch(small.q768, small.q256, p, threads=4)
    for i in range(small.nq):
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])
    assert tm.scanned == wt["scanned"]


def test_cluster_scan_with_tombstones_and_allowlists(small, cm_index):
    """Tombstones are honoured by the MFMA pass; an allow filter sends the batch to the
    per-query scan."""
    q768, q256 = _queries(small, 2)
    p = Params(nprobe=12, cand=250, k=10, rerank=Rerank.FULL)
    probes = cm_index.route(q256, p)
    rng = np.random.default_rng(5)
    dead = rng.choice(small.n, small.n // 5, replace=False)
    allow = rng.choice(small.n, small.n // 2, replace=False)
    views = [ViewSpec(deleted=bitset(dead, small.n)), ViewSpec(deleted=bitset(dead, small.n),
                                                               allow=bitset(allow, small.n))]
    for v in views:
        want = _cands(cm_index, q256, probes, p, 0, view=v)
        cm_index.set_cluster_scan(min_queries=1, seed_slots=1)
        try:
            got = _cands(cm_index, q256, probes, p, CM[0], view=v)
        finally:
            cm_index.set_cluster_scan()
        _same_candidates(want, got)


def test_cluster_scan_over_live_delta_lists(small):
    """Delta lists are segments of their own; results equal the per-query scan's."""
    ix = Index.create(small.host)
    live = Live(ix, max_rows=1000, soar_lambda=1.0)
    new, _ = normalize_rows(small.mix.rows(small.n, 600))
    new = new.numpy().astype(np.float32)
    live.insert((80_000_000 + np.arange(600)).astype(np.uint32), new)
    live.delete(np.arange(0, small.n, 7).astype(np.uint32))
    q768, q256 = _queries(small)
    p = Params(nprobe=16, cand=400, k=10, rerank=Rerank.FULL)
    want = live.search(q768, q256, p)
    ix.set_cluster_scan(min_queries=1, seed_slots=1)
    for v in CM:
        ix.set_scan_variant(v)
        got = live.search(q768, q256, p)
        for w, g in zip(want, got):
            assert np.array_equal(w, g)
    ix.set_scan_variant(0)
    live.close()


def test_cluster_scan_small_batches_equal_the_per_query_scan(small, cm_index):
    """Below min_queries a batch takes the small-batch scan (same results)."""
    p = Params(nprobe=12, cand=300, k=10, rerank=Rerank.FULL)
    probes = cm_index.route(small.q256[:5], p)
    want = _cands(cm_index, small.q256[:5], probes, p, 0)
    got = _cands(cm_index, small.q256[:5], probes, p, CM[0])
    _same_candidates(want, got)
