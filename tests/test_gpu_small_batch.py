"""The small-batch scan (at most 16 unfiltered queries on the cluster-major variants: per-chunk
bounds, a threshold per query below its cand-th best admissible slot, the slots above it
appended, the finish merge) gives the per-query scan's candidates bit for bit: for one query
and up to 16, for every candidate count (also below the 16 bound entries a chunk keeps), with
tombstones, over live delta lists, and with appended lists too small to hold them (those
queries are rescanned)."""
import numpy as np
import pytest

from oracle import oracle as O
from sieve_amd import Params, Rerank, abi
from sieve_amd.build import normalize_rows
from sieve_amd.index import Index, Live, ViewSpec, bitset

from test_gpu_cluster_scan import _cands, _same_candidates
from test_gpu_parity import assert_same_up_to_ties

pytestmark = pytest.mark.gpu

CM = abi.SCAN_VARIANTS_CM


@pytest.fixture(scope="module")
def sb_index(small):
    print(f"[gpu] small-batch index k={small.k} n={small.n}", flush=True)
    ix = Index.create(small.host)
    ix.set_small_batch(16)
    return ix


@pytest.mark.parametrize("nq", [1, 2, 7, 16])
@pytest.mark.parametrize("cand,nprobe", [(300, 16), (1000, 48), (8, 4), (37, 12)])
def test_small_batches_equal_the_per_query_scan(small, sb_index, nq, cand, nprobe):
    q256 = small.q256[:nq]
    p = Params(nprobe=nprobe, cand=cand, k=min(10, cand), rerank=Rerank.FULL)
    probes = sb_index.route(q256, p)
    t0, t1 = abi.Timing(), abi.Timing()
    want = _cands(sb_index, q256, probes, p, 0, timing=t0)
    s = sb_index.stream()
    got = _cands(sb_index, q256, probes, p, CM[0], timing=t1, stream=s)
    _same_candidates(want, got)
    if s.last_stats()["rescanned"] == 0:
        assert t1.scanned == t0.scanned  # every probed slot scored once by the append pass


def test_small_batches_with_tombstones(small, sb_index):
    q256 = small.q256[:6]
    p = Params(nprobe=12, cand=250, k=10, rerank=Rerank.FULL)
    probes = sb_index.route(q256, p)
    dead = np.random.default_rng(5).choice(small.n, small.n // 5, replace=False)
    v = ViewSpec(deleted=bitset(dead, small.n))
    want = _cands(sb_index, q256, probes, p, 0, view=v)
    got = _cands(sb_index, q256, probes, p, CM[1], view=v)
    _same_candidates(want, got)


def test_small_batches_rescan_overflowing_lists(small, sb_index):
    q256 = small.q256[:4]
    p = Params(nprobe=24, cand=150, k=10, rerank=Rerank.FULL)
    probes = sb_index.route(q256, p)
    want = _cands(sb_index, q256, probes, p, 0)
    sb_index.set_cluster_scan(ext_cap=40)
    s = sb_index.stream()
    try:
        got = _cands(sb_index, q256, probes, p, CM[0], stream=s)
    finally:
        sb_index.set_cluster_scan()
    _same_candidates(want, got)
    assert s.last_stats()["rescanned"] == 4


def test_small_batch_search_matches_oracle(small, oracle_index, sb_index):
    p = Params(nprobe=16, cand=700, k=10, rerank=Rerank.FULL)
    sb_index.set_scan_variant(CM[0])
    try:
        tm = abi.Timing()
        s, r, n = sb_index.search(small.q768[:3], small.q256[:3], p, timing=tm)
    finally:
        sb_index.set_scan_variant(0)
    ws, wr, wn, wt = oracle_index.search_batch(small.q768[:3], small.q256[:3], p, threads=3)
    for i in range(3):
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])
    assert tm.scanned == wt["scanned"]


def test_small_batches_over_live_delta_lists(small):
    ix = Index.create(small.host)
    ix.set_small_batch(16)
    live = Live(ix, max_rows=1000, soar_lambda=1.0)
    new, _ = normalize_rows(small.mix.rows(small.n, 600))
    live.insert((80_000_000 + np.arange(600)).astype(np.uint32), new.numpy().astype(np.float32))
    live.delete(np.arange(0, small.n, 7).astype(np.uint32))
    p = Params(nprobe=16, cand=400, k=10, rerank=Rerank.FULL)
    for nq in (1, 9):
        want = live.search(small.q768[:nq], small.q256[:nq], p)
        for v in CM:
            ix.set_scan_variant(v)
            got = live.search(small.q768[:nq], small.q256[:nq], p)
            for w, g in zip(want, got):
                assert np.array_equal(w, g)
        ix.set_scan_variant(0)
    live.close()
