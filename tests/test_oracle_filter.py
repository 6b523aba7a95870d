"""The oracle's restatement of filter.rs: the planner's thresholds (filter.rs:157-165), the
filter's per-cluster counts and rows, and each plan against brute force over allowed rows."""
import numpy as np

from oracle import oracle as O
from sieve_amd import Params, Rerank
from sieve_amd.index import bitset


def test_auto_plan():
    n = 10_000_000
    assert O.plan_resolve("auto", 250_000, n) == O.PLANS["brute"]
    assert O.plan_resolve("auto", 1_000_000, n) == O.PLANS["expand"]
    assert O.plan_resolve("auto", 3_000_000, n) == O.PLANS["post"]
    assert O.plan_resolve("expand", 1, n) == O.PLANS["expand"]


def _allow(small, frac, seed):
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(small.n, int(small.n * frac), replace=False))
    return rows, bitset(rows, small.n)


def test_filter_counts_and_rows(small, oracle_index):
    rows, allow = _allow(small, 0.1, 1)
    f = O.Filter(oracle_index, allow)
    assert f.rows() == len(rows)
    h = small.host
    bo = h["block_off"].astype(np.int64)
    ok = np.zeros(small.n + 1, bool)
    ok[rows] = True
    ids = h["ids"].astype(np.int64)
    want = [int(np.sum(ok[np.where(ids[bo[c] * 32:bo[c + 1] * 32] == 0xFFFFFFFF, small.n,
                                   ids[bo[c] * 32:bo[c + 1] * 32])])) for c in range(small.k)]
    assert f.per_cluster().tolist() == want


def _brute(small, q768, rows, k):
    quant = small.host["quant"]
    wq, s = O.query_i8((q768 / quant).astype(np.float32))
    inv = np.float32(np.float32(1.0) / np.float32(s))
    dots = small.host["rerank"][rows].astype(np.int64) @ wq.astype(np.int64)
    sc = (dots.astype(np.float32) * inv).astype(np.float32)
    o = np.lexsort((rows, -sc))[:k]
    return sc[o], rows[o]


def test_plans_against_brute_force(small, oracle_index):
    rows, allow = _allow(small, 0.2, 2)
    f = O.Filter(oracle_index, allow)
    p = Params(nprobe=small.k, cand=100_000, k=10, rerank=Rerank.FULL)
    for plan, target in (("brute", 0), ("post", 0), ("expand", 10 ** 9)):
        s, r, n, used = f.search(small.q768[:16], small.q256[:16], p, plan=plan, target=target)
        assert (used == O.PLANS[plan]).all()
        for i in range(16):
            ws, wr = _brute(small, small.q768[i], rows, 10)
            cut = ws[-1]
            assert {int(x) for x, y in zip(r[i, :n[i]], s[i, :n[i]]) if y > cut} == \
                   {int(x) for x, y in zip(wr, ws) if y > cut}
            assert np.isin(r[i, :n[i]], rows).all()


def test_expand_takes_clusters_until_the_target(small, oracle_index):
    """Expand(target) probes clusters in route order, skipping empty ones, while fewer than
    target allowed slots are covered: results come only from those clusters."""
    rows, allow = _allow(small, 0.05, 3)
    f = O.Filter(oracle_index, allow)
    per = f.per_cluster()
    p = Params(nprobe=4, cand=2000, k=10, rerank=Rerank.FULL)
    target = 300
    s, r, n, used = f.search(small.q768[:8], small.q256[:8], p, plan="expand", target=target)
    h = small.host
    bo = h["block_off"].astype(np.int64)
    for i in range(8):
        sc, cl = oracle_index.route(small.q256[i], Params(nprobe=small.k, cand=1, k=1))
        cov, taken = 0, []
        for c in cl:
            if per[c] == 0:
                continue
            if cov >= target:
                break
            cov += per[c]
            taken.append(c)
        assert cov >= min(target, per.sum()) and len(taken) < small.k
        in_taken = set()
        for c in taken:
            in_taken |= set(h["ids"][bo[c] * 32:bo[c + 1] * 32].tolist())
        assert set(r[i, :n[i]].tolist()) <= in_taken
