"""Filtered search on the device (filter.rs) against the oracle's restatement: the filter's
per-cluster counts and rows, the planner, and every plan's results."""
import numpy as np
import pytest

from oracle import oracle as O
from sieve_amd import Params, Rerank, parse_config
from sieve_amd.index import Filter, Index, bitset

from test_gpu_parity import assert_same_up_to_ties

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fx(small):
    print(f"[gpu] filter index k={small.k} n={small.n}", flush=True)
    return Index.create(small.host)


def _allow(small, frac, seed):
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(small.n, max(1, int(small.n * frac)), replace=False))
    return rows, bitset(rows, small.n)


@pytest.mark.parametrize("frac", [0.003, 0.1, 0.6])
def test_filter_counts_match_oracle(small, oracle_index, fx, frac):
    rows, allow = _allow(small, frac, 7)
    f, of = Filter(fx, allow), O.Filter(oracle_index, allow)
    assert f.rows() == of.rows() == len(rows)
    assert np.array_equal(f.per_cluster(), of.per_cluster())


def test_plan_resolve_matches_oracle():
    from sieve_amd import abi
    for rows, n in ((250_000, 10_000_000), (250_001, 10_000_000), (3_000_000, 10_000_000), (1, 1), (0, 100)):
        for plan in range(4):
            assert abi.lib().sieve_plan_resolve(plan, rows, n) == O.plan_resolve(plan, rows, n)


@pytest.mark.parametrize("plan,frac,target,cfg", [
    ("post", 0.3, 0, "12:400:full"),
    ("post", 0.05, 0, "16:300:tiered40:b10x1.0"),
    ("expand", 0.05, 400, "12:400:full"),
    ("expand", 0.2, 3000, "8:500:prefix"),
    ("brute", 0.02, 0, "12:400:full"),
    ("brute", 0.4, 0, "12:400:full"),
    ("auto", 0.1, 0, "12:400:full"),
])
def test_filter_search_matches_oracle(small, oracle_index, fx, plan, frac, target, cfg):
    rows, allow = _allow(small, frac, 13)
    f, of = Filter(fx, allow), O.Filter(oracle_index, allow)
    _, p = parse_config(cfg, small.k, k=10)
    s, r, n, used = f.search(small.q768, small.q256, p, plan=plan, target=target)
    ws, wr, wn, wused = of.search(small.q768, small.q256, p, plan=plan, target=target)
    assert (wused == used).all()
    for i in range(small.nq):
        assert np.isin(r[i, :n[i]], rows).all()
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])


@pytest.mark.parametrize("frac,target,cfg", [(0.01, 300, "12:400:full"), (0.05, 400, "12:400:full"),
                                             (0.2, 3000, "8:500:prefix"), (0.12, 200_000, "16:300:full")])
def test_expand_over_listed_slots_equals_whole_clusters(small, oracle_index, fx, frac, target, cfg):
    """A filter below the Post share keeps each cluster's allowed slots; the Expand scan reads
    only those: the same results bit for bit as scanning the probed clusters whole, fewer slots
    read, and the oracle's results up to ties."""
    from sieve_amd import abi
    rows, allow = _allow(small, frac, 29)
    f, of = Filter(fx, allow), O.Filter(oracle_index, allow)
    _, p = parse_config(cfg, small.k, k=10)
    t_list, t_whole = abi.Timing(), abi.Timing()
    got = f.search(small.q768, small.q256, p, plan="expand", target=target, timing=t_list)
    f.set_gather(False)
    want = f.search(small.q768, small.q256, p, plan="expand", target=target, timing=t_whole)
    f.set_gather(True)
    for g, w in zip(got[:3], want[:3]):
        assert np.array_equal(np.asarray(g).view(np.uint8), np.asarray(w).view(np.uint8))
    assert got[3] == want[3] == 1  # Expand
    assert t_list.scanned < t_whole.scanned
    ws, wr, wn, _ = of.search(small.q768, small.q256, p, plan="expand", target=target)
    s, r, n = got[:3]
    for i in range(small.nq):
        assert np.isin(r[i, :n[i]], rows).all()
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])
