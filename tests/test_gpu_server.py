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
