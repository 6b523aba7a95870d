"""Batched search as captured HIP graphs (sieve_stream_set_graphs): the first call with a given
key runs eagerly, the second captures the launches, later ones replay them. Every result equals
the eager stream's bit for bit, for the per-query and the cluster-major scan and for host and
device buffers, with other eager work and synchronisation on the stream between the replays;
a scratch reallocation retires the graphs that hold the old addresses."""
import numpy as np
import pytest
import torch

from sieve_amd import Params, Rerank
from sieve_amd.index import Index, Stream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gix(small):
    print(f"[gpu] graph index k={small.k} n={small.n}", flush=True)
    return Index.create(small.host)


def _host(x):
    return x.cpu().numpy() if hasattr(x, "cpu") else x


def _same(want, got):
    ws, wr, wn = (_host(x) for x in want)
    gs, gr, gn = (_host(x) for x in got)
    assert np.array_equal(wn.view(np.uint32), gn.view(np.uint32))
    for i, c in enumerate(wn.view(np.uint32)):
        assert np.array_equal(ws[i, :c].view(np.uint32), gs[i, :c].view(np.uint32)), i
        assert np.array_equal(wr[i, :c].view(np.uint32), gr[i, :c].view(np.uint32)), i


def _graph_stream():
    s = Stream()
    s.set_graphs(True)
    return s


def test_replays_equal_the_eager_search(small, gix):
    p = Params(nprobe=16, cand=300, k=10, rerank=Rerank.FULL)
    want = gix.search(small.q768, small.q256, p)
    s = _graph_stream()
    modes = []
    for _ in range(5):
        got = gix.search(small.q768, small.q256, p, stream=s)
        modes.append(s.last_stats()["graph"])
        _same(want, got)
    assert modes == [0, 1, 2, 2, 2]


def test_replays_with_eager_work_between(small, gix):
    p = Params(nprobe=12, cand=200, k=10, rerank=Rerank.FULL, margin=0.3)
    p2 = Params(nprobe=8, cand=100, k=5, rerank=Rerank.FULL)
    want = gix.search(small.q768, small.q256, p)
    want_few = gix.search(small.q768[:5], small.q256[:5], p)
    want_other = gix.search(small.q768[5:], small.q256[5:], p2)
    s, eager = _graph_stream(), Stream()
    # Every stream sizes its scratch before the loop (a reallocation on any stream retires the
    # graphs), then both keys capture.
    for _ in range(3):
        gix.search(small.q768, small.q256, p, stream=s)
        gix.search(small.q768[:5], small.q256[:5], p, stream=s)
        gix.search(small.q768[5:], small.q256[5:], p2, stream=eager)
    modes = []
    for _ in range(6):
        got = gix.search(small.q768, small.q256, p, stream=s)
        modes.append(s.last_stats()["graph"])
        _same(want, got)
        # another graphed batch on the same stream, an eager search and torch work elsewhere,
        # then a device-wide synchronisation
        _same(want_few, gix.search(small.q768[:5], small.q256[:5], p, stream=s))
        _same(want_other, gix.search(small.q768[5:], small.q256[5:], p2, stream=eager))
        x = torch.randn(1 << 20, device="cuda")
        assert torch.isfinite(x.sum()).item()
        torch.cuda.synchronize()
    assert modes == [2] * 6, modes


def test_scratch_growth_retires_graphs(small, gix):
    p = Params(nprobe=16, cand=300, k=10, rerank=Rerank.FULL)
    want = gix.search(small.q768[:20], small.q256[:20], p)
    s = _graph_stream()
    modes = []
    for i in range(7):
        if i == 3:  # a larger batch grows the stream's scratch
            big768, big256 = np.tile(small.q768, (4, 1)), np.tile(small.q256, (4, 1))
            _same(gix.search(big768, big256, p), gix.search(big768, big256, p, stream=s))
        got = gix.search(small.q768[:20], small.q256[:20], p, stream=s)
        modes.append(s.last_stats()["graph"])
        _same(want, got)
    assert modes == [0, 1, 2, 0, 1, 2, 2]


def test_cluster_major_batches_replay(small, gix):
    q768, q256 = np.tile(small.q768, (3, 1)), np.tile(small.q256, (3, 1))
    p = Params(nprobe=24, cand=300, k=10, rerank=Rerank.FULL)
    gix.set_scan_variant(10)
    gix.set_cluster_scan(min_queries=1)
    try:
        want = gix.search(q768, q256, p)
        s = _graph_stream()
        modes = []
        for _ in range(4):
            got = gix.search(q768, q256, p, stream=s)
            modes.append(s.last_stats()["graph"])
            _same(want, got)
            torch.cuda.synchronize()
        assert modes == [0, 1, 2, 2]
    finally:
        gix.set_cluster_scan()
        gix.set_scan_variant(0)


def test_device_buffers_replay(small, gix):
    p = Params(nprobe=16, cand=300, k=10, rerank=Rerank.FULL)
    want = gix.search(small.q768, small.q256, p)
    q768 = torch.from_numpy(small.q768).cuda()
    q256 = torch.from_numpy(small.q256).cuda()
    out = (torch.empty(small.nq, 10, dtype=torch.float32, device="cuda"),
           torch.empty(small.nq, 10, dtype=torch.int32, device="cuda"),
           torch.empty(small.nq, dtype=torch.int32, device="cuda"))
    s = _graph_stream()
    modes = []
    for _ in range(4):
        for o in out:
            o.fill_(-1)
        torch.cuda.synchronize()  # the fill runs on torch's stream
        got = gix.search(q768, q256, p, stream=s, out=out)
        s.sync()
        modes.append(s.last_stats()["graph"])
        _same(want, got)
    assert modes == [0, 1, 2, 2]


def test_small_batch_scan_replays(small, gix):
    """A batch of a few queries on the cluster-major variant takes the small-batch scan (bound,
    threshold, append and finish launches): captured and replayed, every result equal."""
    p = Params(nprobe=16, cand=300, k=10, rerank=Rerank.FULL)
    gix.set_scan_variant(10)
    try:
        for nq in (1, 5):
            want = gix.search(small.q768[:nq], small.q256[:nq], p)
            s = _graph_stream()
            modes = []
            for _ in range(4):
                got = gix.search(small.q768[:nq], small.q256[:nq], p, stream=s)
                modes.append(s.last_stats()["graph"])
                _same(want, got)
                torch.cuda.synchronize()
            assert modes == [0, 1, 2, 2]
    finally:
        gix.set_scan_variant(0)
