"""sieve_search_batch_noalloc (all scratch in one fixed device array) against sieve_search_batch,
through the C ABI: bit-identical results on two calls in a row, the second with a smaller
batch, over the per-query, cluster-major and small-batch scans and with a tombstone view; no
device memory allocated by the calls; a batch too large for the array rejected."""
import ctypes

import numpy as np
import pytest

from sieve_amd import Params, Rerank, abi
from sieve_amd.index import Index, Stream, bitset

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    from conftest import SmallWorkload
    print("[gpu] no-alloc entry point: building a 24k-row index with 80 queries", flush=True)
    w = SmallWorkload(nq=80)
    ix = Index.create(w.host)
    q768 = torch.tensor(w.q768, device="cuda")
    q256 = torch.tensor(w.q256, device="cuda")
    return w, ix, q768, q256


def _outputs(nq, k):
    import torch
    return (torch.full((nq, k), -1.0, dtype=torch.float32, device="cuda"),
            torch.full((nq, k), -1, dtype=torch.int32, device="cuda"),
            torch.full((nq,), -1, dtype=torch.int32, device="cuda"))


def _noalloc(ix, st, q768, q256, p, view=None, out=None):
    import torch
    nq = q768.shape[0]
    s, r, n = out if out is not None else _outputs(nq, p.k)
    torch.cuda.synchronize()
    rc = abi.lib().sieve_search_batch_noalloc(ix._h, st._h, abi.ptr(q768), abi.ptr(q256), nq,
                                              ctypes.byref(abi.make_params(p)),
                                              ctypes.byref(view) if view is not None else None,
                                              abi.ptr(s), abi.ptr(r), abi.ptr(n))
    abi.check(rc)
    st.sync()
    return s, r, n


def _host(out):
    s, r, n = out
    return s.cpu().numpy(), r.cpu().numpy().view(np.uint32), n.cpu().numpy().view(np.uint32)


def _same(a, b):
    return all(np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32)) for x, y in zip(a, b))


@pytest.mark.parametrize("variant", [0, 10])
def test_noalloc_equals_search_batch_twice(setup, variant):
    import torch
    w, ix, q768, q256 = setup
    ix.set_scan_variant(variant)
    p = Params(nprobe=12, cand=400, k=10, rerank=Rerank.tiered(60), prune=(10, 1.0))
    try:
        want_full = [x.cpu().numpy() for x in ix.search(q768, q256, p)]
        want_small = [x.cpu().numpy() for x in ix.search(q768[:7], q256[:7], p)]
        st = Stream()
        out_full, out_small = _outputs(80, p.k), _outputs(7, p.k)  # before measuring: no allocation after
        q7s, q2s = q768[:7].contiguous(), q256[:7].contiguous()
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        _noalloc(ix, st, q768, q256, p, out=out_full)      # 80 queries (cluster-major under variant 10)
        _noalloc(ix, st, q7s, q2s, p, out=out_small)  # 7 queries (small-batch scan under 10)
        free1 = torch.cuda.mem_get_info()[0]
        got_full, got_small = _host(out_full), _host(out_small)
    finally:
        ix.set_scan_variant(0)
    print(f"[gpu] variant {variant}: device free bytes {free0} -> {free1}", flush=True)
    assert _same(got_full, want_full)
    assert _same(got_small, want_small)
    assert free1 == free0


def test_noalloc_with_tombstones_and_reuse_after_other_calls(setup):
    import torch
    w, ix, q768, q256 = setup
    p = Params(nprobe=10, cand=300, k=10, rerank=Rerank.FULL)
    dead = np.random.default_rng(4).choice(w.n, w.n // 6, replace=False)
    words = bitset(dead, w.n)
    dwords = torch.from_numpy(words.view(np.int64)).cuda()
    v = abi.View(deleted=abi.ptr(dwords).value, deleted_words=words.shape[0])
    from sieve_amd.index import ViewSpec
    want = [x.cpu().numpy() for x in ix.search(q768[:33], q256[:33], p, view=ViewSpec(deleted=words))]
    st = Stream()
    a = _host(_noalloc(ix, st, q768[:33].contiguous(), q256[:33].contiguous(), p, view=v))
    # ordinary calls in between, then the same fixed-scratch call again
    ix.search(q768, q256, Params(nprobe=20, cand=1000, k=10))
    b = _host(_noalloc(ix, st, q768[:33].contiguous(), q256[:33].contiguous(), p, view=v))
    assert _same(a, want) and _same(b, want)
    assert not set(a[1][a[1] != 0xFFFFFFFF].tolist()) & set(dead.tolist())


def test_noalloc_rejects_a_batch_beyond_the_array(setup):
    import torch
    w, ix, _, _ = setup
    ix.set_scan_variant(10)
    try:
        nq = 4000  # appended lists of 64 x 2048 entries per query: far beyond 64 MiB
        q = torch.zeros(nq, 768, device="cuda")
        q2 = torch.zeros(nq, 256, device="cuda")
        p = Params(nprobe=8, cand=2048, k=10)
        with pytest.raises(abi.SieveError, match="fixed"):
            _noalloc(ix, Stream(), q, q2, p)
    finally:
        ix.set_scan_variant(0)


def test_noalloc_requires_device_arrays(setup):
    w, ix, _, _ = setup
    p = Params(nprobe=8, cand=100, k=10)
    s = np.zeros((2, 10), np.float32)
    r = np.zeros((2, 10), np.uint32)
    n = np.zeros(2, np.uint32)
    st = Stream()
    rc = abi.lib().sieve_search_batch_noalloc(ix._h, st._h, abi.ptr(w.q768[:2].copy()), abi.ptr(w.q256[:2].copy()),
                                              2, ctypes.byref(abi.make_params(p)), None, abi.ptr(s), abi.ptr(r),
                                              abi.ptr(n))
    assert rc == -1 and b"device memory" in abi.lib().sieve_last_error()
