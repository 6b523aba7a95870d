"""Live updates on the device against the oracle's Live (src/live.rs): the device encodes
inserted rows bit for bit as the oracle does, lays each cluster's delta list out in insertion
order, and a live search (delta lists, tombstones, external ids) equals the oracle's."""
import numpy as np
import pytest

from oracle import oracle as O
from sieve_amd import Params, Rerank, parse_config
from sieve_amd.build import normalize_rows
from sieve_amd.index import Index, Live, ViewSpec, bitset

from test_gpu_parity import assert_same_up_to_ties

pytestmark = pytest.mark.gpu

PAD = 0xFFFFFFFF


def _ops(small):
    """A seeded sequence of live writes: new rows, upserts of corpus ids and of new ids,
    deletes of corpus, new and unknown ids."""
    rng = np.random.default_rng(11)
    new, _ = normalize_rows(small.mix.rows(small.n, 400))
    new = new.numpy().astype(np.float32)
    ext_new = (70_000_000 + np.arange(400)).astype(np.uint32)
    up = rng.choice(small.n, 25, replace=False).astype(np.uint32)
    dels = np.concatenate([rng.choice(small.n, 60, replace=False), ext_new[5:15], [4_000_000_000]]).astype(np.uint32)
    return [("insert", ext_new[:250], new[:250]), ("insert", up, -new[250:275]), ("delete", dels),
            ("insert", ext_new[250:], new[250:]), ("insert", ext_new[:3], new[300:303])]


@pytest.fixture(scope="module")
def live_pair(small, oracle_index):
    print(f"[gpu] live index k={small.k} n={small.n}", flush=True)
    ix = Index.create(small.host)
    base = ix.search(small.q768, small.q256, Params(nprobe=12, cand=400, k=10, rerank=Rerank.FULL))
    live = Live(ix, max_rows=2000, soar_lambda=1.0)
    olive = O.Live(oracle_index, soar_lambda=1.0)
    for op in _ops(small):
        if op[0] == "insert":
            live.insert(op[1], op[2])
            olive.insert(op[1], op[2])
        else:
            live.delete(op[1])
            olive.delete(op[1])
    return ix, live, olive, base


def test_live_state_matches_oracle(small, live_pair):
    ix, live, olive, _ = live_pair
    info = live.info()
    assert info["rows"] == olive.rows() == 428
    d = live.export()
    for j in range(info["rows"]):
        row, norm, t2, t2n, e = olive.row(j)
        assert np.array_equal(d["rows"][j], row), j
        assert np.float32(d["norm256"][j]).view(np.uint32) == np.float32(norm).view(np.uint32)
        assert np.array_equal(d["t2"][j], t2)
        assert np.float32(d["t2n"][j]).view(np.uint32) == np.float32(t2n).view(np.uint32)
        assert d["ext"][j] == e
    off = d["dblock_off"].astype(np.int64)
    assert off[-1] == info["arena_blocks"]
    for c in range(small.k):
        codes, fac, ids = olive.delta_list(c)
        lo, hi = off[c] * 32, off[c + 1] * 32
        assert hi - lo == (len(ids) + 31) // 32 * 32
        assert np.array_equal(d["ids"][lo:lo + len(ids)], ids), c
        assert (d["ids"][lo + len(ids):hi] == PAD).all()
        assert np.array_equal(d["codes"][lo:lo + len(ids)], codes), c
        assert np.array_equal(d["mul"][lo:lo + len(ids)].view(np.uint32), fac[:, 0].view(np.uint32))
        assert np.array_equal(d["err"][lo:lo + len(ids)].view(np.uint32), fac[:, 1].view(np.uint32))
    words = olive.deleted()
    assert info["deleted"] == int(sum(bin(int(w)).count("1") for w in words))
    for e in [int(x) for x in _ops(small)[2][1][:20]] + [70_000_000, 70_000_004, 70_000_007]:
        assert live.current(e) == olive.current(e)


@pytest.mark.parametrize("cfg", ["12:400:full", "48:2000:full", "8:300:prefix", "12:500:tiered60:b10x1.0",
                                 "16:400:ssd50", "12:400:none:p0.1"])
def test_live_search_matches_oracle(small, live_pair, cfg):
    ix, live, olive, _ = live_pair
    _, p = parse_config(cfg, small.k, k=10)
    s, r, n = live.search(small.q768, small.q256, p)
    ws, wr, wn = olive.search_batch(small.q768, small.q256, p, threads=4)
    for i in range(small.nq):
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])


def test_live_search_with_split_scan_and_allowlists(small, live_pair):
    """The delta segments go through the split scan's chunks and merge too, and per-query
    allowlists combine with the tombstones."""
    ix, live, olive, _ = live_pair
    p = Params(nprobe=16, cand=500, k=10, rerank=Rerank.FULL)
    want = olive.search_batch(small.q768, small.q256, p, threads=4)
    try:
        ix.set_scan_chunk(2)
        s, r, n = live.search(small.q768, small.q256, p)
    finally:
        ix.set_scan_chunk(0)
    for i in range(small.nq):
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], want[1][i, :want[2][i]], want[0][i, :want[2][i]])
    rng = np.random.default_rng(5)
    total = small.n + live.info()["rows"]
    allow = np.stack([bitset(rng.choice(total, total // 3, replace=False), total) for _ in range(small.nq)])
    view = ViewSpec(allow=allow)
    s, r, n = live.search(small.q768, small.q256, p, view=view)
    ws, wr, wn = olive.search_batch(small.q768, small.q256, p, view=O.make_view(allow=allow, allow_stride=allow.shape[1]),
                                    threads=4)
    for i in range(small.nq):
        assert_same_up_to_ties(r[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])


def test_inserted_vectors_find_themselves(small, live_pair):
    ix, live, olive, _ = live_pair
    new, _ = normalize_rows(small.mix.rows(small.n + 300, 8))
    q768 = new.numpy().astype(np.float32)
    q256 = normalize_rows(new[:, :256].contiguous())[0].numpy().astype(np.float32)
    s, r, n = live.search(q768, q256, Params(nprobe=8, cand=300, k=5, rerank=Rerank.FULL))
    assert (r[:, 0] == 70_000_000 + 300 + np.arange(8)).all()


def test_plain_search_sees_the_base_index_only(small, live_pair):
    ix, live, olive, base = live_pair
    s, r, n = ix.search(small.q768, small.q256, Params(nprobe=12, cand=400, k=10, rerank=Rerank.FULL))
    assert np.array_equal(r, base[1]) and np.array_equal(s, base[0]) and np.array_equal(n, base[2])


def test_live_errors(small):
    from sieve_amd.abi import SieveError
    ix = Index.create(small.host)
    live = Live(ix, max_rows=10)
    with pytest.raises(SieveError, match="already has a live state"):
        Live(ix, max_rows=10)
    vec = np.zeros((11, 768), np.float32)
    vec[:, 0] = 1.0
    with pytest.raises(SieveError, match="capacity"):
        live.insert(np.arange(11, dtype=np.uint32), vec)
    live.insert(np.arange(5, dtype=np.uint32) + 90_000_000, vec[:5])
    assert live.info()["rows"] == 5
    with pytest.raises(SieveError, match="tombstones"):
        live.search(small.q768, small.q256, Params(nprobe=4, cand=50, k=5),
                    view=ViewSpec(deleted=bitset(np.array([1]), small.n)))
    live.close()
    Live(ix, max_rows=4).close()
