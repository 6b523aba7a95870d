"""Durability of the device live state against the oracle (src/live.rs): the write-ahead log
the device writes is byte-identical with the oracle's, a restart that replays it rebuilds the
state bit for bit (a torn tail cut, an unknown op refused), and compaction writes the oracle's
merged lists, the postings records, meta.json and the row segment; the compacted index, opened
from its files, searches as the live state did."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from sieve_amd import Params, Rerank, parse_config
from sieve_amd.abi import SieveError
from sieve_amd.build import normalize_rows, write_index
from sieve_amd.index import Index, Live

from test_gpu_parity import assert_same_up_to_ties

pytestmark = pytest.mark.gpu


def _ops(small):
    rng = np.random.default_rng(29)
    new, _ = normalize_rows(small.mix.rows(small.n + 1000, 420))
    new = new.numpy().astype(np.float32)
    ext = (80_000_000 + np.arange(420)).astype(np.uint32)
    up = rng.choice(small.n, 30, replace=False).astype(np.uint32)
    dels = np.concatenate([rng.choice(small.n, 70, replace=False), ext[7:19], [4_000_000_002]]).astype(np.uint32)
    return [("insert", ext[:260], new[:260]), ("insert", up, -new[260:290]), ("delete", dels),
            ("insert", ext[260:], new[260:]), ("delete", ext[300:305]), ("insert", ext[:4], new[290:294])]


def _apply(live, ops):
    for op in ops:
        if op[0] == "insert":
            live.insert(op[1], op[2])
        else:
            live.delete(op[1])


@pytest.fixture(scope="module")
def durable(small, tmp_path_factory):
    d = tmp_path_factory.mktemp("livefs")
    ix_dir, ing_dir = str(d / "index"), str(d / "ingest")
    write_index(small.arrays, ix_dir, ing_dir)
    print(f"[gpu] livefs: index k={small.k} n={small.n} in {d}", flush=True)
    ix = Index.open(ix_dir, ing_dir)
    live = Live(ix, max_rows=3000, index_dir=ix_dir, wal=str(d / "wal"))
    oix = O.Index(small.host)
    olive = O.Live(oix, soar_lambda=1.0)
    olive.open_wal(str(d / "owal"))
    ops = _ops(small)
    _apply(live, ops)
    _apply(olive, ops)
    return dict(d=d, ix_dir=ix_dir, ing_dir=ing_dir, ix=ix, live=live, oix=oix, olive=olive, ops=ops)


def test_log_is_byte_identical_with_the_oracle(durable):
    a = (durable["d"] / "wal").read_bytes()
    b = (durable["d"] / "owal").read_bytes()
    ins = sum(len(o[1]) for o in durable["ops"] if o[0] == "insert")
    dels = sum(len(o[1]) for o in durable["ops"] if o[0] == "delete")
    assert len(a) == ins * (9 + 768) + dels * 5
    assert a == b


def _exports_equal(a, b):
    ea, eb = a.export(), b.export()
    for f in ea:
        assert np.array_equal(np.asarray(ea[f]).view(np.uint8), np.asarray(eb[f]).view(np.uint8)), f
    ia, ib = a.info(), b.info()
    for f in ("rows", "slots", "arena_blocks", "deleted"):
        assert ia[f] == ib[f], f


def test_replay_rebuilds_the_state(small, durable):
    """A fresh index and live state from the files and the log alone: the same delta rows,
    slots, tombstones and external ids, and bit-identical searches."""
    ix2 = Index.open(durable["ix_dir"], durable["ing_dir"])
    wal_copy = durable["d"] / "wal_copy"
    wal_copy.write_bytes((durable["d"] / "wal").read_bytes())
    live2 = Live(ix2, max_rows=3000, index_dir=durable["ix_dir"], wal=str(wal_copy))
    ins = sum(len(o[1]) for o in durable["ops"] if o[0] == "insert")
    dels = sum(len(o[1]) for o in durable["ops"] if o[0] == "delete")
    assert (live2.replay["inserts"], live2.replay["deletes"], live2.replay["torn"]) == (ins, dels, 0)
    _exports_equal(durable["live"], live2)
    for e in [int(x) for x in durable["ops"][2][1][:20]] + [80_000_000, 80_000_008, 80_000_301, 80_000_419]:
        assert live2.current(e) == durable["live"].current(e) == durable["olive"].current(e)
    for cfg in ("12:400:full", "8:300:tiered60:b10x1.0"):
        _, p = parse_config(cfg, small.k, k=10)
        a = durable["live"].search(small.q768, small.q256, p)
        b = live2.search(small.q768, small.q256, p)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32)), cfg
    size = wal_copy.stat().st_size
    live2.delete(np.array([80_000_100], np.uint32))
    assert wal_copy.stat().st_size == size + 5
    live2.close()


@pytest.mark.parametrize("cut", [3, 400])
def test_replay_cuts_a_torn_tail(small, durable, cut):
    """The log cut inside its last insert record: replay drops the partial record, truncates the
    file to the last whole record and gives the oracle's state for the same bytes."""
    full = (durable["d"] / "wal").read_bytes()
    last = len(full) - 4 * (9 + 768)  # the final op inserts four rows
    torn = durable["d"] / f"torn{cut}"
    torn.write_bytes(full[:last + 3 * (9 + 768) + cut])
    otorn = durable["d"] / f"otorn{cut}"
    otorn.write_bytes(torn.read_bytes())
    ix2 = Index.open(durable["ix_dir"], durable["ing_dir"])
    live2 = Live(ix2, max_rows=3000, index_dir=durable["ix_dir"], wal=str(torn))
    assert live2.replay["torn"] == cut and torn.stat().st_size == last + 3 * (9 + 768)
    ol = O.Live(durable["oix"], soar_lambda=1.0)
    assert ol.open_wal(str(otorn))["torn"] == cut
    d = live2.export()
    assert live2.info()["rows"] == ol.rows()
    for j in range(ol.rows()):
        row, norm, t2, t2n, e = ol.row(j)
        assert np.array_equal(d["rows"][j], row) and d["ext"][j] == e
    assert live2.current(80_000_003) == ol.current(80_000_003)
    live2.close()


def test_replay_refuses_an_unknown_op(durable):
    bad = durable["d"] / "bad"
    bad.write_bytes(bytes([2, 1, 0, 0, 0, 9, 0, 0]))
    ix2 = Index.open(durable["ix_dir"], durable["ing_dir"])
    with pytest.raises(SieveError, match="bad op 9 at byte 5"):
        Live(ix2, max_rows=10, index_dir=durable["ix_dir"], wal=str(bad))
    Live(ix2, max_rows=10).close()  # the failed open left no live state attached


@pytest.fixture(scope="module")
def compacted(small, durable):
    out = durable["d"] / "compacted"
    st = durable["live"].compact(str(out))
    m = durable["olive"].compact()
    return dict(out=str(out), st=st, m=m)


def test_compaction_writes_the_oracles_lists(small, durable, compacted):
    out, st, m = compacted["out"], compacted["st"], compacted["m"]
    rd = lambda f, dt: np.fromfile(os.path.join(out, f), dt)
    assert np.array_equal(rd("block_off.u64", np.uint64), m["block_off"])
    assert np.array_equal(rd("blocks.bin", np.uint8).reshape(-1, 1024), m["blocks"])
    assert np.array_equal(rd("factors.f32", np.uint32), m["factors"].view(np.uint32).ravel())
    assert np.array_equal(rd("ids.u32", np.uint32), m["ids"])
    assert (st["slots"], st["dropped"], st["blocks"]) == (m["slots"], m["dropped"], len(m["blocks"]))
    # postings::pack: record c = blocks | factors | ids, each record padded to 4 KiB
    off = rd("postings.off", np.uint64).astype(np.int64)
    pb = rd("postings.bin", np.uint8)
    bo = m["block_off"].astype(np.int64)
    assert len(off) == small.k + 1 and off[-1] == len(pb)
    for c in range(small.k):
        b0, b1 = bo[c], bo[c + 1]
        rec = np.concatenate([m["blocks"][b0:b1].ravel(), m["factors"][b0 * 32:b1 * 32].ravel().view(np.uint8),
                              m["ids"][b0 * 32:b1 * 32].view(np.uint8)])
        assert off[c] % 4096 == 0 and off[c + 1] - off[c] == (len(rec) + 4095) // 4096 * 4096
        assert np.array_equal(pb[off[c]:off[c] + len(rec)], rec), c
        assert not pb[off[c] + len(rec):off[c + 1]].any()


def test_compaction_writes_the_segment_and_meta(small, durable, compacted):
    out, ol = compacted["out"], durable["olive"]
    rd = lambda f, dt: np.fromfile(os.path.join(out, f), dt)
    r = ol.rows()
    rows, ext = rd("live.rows.i8", np.int8).reshape(r, 768), rd("live.ext.u32", np.uint32)
    norm, t2, t2n = rd("live.norm256.f32", np.float32), rd("live.t2.bin", np.uint8).reshape(r, 64), rd(
        "live.t2.norm", np.float32)
    for j in range(r):
        w = ol.row(j)
        assert np.array_equal(rows[j], w[0]) and norm[j].tobytes() == np.float32(w[1]).tobytes()
        assert np.array_equal(t2[j], w[2]) and t2n[j].tobytes() == np.float32(w[3]).tobytes() and ext[j] == w[4]
    assert np.array_equal(rd("live.deleted.u64", np.uint64), ol.deleted())
    meta = json.load(open(os.path.join(out, "meta.json")))
    src = json.load(open(os.path.join(durable["ix_dir"], "meta.json")))
    assert meta["compacted_from"] == os.path.realpath(durable["ix_dir"])
    assert meta["live_rows"] == r and meta["deleted"] == int(sum(bin(int(w)).count("1") for w in ol.deleted()))
    assert all(meta[k] == v for k, v in src.items())
    for f in ("centroids.f32", "rotation.f32", "tier2.sigma", "tier2.bin", "tier2.norm"):
        p = os.path.join(out, f)
        assert os.path.islink(p) and os.path.realpath(p) == os.path.realpath(os.path.join(durable["ix_dir"], f))
    with pytest.raises(SieveError, match="new directory"):
        durable["live"].compact(durable["ix_dir"])


def test_compacted_index_searches_like_the_live_state(small, durable, compacted):
    """The compacted directory opened from its files (Index::open + Live::open with the row
    segment): its searches equal the live state's and the oracle's over the same merged lists."""
    out = compacted["out"]
    ix3 = Index.open(out, durable["ing_dir"])
    with pytest.raises(SieveError, match="live segment"):
        ix3.search(small.q768, small.q256, Params(nprobe=8, cand=100, k=10, rerank=Rerank.FULL))
    with pytest.raises(SieveError, match="sieve_live_open_dir"):
        Live(ix3, max_rows=10)
    live3 = Live(ix3, max_rows=3000, index_dir=out)
    assert live3.info()["rows"] == durable["olive"].rows() and live3.info()["slots"] == 0
    m = compacted["m"]
    h = dict(small.host, block_off=m["block_off"], blocks=m["blocks"], ids=m["ids"], factors=m["factors"])
    omerged = O.Index(h)
    oseg = O.Live(omerged, soar_lambda=1.0)
    rd = lambda f, dt: np.fromfile(os.path.join(out, f), dt)
    r = durable["olive"].rows()
    oseg.load_segment(rd("live.rows.i8", np.int8).reshape(r, 768), rd("live.norm256.f32", np.float32),
                      rd("live.t2.bin", np.uint8).reshape(r, 64), rd("live.t2.norm", np.float32),
                      rd("live.ext.u32", np.uint32), rd("live.deleted.u64", np.uint64))
    for cfg in ("12:400:full", "48:2000:full", "8:300:prefix", "12:500:tiered60:b10x1.0"):
        _, p = parse_config(cfg, small.k, k=10)
        s, rr, n = live3.search(small.q768, small.q256, p)
        ws, wr, wn = oseg.search_batch(small.q768, small.q256, p, threads=4)
        ls, lr, ln = durable["live"].search(small.q768, small.q256, p)
        for i in range(small.nq):
            assert_same_up_to_ties(rr[i, :n[i]], s[i, :n[i]], wr[i, :wn[i]], ws[i, :wn[i]])
            assert_same_up_to_ties(rr[i, :n[i]], s[i, :n[i]], lr[i, :ln[i]], ls[i, :ln[i]])
    # writes continue on the compacted index: new rows after the segment's
    v, _ = normalize_rows(small.mix.rows(small.n + 5000, 3))
    v = v.numpy().astype(np.float32)
    e = np.array([80_000_000, 90_000_000, 5], np.uint32)
    live3.insert(e, v)
    oseg.insert(e, v)
    for x in (80_000_000, 90_000_000, 5, 80_000_001):
        assert live3.current(x) == oseg.current(x)
    assert live3.current(90_000_000) == small.n + r + 1
    s, rr, n = live3.search(v, normalize_rows(torch.from_numpy(v[:, :256].copy()))[0].numpy(),
                            Params(nprobe=8, cand=300, k=5, rerank=Rerank.FULL))
    assert rr[:, 0].tolist() == e.tolist()
    live3.close()


def test_truncate_wal_empties_the_log(durable):
    wal = durable["d"] / "wal"
    assert wal.stat().st_size > 0
    durable["live"].truncate_wal()
    assert wal.stat().st_size == 0
    durable["live"].delete(np.array([80_000_200], np.uint32))
    assert wal.stat().st_size == 5


def test_compaction_of_an_index_from_arrays(small, tmp_path):
    """An index created from arrays has no directory: compaction writes centroids, rotation and
    tier-2 files itself and a meta.json with the live state's soar_lambda; the result opens."""
    ix = Index.create(small.host)
    live = Live(ix, max_rows=200, soar_lambda=1.0)
    v, _ = normalize_rows(small.mix.rows(small.n + 7000, 50))
    live.insert((70_000_000 + np.arange(50)).astype(np.uint32), v.numpy().astype(np.float32))
    live.delete(np.arange(0, 40, dtype=np.uint32))
    out = tmp_path / "c"
    st = live.compact(str(out))
    assert st["dropped"] == sum(1 for i in small.host["ids"] if i < 40) and st["slots"] > 0
    meta = json.load(open(out / "meta.json"))
    assert meta["soar_lambda"] == 1.0 and meta["compacted_from"] is None and meta["live_rows"] == 50
    assert np.array_equal(np.fromfile(out / "centroids.f32", np.float32), small.host["centroids"].ravel())
    assert np.array_equal(np.fromfile(out / "tier2.bin", np.uint8), small.host["tier2_codes"].ravel())
    ing = tmp_path / "ing"
    write_index(small.arrays, str(tmp_path / "unused"), str(ing))
    ix2 = Index.open(str(out), str(ing))
    live2 = Live(ix2, max_rows=200, index_dir=str(out))
    p = Params(nprobe=12, cand=400, k=10, rerank=Rerank.FULL)
    a = live.search(small.q768, small.q256, p)
    b = live2.search(small.q768, small.q256, p)
    for i in range(small.nq):
        assert_same_up_to_ties(b[1][i, :b[2][i]], b[0][i, :b[2][i]], a[1][i, :a[2][i]], a[0][i, :a[2][i]])
    live2.close()
    live.close()
