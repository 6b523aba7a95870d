"""The oracle's restatement of Live's durability (src/live.rs): the log's record layout
(src/live.rs:131-135, :276-306), replay (src/live.rs:209-249: consecutive inserts as a batch,
torn tails cut, unknown ops refused), truncate_wal (src/live.rs:470-479) and the merge of
compaction (src/live.rs:395-468), pinned against bytes written by hand, the same writes applied
without a log, a numpy rebuild of the merged lists, and searches of the compacted index."""
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O
from sieve_amd import Params, Rerank
from sieve_amd.build import normalize_rows


def _ops(small):
    rng = np.random.default_rng(21)
    new, _ = normalize_rows(small.mix.rows(small.n, 260))
    new = new.numpy().astype(np.float32)
    ext_new = (60_000_000 + np.arange(260)).astype(np.uint32)
    up = rng.choice(small.n, 15, replace=False).astype(np.uint32)
    dels = np.concatenate([rng.choice(small.n, 40, replace=False), ext_new[3:9], [4_000_000_001]]).astype(np.uint32)
    return [("insert", ext_new[:150], new[:150]), ("insert", up, -new[150:165]), ("delete", dels),
            ("insert", ext_new[150:], new[150:]), ("delete", ext_new[200:203]), ("insert", ext_new[:2], new[165:167])]


def _apply(live, ops):
    for op in ops:
        if op[0] == "insert":
            live.insert(op[1], op[2])
        else:
            live.delete(op[1])


def _state(small, live):
    rows = [live.row(j) for j in range(live.rows())]
    lists = [live.delta_list(c) for c in range(small.k)]
    return rows, lists, live.deleted()


def _same_state(small, a, b):
    ra, la, da = _state(small, a)
    rb, lb, db = _state(small, b)
    assert len(ra) == len(rb)
    for x, y in zip(ra, rb):
        assert np.array_equal(x[0], y[0]) and np.float32(x[1]).tobytes() == np.float32(y[1]).tobytes()
        assert np.array_equal(x[2], y[2]) and x[4] == y[4]
    for (ca, fa, ia), (cb, fb, ib) in zip(la, lb):
        assert np.array_equal(ca, cb) and np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
        assert np.array_equal(ia, ib)
    assert np.array_equal(da, db)


def test_log_records_have_the_reference_layout(small, oracle_index, tmp_path):
    """An insert record is op 1, the external id (u32 LE), the prefix norm (f32 LE) and the 768
    int8 row bytes; a delete is op 2 and the id (src/live.rs:131-135, :284-290, :298-302)."""
    wal = tmp_path / "wal"
    live = O.Live(oracle_index, soar_lambda=1.0)
    assert live.open_wal(wal) == dict(inserts=0, deletes=0, torn=0)
    v, _ = normalize_rows(small.mix.rows(small.n + 7, 1))
    v = v.numpy().astype(np.float32)
    live.insert(np.array([0x01020304], np.uint32), v)
    live.delete(np.array([7, 0xA0B0C0D0], np.uint32))
    b = wal.read_bytes()
    assert len(b) == 1 + 4 + 4 + 768 + 2 * 5
    assert b[:5] == bytes([1, 4, 3, 2, 1])
    # the prefix norm: sqrt of the sequential f32 sum of squares of the first 256 values
    s = np.float32(-0.0)
    for x in v[0, :256]:
        s = np.float32(s + np.float32(x * x))
    assert b[5:9] == struct.pack("<f", np.sqrt(s, dtype=np.float32))
    assert b[9:777] == O.int8_encode(v[0], small.host["quant"]).tobytes()
    assert b[777:] == bytes([2, 7, 0, 0, 0, 2, 0xD0, 0xC0, 0xB0, 0xA0])


def test_replay_reproduces_the_state(small, oracle_index, tmp_path):
    """A state rebuilt from the log alone equals the one the writes built (rows, slots,
    tombstones, external ids) and searches the same; replay then appends to the same log."""
    wal = tmp_path / "wal"
    a = O.Live(oracle_index, soar_lambda=1.0)
    a.open_wal(wal)
    ops = _ops(small)
    _apply(a, ops)
    direct = O.Live(oracle_index, soar_lambda=1.0)
    _apply(direct, ops)
    _same_state(small, a, direct)
    b = O.Live(oracle_index, soar_lambda=1.0)
    st = b.open_wal(wal)
    n_ins = sum(len(o[1]) for o in ops if o[0] == "insert")
    n_del = sum(len(o[1]) for o in ops if o[0] == "delete")
    assert st == dict(inserts=n_ins, deletes=n_del, torn=0)
    _same_state(small, a, b)
    for e in [int(x) for x in ops[2][1]] + [60_000_000, 60_000_001, 60_000_201, 60_000_250]:
        assert a.current(e) == b.current(e)
    p = Params(nprobe=16, cand=600, k=10, rerank=Rerank.FULL)
    sa = a.search_batch(small.q768[:16], small.q256[:16], p)
    sb = b.search_batch(small.q768[:16], small.q256[:16], p)
    for x, y in zip(sa, sb):
        assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
    # the reopened log is appended to
    size = wal.stat().st_size
    b.delete(np.array([60_000_100], np.uint32))
    assert wal.stat().st_size == size + 5


@pytest.mark.parametrize("cut", [1, 5, 300, 776])
def test_replay_cuts_a_torn_tail(small, oracle_index, tmp_path, cut):
    wal = tmp_path / "wal"
    a = O.Live(oracle_index, soar_lambda=1.0)
    a.open_wal(wal)
    ops = _ops(small)[:3]
    _apply(a, ops)
    good = wal.read_bytes()
    v, _ = normalize_rows(small.mix.rows(small.n + 900, 1))
    a.insert(np.array([99], np.uint32), v.numpy().astype(np.float32))
    torn = wal.read_bytes()[:len(good) + cut]
    wal.write_bytes(torn)
    b = O.Live(oracle_index, soar_lambda=1.0)
    st = b.open_wal(wal)
    assert st["torn"] == cut and wal.stat().st_size == len(good)
    direct = O.Live(oracle_index, soar_lambda=1.0)
    _apply(direct, ops)
    _same_state(small, b, direct)


def test_replay_refuses_an_unknown_op(small, oracle_index, tmp_path):
    wal = tmp_path / "wal"
    wal.write_bytes(bytes([2, 5, 0, 0, 0, 9, 1, 1]))
    b = O.Live(oracle_index, soar_lambda=1.0)
    with pytest.raises(ValueError, match="bad op at byte 5"):
        b.open_wal(wal)
    assert wal.stat().st_size == 8


def test_truncate_wal_empties_the_log(small, oracle_index, tmp_path):
    wal = tmp_path / "wal"
    a = O.Live(oracle_index, soar_lambda=1.0)
    a.open_wal(wal)
    a.delete(np.array([3, 4], np.uint32))
    assert wal.stat().st_size == 10
    a.truncate_wal()
    assert wal.stat().st_size == 0
    a.delete(np.array([5], np.uint32))
    assert wal.stat().st_size == 5


def _merge_by_numpy(small, live):
    """The merged lists rebuilt from the base arrays and the delta lists: per cluster the base
    slots in block order, then the delta slots, PAD and tombstoned ones dropped, 32 per block."""
    h = small.host
    bo = h["block_off"].astype(np.int64)
    words = live.deleted()

    def dead(i):
        w = int(i) >> 6
        return w < len(words) and (int(words[w]) >> (int(i) & 63)) & 1 == 1
    off, blocks, fac, ids = [0], [], [], []
    dropped = 0
    for c in range(small.k):
        codes, fs, rs = [], [], []
        for s in range(bo[c] * 32, bo[c + 1] * 32):
            r = h["ids"][s]
            if r == O.PAD:
                continue
            if dead(r):
                dropped += 1
                continue
            codes.append(O.unpack_slot(h["blocks"][s // 32], s % 32))
            fs.append(h["factors"][s])
            rs.append(r)
        dc, df, di = live.delta_list(c)
        for j in range(len(di)):
            if dead(di[j]):
                dropped += 1
                continue
            codes.append(dc[j])
            fs.append(df[j])
            rs.append(di[j])
        for o in range(0, len(rs), 32):
            m = min(32, len(rs) - o)
            blocks.append(O.pack_block(np.stack(codes[o:o + m]), 64).ravel())
            f = np.zeros((32, 2), np.float32)
            f[:m] = np.stack(fs[o:o + m])
            i = np.full(32, O.PAD, np.uint32)
            i[:m] = rs[o:o + m]
            fac.append(f)
            ids.append(i)
        off.append(len(blocks))
    return (np.array(off, np.uint64), np.stack(blocks), np.concatenate(fac), np.concatenate(ids), dropped)


def test_compaction_merge_matches_numpy(small, oracle_index):
    live = O.Live(oracle_index, soar_lambda=1.0)
    _apply(live, _ops(small))
    got = live.compact()
    off, blocks, fac, ids, dropped = _merge_by_numpy(small, live)
    assert np.array_equal(got["block_off"], off)
    assert np.array_equal(got["blocks"], blocks)
    assert np.array_equal(got["factors"].view(np.uint32), fac.view(np.uint32))
    assert np.array_equal(got["ids"], ids)
    assert got["dropped"] == dropped and got["slots"] == int((ids != O.PAD).sum())


def _segment(small, live):
    rows = [live.row(j) for j in range(live.rows())]
    return (np.stack([r[0] for r in rows]), np.array([r[1] for r in rows], np.float32),
            np.stack([r[2] for r in rows]), np.array([r[3] for r in rows], np.float32),
            np.array([r[4] for r in rows], np.uint32), live.deleted())


def test_compacted_index_searches_like_the_live_state(small, oracle_index):
    """The compacted lists plus the row segment (a fresh Live over the merged index) answer
    every query as the live state does: same rows, same int8 scores, up to ties at the cut."""
    live = O.Live(oracle_index, soar_lambda=1.0)
    _apply(live, _ops(small))
    m = live.compact()
    h = dict(small.host, block_off=m["block_off"], blocks=m["blocks"], ids=m["ids"], factors=m["factors"])
    merged = O.Index(h)
    seg = O.Live(merged, soar_lambda=1.0)
    seg.load_segment(*_segment(small, live))
    for e in [int(x) for x in _ops(small)[2][1]] + [60_000_000, 60_000_150, 60_000_259]:
        assert seg.current(e) == live.current(e)
    for cfg in (Params(nprobe=12, cand=500, k=10, rerank=Rerank.FULL),
                Params(nprobe=small.k, cand=200_000, k=10, rerank=Rerank.FULL)):
        a = live.search_batch(small.q768, small.q256, cfg, threads=4)
        b = seg.search_batch(small.q768, small.q256, cfg, threads=4)
        for i in range(small.nq):
            ga, gb = a[2][i], b[2][i]
            assert ga == gb
            cut = np.float32(a[0][i, ga - 1])
            sa = {int(r): s for s, r in zip(a[0][i, :ga], a[1][i, :ga]) if s > cut}
            sb = {int(r): s for s, r in zip(b[0][i, :gb], b[1][i, :gb]) if s > cut}
            assert sa.keys() == sb.keys(), i
            assert all(np.float32(sa[r]).tobytes() == np.float32(sb[r]).tobytes() for r in sa)
