"""The oracle's Live restatement (src/live.rs): delta lists, inserts encoded as the build
encodes a row, tombstones, upserts, external ids, and live search against brute force."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from sieve_amd import Params, Rerank
from sieve_amd.build import normalize_rows


def test_delta_list_matches_build_layout():
    """live.rs:493-520: 40 pushes fill one block and start a second in pack_block's layout."""
    i = np.arange(40, dtype=np.uint32)
    codes = np.array([(np.full(32, (k * 37) & 255, np.uint8) ^ np.uint8(k)) for k in range(40)], np.uint8)
    f = np.stack([i.astype(np.float32), -i.astype(np.float32)], 1)
    blocks, fac, ids = O.delta_list_fixture(codes, f, i)
    assert blocks.shape == (2, 1024)
    want = O.pack_block(codes[32:], 64)
    assert np.array_equal(blocks[1], want.ravel())
    assert ids[39] == 39 and ids[40] == O.PAD
    assert fac[39].tolist() == [39.0, -39.0]
    for k in range(40):
        assert np.array_equal(O.unpack_slot(blocks[k // 32], k % 32), codes[k])


def _corpus_vectors(small, rows):
    """The unit rows ingest quantised (ingest_batch normalises each generated row)."""
    out = []
    for r in rows:
        v, _ = normalize_rows(small.mix.rows(int(r), 1))
        out.append(v[0].numpy())
    return np.stack(out).astype(np.float32)


def _build_slots(small):
    """row -> {cluster: (code 32 B, [mul, err])} of the built index."""
    h = small.host
    bo = h["block_off"].astype(np.int64)
    cl_of_block = np.repeat(np.arange(small.k), np.diff(bo))
    return h, bo, cl_of_block


def test_insert_encodes_like_the_build(small, oracle_index):
    """live.rs:2-6: an insert is encoded exactly as the build encodes a row. Re-inserting corpus
    rows must reproduce their int8 row, prefix norm, tier-2 code and, wherever the assignment
    picks the same clusters, their slots' RaBitQ codes and factors bit for bit."""
    live = O.Live(oracle_index, soar_lambda=1.0)
    rows = np.arange(0, small.n, small.n // 64)[:64]
    vecs = _corpus_vectors(small, rows)
    ext = (10_000_000 + np.arange(len(rows))).astype(np.uint32)
    live.insert(ext, vecs)
    assert live.rows() == len(rows)
    h, bo, cl_of_block = _build_slots(small)
    ids = h["ids"]
    same_pair = 0
    for j, r in enumerate(rows):
        row, norm, t2, t2n, e = live.row(j)
        assert e == ext[j]
        assert np.array_equal(row, h["rerank"][r])
        assert np.float32(norm).view(np.uint32) == np.float32(h["norm256"][r]).view(np.uint32)
        assert np.array_equal(t2, h["tier2_codes"][r])
        assert np.float32(t2n).view(np.uint32) == np.float32(h["tier2_norm"][r]).view(np.uint32)
        built = {}
        for s in np.nonzero(ids == r)[0]:
            c = int(cl_of_block[s // 32])
            code = O.unpack_slot(h["blocks"][s // 32], s % 32)
            built[c] = (code, h["factors"].reshape(-1, 2)[s])
        mine = {}
        for c in range(small.k):
            codes, fac, dids = live.delta_list(c)
            for k in np.nonzero(dids == small.n + j)[0]:
                mine[c] = (codes[k], fac[k])
        assert len(mine) == 2
        if set(mine) == set(built):
            same_pair += 1
        for c in set(mine) & set(built):
            assert np.array_equal(mine[c][0], built[c][0])
            assert np.array_equal(mine[c][1].view(np.uint32), built[c][1].view(np.uint32))
    assert same_pair >= 0.9 * len(rows)


class Model:
    """Independent bookkeeping of the live row set: internal id -> (ext, int8 row)."""

    def __init__(self, small):
        self.n = small.n
        self.alive = {r: (r, small.host["rerank"][r]) for r in range(small.n)}
        self.next = small.n
        self.by_ext = {r: r for r in range(small.n)}

    def insert(self, ext, rows):
        for e, row in zip(ext, rows):
            old = self.by_ext.get(int(e))
            if old is not None:
                del self.alive[old]
            self.alive[self.next] = (int(e), row)
            self.by_ext[int(e)] = self.next
            self.next += 1

    def delete(self, ext):
        for e in ext:
            old = self.by_ext.pop(int(e), None)
            if old is not None:
                del self.alive[old]

    def brute(self, small, q768, k):
        """Exact int8 top-k over the live rows (the full rerank's scores), as external ids."""
        quant = small.host["quant"]
        w = (q768.astype(np.float32) / quant).astype(np.float32)
        wq, s = O.query_i8(w)
        inv = np.float32(np.float32(1.0) / np.float32(s))
        ids = np.array(sorted(self.alive), np.int64)
        mat = np.stack([self.alive[i][1] for i in ids]).astype(np.int64)
        dots = mat @ wq.astype(np.int64)
        sc = (dots.astype(np.float32) * inv).astype(np.float32)
        order = np.lexsort((ids, -sc))[:k]
        return sc[order], [self.alive[int(ids[o])][0] for o in order]


def _same_up_to_ties(got_s, got_r, want_s, want_r):
    cut = np.float32(want_s[-1])
    g = {int(r) for s, r in zip(got_s, got_r) if s > cut}
    w = {int(r) for s, r in zip(want_s, want_r) if s > cut}
    return g == w and len(got_r) == len(want_r)


def test_live_search_matches_brute_force(small, oracle_index):
    """With every slot admitted (all clusters probed, cand above the slot count, full rerank)
    the live search is the exact int8 top-k over the live rows: inserts are found through the
    delta lists, deletes and replaced rows never come back, ids are external."""
    live = O.Live(oracle_index, soar_lambda=1.0)
    model = Model(small)
    rng = np.random.default_rng(3)
    new = small.mix.rows(small.n, 300)
    new, _ = normalize_rows(new)
    new = new.numpy().astype(np.float32)
    ext_new = (50_000_000 + np.arange(300)).astype(np.uint32)
    up = rng.choice(small.n, 20, replace=False).astype(np.uint32)       # upserts of corpus ids
    up_vecs = new[:20] * np.float32(-1.0)                                 # other vectors, unit length
    quant = small.host["quant"]
    enc = lambda vs: [O.int8_encode(v, quant) for v in vs]
    live.insert(ext_new, new)
    model.insert(ext_new, enc(new))
    live.insert(up, up_vecs)
    model.insert(up, enc(up_vecs))
    dels = np.concatenate([rng.choice(small.n, 50, replace=False), ext_new[:10], [4_000_000_000]]).astype(np.uint32)
    live.delete(dels)
    model.delete(dels)
    assert live.rows() == 320
    for e in up:
        assert live.current(e) is not None and live.current(e) >= small.n
    assert live.current(int(ext_new[0])) is None
    # queries: the workload's, plus some of the inserted vectors themselves
    q768 = np.concatenate([small.q768[:12], new[10:16]])
    q256 = np.concatenate([small.q256[:12], normalize_rows(torch.from_numpy(new[10:16, :256].copy()))[0].numpy()])
    p = Params(nprobe=small.k, cand=200_000, k=10, rerank=Rerank.FULL)
    s, r, n = live.search_batch(q768, q256, p, threads=4)
    for i in range(len(q768)):
        ws, wr = model.brute(small, q768[i], 10)
        assert _same_up_to_ties(s[i, :n[i]], r[i, :n[i]], ws, wr), i
    # an inserted vector finds itself first
    for i in range(12, 18):
        assert r[i, 0] == ext_new[10 + i - 12]


def test_delete_of_unknown_id_and_bitset(small, oracle_index):
    live = O.Live(oracle_index, soar_lambda=1.0)
    live.delete(np.array([small.n + 5, 4_000_000_000], np.uint32))
    assert live.deleted().sum() == 0
    live.delete(np.array([130], np.uint32))
    w = live.deleted()
    assert (w[130 >> 6] >> np.uint64(130 & 63)) & np.uint64(1) == 1 and int(w.sum()) == 1 << (130 & 63)
    assert live.current(130) is None and live.current(129) == 129
