"""Pins the oracle's route / candidates / rerank (src/ivf.rs:587-815) against brute force on a
small seeded index: float64 probe sets, full sorts instead of the amortised TopK, and int64
dot products. Runs on the CPU."""
import numpy as np
import pytest

from oracle import oracle as O
from sieve_amd import Params, Rerank

PAD = 0xFFFFFFFF


def test_route_probe_set_matches_float64_and_f32_scores(small, oracle_index):
    h = small.host
    C = h["centroids"].reshape(-1, 256)
    for i in range(small.nq):
        q = small.q256[i]
        p = Params(nprobe=7, cand=100)
        s, c = oracle_index.route(q, p)
        assert len(c) == 7
        exact = C.astype(np.float64) @ q.astype(np.float64)
        order = np.argsort(-exact, kind="stable")
        gap = exact[order[6]] - exact[order[7]]
        if gap > 1e-5:
            assert set(c.tolist()) == set(order[:7].tolist())
        seq = O.gemm_abt(q[None, :], C)[0]
        assert np.array_equal(s, seq[c])
        assert np.all(np.diff(s) <= 0)


def test_route_margin_and_budget(small, oracle_index):
    h = small.host
    bo = h["block_off"].astype(np.int64)
    slots = (bo[1:] - bo[:-1]) * 32
    q = small.q256[3]
    s_all, c_all = oracle_index.route(q, Params(nprobe=20, cand=10))
    s, c = oracle_index.route(q, Params(nprobe=20, cand=10, margin=0.05))
    assert np.array_equal(c, c_all[s_all >= np.float32(s_all[0] - np.float32(0.05))])
    b = int(slots[c_all[:3]].sum()) + 1
    s, c = oracle_index.route(q, Params(nprobe=20, cand=10, budget=b))
    assert np.array_equal(c, c_all[:3])
    s, c = oracle_index.route(q, Params(nprobe=20, cand=10, budget=1))
    assert np.array_equal(c, c_all[:1])  # the best cluster is always scanned


def brute_candidates(small, q, probe_s, probe_c, cand, hidden=None):
    """Every probed slot's estimate, top-`cand` slots by a full sort, dedup by row."""
    h = small.host
    bo = h["block_off"].astype(np.int64)
    tables, scale, bias, a, b = O.query_new(h["rotation"], q)
    a, b = np.float32(a), np.float32(b)
    rows_all, est_all, err_all = [], [], []
    for cent, c in zip(probe_s, probe_c):
        b0, b1 = bo[c], bo[c + 1]
        sc = O.scan(tables, h["blocks"][b0:b1].ravel()).astype(np.float32)
        f = h["factors"].reshape(-1, 2)[b0 * 32:b1 * 32]
        ids = h["ids"][b0 * 32:b1 * 32]
        est = (np.float32(cent) + f[:, 0] * (a * sc + b)).astype(np.float32)
        keep = ids != PAD
        if hidden is not None:
            keep &= ~np.isin(ids, hidden)
        rows_all.append(ids[keep]); est_all.append(est[keep]); err_all.append(f[keep, 1])
    rows = np.concatenate(rows_all); est = np.concatenate(est_all); err = np.concatenate(err_all)
    order = np.argsort(-est, kind="stable")
    kth = est[order[cand - 1]] if len(order) >= cand else -np.inf
    top = order[:cand]
    d = {}
    for i in top:
        r = int(rows[i])
        if r not in d or est[i] > d[r][0]:
            d[r] = (est[i], err[i])
    return d, kth


def test_candidates_match_brute_force(small, oracle_index):
    for i in range(0, small.nq, 3):
        q = small.q256[i]
        p = Params(nprobe=6, cand=150)
        ps, pc = oracle_index.route(q, p)
        est, rows, errs, scanned = oracle_index.candidates(q, ps, pc, p)
        assert np.all(np.diff(rows.astype(np.int64)) > 0)  # row order, deduplicated
        want, kth = brute_candidates(small, q, ps, pc, p.cand)
        bo = small.host["block_off"].astype(np.int64)
        assert scanned == int(sum((bo[c + 1] - bo[c]) * 32 for c in pc))
        got = {int(r): (e, x) for r, e, x in zip(rows, est, errs)}
        # identical up to ties at the cand-th estimate
        strict_w = {r for r, v in want.items() if v[0] > kth}
        strict_g = {r for r, v in got.items() if v[0] > kth}
        assert strict_w == strict_g
        for r in strict_g:
            assert got[r][0] == want[r][0] and got[r][1] == want[r][1]


def test_candidates_respect_tombstones_and_allowlist(small, oracle_index):
    q = small.q256[5]
    p = Params(nprobe=6, cand=150)
    ps, pc = oracle_index.route(q, p)
    est, rows, _, _ = oracle_index.candidates(q, ps, pc, p)
    dead = rows[::3].astype(np.uint32)
    deleted = np.zeros((small.n + 63) // 64, np.uint64)
    for r in dead:
        deleted[r >> 6] |= np.uint64(1) << np.uint64(r & 63)
    v = O.make_view(deleted=deleted)
    est2, rows2, _, _ = oracle_index.candidates(q, ps, pc, p, view=v)
    assert not set(rows2.tolist()) & set(dead.tolist())
    want, kth = brute_candidates(small, q, ps, pc, p.cand, hidden=dead)
    assert {r for r, e in want.items() if e[0] > kth} == {int(r) for r, e in zip(rows2, est2) if e > kth}
    allow = np.zeros_like(deleted)
    ok = np.arange(0, small.n, 7)
    for r in ok:
        allow[r >> 6] |= np.uint64(1) << np.uint64(r & 63)
    est3, rows3, _, _ = oracle_index.candidates(q, ps, pc, p, view=O.make_view(allow=allow))
    assert set(rows3.tolist()) <= set(ok.tolist())


def _int8_scores(small, q, rows, dims):
    h = small.host
    w = (q[:dims] / h["quant"][:dims]).astype(np.float32)
    wq, s = O.query_i8(w)
    inv = np.float32(np.float32(1.0) / np.float32(s))
    R = h["rerank"].reshape(-1, 768)[rows, :dims].astype(np.int64)
    dots = R @ wq.astype(np.int64)
    return dots.astype(np.float32) * inv


def _top(scores, ids, k):
    order = sorted(range(len(ids)), key=lambda i: (-scores[i], ids[i]))[:k]
    return [int(ids[i]) for i in order], [scores[i] for i in order]


@pytest.mark.parametrize("mode", ["full", "prefix", "none", "tiered", "ssd", "tiered_noT2", "pruned"])
def test_rerank_modes_match_brute_force(small, oracle_index, mode):
    h = small.host
    for i in range(0, small.nq, 5):
        q7, q2 = small.q768[i], small.q256[i]
        base = Params(nprobe=6, cand=200, k=10)
        ps, pc = oracle_index.route(q2, base)
        est, rows, errs, _ = oracle_index.candidates(q2, ps, pc, base)
        rows64 = rows.astype(np.int64)
        if mode == "full":
            p = base
            sc = _int8_scores(small, q7, rows64, 768)
            want = _top(sc, rows64, 10)
        elif mode == "pruned":
            p = base.with_(prune=(10, 0.8))
            lo = (est - np.float32(0.8) * errs).astype(np.float32)
            tau = np.sort(lo)[::-1][9]
            keep = (est + np.float32(0.8) * errs).astype(np.float32) >= tau
            sc = _int8_scores(small, q7, rows64[keep], 768)
            want = _top(sc, rows64[keep], 10)
        elif mode == "none":
            p = base.with_(rerank=Rerank.NONE)
            want = _top(est, rows64, 10)
        elif mode == "prefix":
            p = base.with_(rerank=Rerank.PREFIX)
            sc = (_int8_scores(small, q2, rows64, 256) / h["norm256"][rows64]).astype(np.float32)
            want = _top(sc, rows64, 10)
        elif mode == "ssd":
            p = base.with_(rerank=Rerank.ssd(40))
            ids40, _ = _top(est, rows64, 40)
            ids40 = np.array(ids40)
            want = _top(_int8_scores(small, q7, ids40, 768), ids40, 10)
        elif mode == "tiered":
            p = base.with_(rerank=Rerank.tiered(40))
            tab = O.tier2_tables(h["tier2_sigma"], q2)
            codes = h["tier2_codes"].reshape(-1, 64)
            t2 = np.array([O.tier2_score(tab, codes[r], h["tier2_norm"][r]) for r in rows64], np.float32)
            ids40 = np.array(_top(t2, rows64, 40)[0])
            want = _top(_int8_scores(small, q7, ids40, 768), ids40, 10)
        else:  # tiered without a 2-bit tier: prefix rerank to 40, then full
            p = base.with_(rerank=Rerank.tiered(40))
            sc = (_int8_scores(small, q2, rows64, 256) / h["norm256"][rows64]).astype(np.float32)
            ids40 = np.array(_top(sc, rows64, 40)[0])
            want = _top(_int8_scores(small, q7, ids40, 768), ids40, 10)
        ix = oracle_index
        if mode == "tiered_noT2":
            hh = dict(h)
            hh.pop("tier2_sigma"); hh.pop("tier2_codes"); hh.pop("tier2_norm")
            ix = O.Index(hh)
        s, r, _ = ix.rerank(q7, q2, est, rows, errs, p)
        assert r.tolist() == want[0]
        assert np.array_equal(s, np.array(want[1], np.float32))
