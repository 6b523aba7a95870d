"""The two lower bounds the GPU path's selection rests on, checked on random instances (CPU):

* routing (k_route_select, tile mode): with T0 the nprobe-th largest of the maxima of 16
  consecutive approximate scores, every score that the top nprobe or the BF16 error band
  (>= kth - 2.2 e - 1e-6 (1 + |kth|)) needs lies in a tile whose maximum reaches
  lim = T0 - 2.2 e - 1e-6 (1 + max(|T0|, |max|)), so collecting those tiles finds the same
  n-th score and band as a pass over all scores;
* the small-batch scan (SCAN_BOUND + k_scan_bound_select): the (cand + 1)-th best of the union
  of every chunk's best 16 admissible (estimate, slot) pairs lies strictly below the cand-th
  best admissible pair, so appending every pair above it keeps the exact top cand.

Scores are drawn with heavy ties (few distinct values) as well as continuous ones."""
import numpy as np
import pytest


def _f32(x):
    return np.asarray(x, np.float32)


@pytest.mark.parametrize("seed", range(40))
def test_route_tile_collection_holds_top_n_and_band(seed):
    rng = np.random.default_rng(seed)
    k = int(rng.integers(16, 3000))
    n = int(rng.integers(1, min(k // 16, 300) + 1))
    if seed % 3 == 0:
        s = _f32(rng.integers(-20, 20, k) / 8.0)  # many ties
    else:
        s = _f32(rng.standard_normal(k))
    e = np.float32(rng.choice([0.0, 1e-3, 0.05]))
    # reference: a pass over every score
    kth = np.sort(s)[::-1][n - 1]
    thr = np.float32(kth - np.float32(2.2) * e - np.float32(1e-6) * (np.float32(1.0) + abs(kth)))
    band = set(np.nonzero(s >= thr)[0].tolist())
    # tile mode
    nt = (k + 15) // 16
    tmax = _f32([s[16 * t:16 * t + 16].max() for t in range(nt)])
    t0 = np.sort(tmax)[::-1][n - 1]
    top = tmax.max()
    lim = np.float32(t0 - np.float32(2.2) * e - np.float32(1e-6) * (np.float32(1.0) + max(abs(t0), abs(top))))
    tiles = np.nonzero(tmax >= lim)[0]
    got = np.concatenate([np.arange(16 * t, min(16 * t + 16, k)) for t in tiles])
    got = got[s[got] >= lim]
    assert len(got) >= n
    kth_t = np.sort(s[got])[::-1][n - 1]
    assert kth_t == kth                         # the same n-th score
    assert band <= set(got.tolist())            # every band member collected


@pytest.mark.parametrize("seed", range(40))
def test_small_batch_threshold_is_below_the_cand_th(seed):
    rng = np.random.default_rng(seed)
    slots = int(rng.integers(200, 40000))
    cand = int(rng.choice([8, 37, 300, 1000]))
    if seed % 3 == 0:
        key = rng.integers(0, 50, slots).astype(np.uint64)  # heavy ties on the estimate key
    else:
        key = rng.integers(0, 1 << 32, slots).astype(np.uint64)
    ok = rng.random(slots) > rng.choice([0.0, 0.3, 0.9])  # admissible (id present, not hidden)
    pair = (key << np.uint64(32)) | (~np.arange(slots, dtype=np.uint64) & np.uint64(0xFFFFFFFF))
    adm = np.sort(pair[ok])[::-1]
    # chunks of 32 blocks x 32 slots, each keeping its best min(16, cand) admissible pairs
    union = []
    for c0 in range(0, slots, 1024):
        p = np.sort(pair[c0:c0 + 1024][ok[c0:c0 + 1024]])[::-1]
        union.extend(p[:min(16, cand)].tolist())
    union = np.sort(np.array(union, np.uint64))[::-1]
    w = union[cand] if len(union) >= cand + 1 else np.uint64(0)
    if len(adm) >= cand:
        v = adm[cand - 1]  # the cand-th best admissible pair
        assert w < v
    kept = np.sort(pair[ok & (pair > w)])[::-1][:cand]
    assert np.array_equal(kept, adm[:cand])
