"""The oracle's k-means and build assignment (oracle.kmeans, oracle.build_pairs) pinned against
independent numpy restatements of src/ivf.rs:137-225 and src/blas.rs:47-53: every score a
left-to-right f32 sum of f32 products from -0.0, the first maximum, the members' sums in sample
order from +0.0, quant::normalize (src/quant.rs:5-13), splitmix64 draws (src/ivf.rs:120-135).
Duplicate sample rows force exact ties and empty clusters. CPU only."""
import numpy as np

from oracle import oracle as O

M64 = (1 << 64) - 1


class _Rng:
    def __init__(self, seed):
        self.s = seed

    def below(self, n):
        self.s = (self.s + 0x9E3779B97F4A7C15) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return (z ^ (z >> 31)) % n


def _scores(x, c):
    acc = np.full((x.shape[0], c.shape[0]), -0.0, np.float32)
    for d in range(x.shape[1]):
        acc = acc + x[:, d:d + 1] * c[None, :, d]
    return acc


def _normalize(v):
    ss = np.float32(-0.0)
    for t in v:
        ss = np.float32(ss + np.float32(t * t))
    n = np.sqrt(ss)
    if n > 0:
        v = (v * (np.float32(1.0) / n)).astype(np.float32)
    return v


def _kmeans(sample, k, iters, seed):
    n = sample.shape[0]
    rng = _Rng(seed)
    cents = np.stack([sample[rng.below(n)] for _ in range(k)]).astype(np.float32)
    for _ in range(iters):
        a = np.argmax(_scores(sample, cents), axis=1)  # the first maximum
        sums = np.zeros_like(cents)
        counts = np.zeros(k, np.int64)
        for i in range(n):
            counts[a[i]] += 1
            sums[a[i]] = sums[a[i]] + sample[i]
        for c in range(k):
            cents[c] = sample[rng.below(n)] if counts[c] == 0 else _normalize(sums[c].copy())
    return cents


def _unit_rows(rng, n):
    x = rng.standard_normal((n, 256)).astype(np.float32)
    return np.stack([O.normalize(r)[0] for r in x])


def test_kmeans_matches_numpy_restatement():
    rng = np.random.default_rng(3)
    sample = _unit_rows(rng, 600)
    want = _kmeans(sample, 12, 4, 7)
    got = O.kmeans(sample, 12, 4, seed=7)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_kmeans_ties_and_empty_clusters():
    # 9 distinct rows repeated: drawn centroids repeat, so rows tie between equal centroids
    # (the lower id wins) and the higher-id twins go empty and are reseeded.
    rng = np.random.default_rng(4)
    base = _unit_rows(rng, 9)
    sample = base[rng.integers(0, 9, 240)]
    want = _kmeans(sample, 14, 5, 11)
    got = O.kmeans(sample, 14, 5, seed=11)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    a = np.argmax(_scores(sample, want), axis=1)
    assert len(np.unique(a)) < 14  # some clusters empty in the last assignment


def test_kmeans_mutations_fail():
    # A plausible slip (sums from the last member backwards, or no renormalisation) changes bits.
    rng = np.random.default_rng(5)
    sample = _unit_rows(rng, 400)
    got = O.kmeans(sample, 8, 2, seed=7)
    rev = _kmeans(sample[::-1].copy(), 8, 2, 7)
    assert not np.array_equal(got.view(np.uint32), rev.view(np.uint32))


def test_build_pairs_match_assign_top_and_soar():
    rng = np.random.default_rng(6)
    for k in (10, 40):  # k < 16: the candidate list's tail reads cluster 0
        cents = _unit_rows(rng, k)
        rows = _unit_rows(rng, 150)
        pairs = O.build_pairs(rows, cents, 1.0)
        sc = _scores(rows, cents)
        m = min(16, k)
        for i in range(len(rows)):
            order = sorted(range(k), key=lambda j: (-sc[i, j], j))[:m]
            cand = np.zeros(16, np.uint32)
            cand[:m] = order
            assert pairs[i, 0] == order[0]
            assert pairs[i, 1] == O.soar_secondary(rows[i], cents, cand, 1.0)
        nosoar = O.build_pairs(rows, cents, None)
        assert np.array_equal(nosoar[:, 0], pairs[:, 0]) and np.all(nosoar[:, 1] == 0xFFFFFFFF)


# ----------------------------------------------------------------- the 1B build ----------
def _scores_l2(x, c):
    acc = np.full((x.shape[0], c.shape[0]), -0.0, np.float32)
    for d in range(x.shape[1]):
        acc = acc + x[:, d:d + 1] * c[None, :, d]
    cn = np.full(c.shape[0], -0.0, np.float32)
    for d in range(c.shape[1]):
        cn = cn + c[:, d] * c[:, d]
    return np.float32(2.0) * acc - cn[None, :]


def _kmeans_l2(sample, k, iters, seed):
    n = sample.shape[0]
    rng = _Rng(seed)
    cents = np.stack([sample[rng.below(n)] for _ in range(k)]).astype(np.float32)
    for _ in range(iters):
        a = np.argmax(_scores_l2(sample, cents), axis=1)
        sums = np.zeros((k, sample.shape[1]), np.float64)
        counts = np.zeros(k, np.int64)
        for i in range(n):
            counts[a[i]] += 1
            sums[a[i]] = sums[a[i]] + sample[i].astype(np.float64)
        for c in range(k):
            cents[c] = sample[rng.below(n)] if counts[c] == 0 else (sums[c] / np.float64(counts[c])).astype(np.float32)
    return cents


def _int8_rows(rng, n, centers=6):
    c = rng.integers(-80, 80, (centers, 100))
    return np.clip(c[rng.integers(0, centers, n)] + rng.integers(-9, 10, (n, 100)), -127, 127).astype(np.int8)


def test_kmeans_l2_matches_numpy_restatement():
    rng = np.random.default_rng(8)
    sample = _int8_rows(rng, 500).astype(np.float32)
    want = _kmeans_l2(sample, 9, 4, 3)
    got = O.kmeans_l2(sample, 9, 4, 3)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # duplicate rows: equal centroids tie (the lower id wins), their twins go empty
    dup = _int8_rows(rng, 7).astype(np.float32)[rng.integers(0, 7, 200)]
    assert np.array_equal(O.kmeans_l2(dup, 12, 3, 5).view(np.uint32), _kmeans_l2(dup, 12, 3, 5).view(np.uint32))


def test_kmeans_finds_separated_clusters():
    # bigann.rs:818-829 with its seed: rows alternate between centres +30 and -30.
    rng = _Rng(3)
    rows = np.array([[(30.0 if i % 2 == 0 else -30.0) + (rng.below(5) - 2.0) for _ in range(100)] for i in range(400)],
                    np.float32)
    cents = O.kmeans_l2(rows, 2, 5, 1)
    a, _ = O.top_l2(rows, cents, 1)
    assert np.all(a[0::2, 0] != a[1::2, 0])


def test_top_l2_matches_numpy():
    rng = np.random.default_rng(9)
    x = _int8_rows(rng, 60).astype(np.float32)
    c = rng.standard_normal((40, 100)).astype(np.float32) * 30
    ids, sc = O.top_l2(x, c, 4)
    s = _scores_l2(x, c)
    for i in range(len(x)):
        order = sorted(range(40), key=lambda j: (-s[i, j], j))[:4]
        assert ids[i].tolist() == order
        assert np.array_equal(sc[i].view(np.uint32), s[i, order].view(np.uint32))


def test_l2_train_and_assign_restate_the_reference():
    """train's sample, coarse cells and per-cell leaves, and Model::assign, against the numpy
    restatement on a small base."""
    rng = np.random.default_rng(10)
    base = _int8_rows(rng, 3000, centers=12)
    k1, k2, sample_n, iters = 4, 6, 900, 3
    coarse, leaves = O.l2_train(base, k1, k2, sample_n, iters)
    r = _Rng(42)
    idx = sorted(r.below(len(base)) for _ in range(sample_n))
    sample = base[idx].astype(np.float32)
    coarse_n = min(sample_n, k1 * 512)
    crow = sample[::sample_n // coarse_n][:coarse_n]
    want_coarse = _kmeans_l2(crow, k1, iters, 7)
    assert np.array_equal(coarse.view(np.uint32), want_coarse.view(np.uint32))
    a = np.argmax(_scores_l2(sample, want_coarse), axis=1)
    for c in range(k1):
        rows = sample[a == c]
        want = _kmeans_l2(rows, k2, iters, c) if len(rows) else np.repeat(want_coarse[c:c + 1], k2, 0)
        assert np.array_equal(leaves[c * k2:(c + 1) * k2].view(np.uint32), want.view(np.uint32))
    x = base[:400].astype(np.float32)
    got = O.l2_assign(x, coarse, leaves, k2)
    sc = _scores_l2(x, coarse)
    sl = _scores_l2(x, leaves)
    for i in range(len(x)):
        cells = sorted(sorted(range(k1), key=lambda j: (-sc[i, j], j))[:4])
        best, bs = 0, -np.inf
        for c in cells:
            j = max(range(k2), key=lambda j: (sl[i, c * k2 + j], -j))
            if sl[i, c * k2 + j] > bs:
                best, bs = c * k2 + j, sl[i, c * k2 + j]
        assert got[i] == best
