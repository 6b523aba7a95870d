"""The index build on the GPU (sieve_amd.build over sieve_build_kmeans, sieve_build_assign and
sieve_build_tier2_fit) against the oracle's restatement of ivf::build (src/ivf.rs:264-395) on the
same corpus bytes: the sample drawn with Rng(42), the k-means centroids, the tier-2 sigma and
codes, every row's (primary, SOAR secondary) pair, the cluster-major layout and every slot's
RaBitQ code and factors, bit for bit."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PAD = 0xFFFFFFFF


def say(msg):
    print(f"[gpu-build] {msg}", flush=True)


@pytest.fixture(scope="module")
def built():
    import torch
    from sieve_amd import build as B, synth
    n, k, sample_n, iters = 60_000, 200, 12_000, 5
    mix = synth.Mixture(synth.MixtureSpec(components=48, seed=5), device="cuda")
    corpus = B.ingest(mix, n, batch=1 << 15)
    say(f"building {n} rows into {k} clusters on the GPU")
    arrays = B.build(corpus, k, sample_n=sample_n, iters=iters, chunk=1 << 14)
    torch.cuda.synchronize()
    h = arrays.numpy()
    # Every row's normalised prefix rebuilt from its int8 row (Corpus::prefix_f32).
    R = h["rerank"].reshape(-1, 768)[:, :256].astype(np.float32)
    X = np.stack([O.normalize((r / h["quant"][:256]).astype(np.float32))[0] for r in R])
    rng = B.Rng(42)
    pick = np.array([rng.below(n) for _ in range(sample_n)])
    return dict(h=h, X=X, sample=X[pick], n=n, k=k, iters=iters)


def test_centroids_and_tier2_sigma_are_the_oracles(built):
    h, k = built["h"], built["k"]
    want = O.kmeans(built["sample"], k, built["iters"], seed=7, threads=16)
    got = h["centroids"].reshape(k, 256)
    bad = int(np.sum(np.any(got.view(np.uint32) != want.view(np.uint32), axis=1)))
    say(f"k-means: {bad} of {k} centroids differ from the oracle's")
    assert bad == 0
    sig = O.tier2_fit(built["sample"])
    assert np.array_equal(h["tier2_sigma"].view(np.uint32), sig.view(np.uint32))


def test_assignment_and_layout_are_the_oracles(built):
    h, X, n, k = built["h"], built["X"], built["n"], built["k"]
    cents = h["centroids"].reshape(k, 256)
    pairs = O.build_pairs(X, cents, 1.0, threads=16)
    bo = h["block_off"].astype(np.int64)
    ids = h["ids"]
    want_members = [[] for _ in range(k)]
    for r in range(n):
        for c in pairs[r]:
            want_members[int(c)].append(r)
    bad = 0
    for c in range(k):
        s = ids[bo[c] * 32:bo[c + 1] * 32]
        real = s[s != PAD]
        if real.tolist() != want_members[c] or len(s) != 32 * ((len(real) + 31) // 32):
            bad += 1
    say(f"assignment: {bad} of {k} cluster member lists differ from the oracle's")
    assert bad == 0


def test_slot_codes_factors_and_tier2_codes_are_the_oracles(built):
    h, X, k = built["h"], built["X"], built["k"]
    cents = h["centroids"].reshape(k, 256)
    signs = h["rotation"]
    crot = np.stack([O.rotation_apply(signs, c) for c in cents])
    bo = h["block_off"].astype(np.int64)
    f = h["factors"].reshape(-1, 2)
    bad = 0
    checked = 0
    for c in range(k):
        for s in range(bo[c] * 32, bo[c + 1] * 32, 3):
            row = h["ids"][s]
            code = O.unpack_slot(h["blocks"][s // 32], s % 32)
            if row == PAD:
                bad += int(code.any() or f[s, 0] != 0 or f[s, 1] != 0)
                continue
            want_code, want_f = O.rabitq_encode(O.rotation_apply(signs, X[row]), crot[c])
            bad += int(not np.array_equal(code, want_code) or not np.array_equal(f[s].view(np.uint32),
                                                                                  want_f.view(np.uint32)))
            checked += 1
    say(f"slots: {bad} of {checked} codes / factors differ")
    assert bad == 0
    for r in range(0, built["n"], 37):
        code, norm = O.tier2_encode(h["tier2_sigma"], X[r])
        assert np.array_equal(code, h["tier2_codes"].reshape(-1, 64)[r])
        assert np.float32(norm).view(np.uint32) == h["tier2_norm"][r].view(np.uint32)


def test_build_entry_points_from_host_arrays(built):
    """sieve_build_assign / _kmeans / _tier2_fit on host arrays: the same results as on device
    arrays (and the no-SOAR assignment)."""
    import ctypes
    from sieve_amd import abi
    L = abi.lib()
    X = np.ascontiguousarray(built["X"][:5000])
    cents = np.ascontiguousarray(built["h"]["centroids"].reshape(built["k"], 256))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    pairs = np.empty((5000, 2), np.uint32)
    abi.check(L.sieve_build_assign(p(X), 5000, p(cents), built["k"], 1.0, p(pairs)))
    assert np.array_equal(pairs, O.build_pairs(X, cents, 1.0, threads=16))
    abi.check(L.sieve_build_assign(p(X), 5000, p(cents), built["k"], float("nan"), p(pairs)))
    assert np.array_equal(pairs, O.build_pairs(X, cents, None, threads=16))
    km = np.empty((24, 256), np.float32)
    abi.check(L.sieve_build_kmeans(p(X), 5000, 24, 3, 9, p(km)))
    assert np.array_equal(km.view(np.uint32), O.kmeans(X, 24, 3, seed=9).view(np.uint32))
    sig = np.empty(256, np.float32)
    abi.check(L.sieve_build_tier2_fit(p(X), 5000, p(sig)))
    assert np.array_equal(sig.view(np.uint32), O.tier2_fit(X).view(np.uint32))


# ----------------------------------------------------------------- the 1B build ----------
@pytest.fixture(scope="module")
def l2built():
    import torch
    from sieve_amd.l2build import L2Mixture, build_l2_large
    n = 400_000
    mix = L2Mixture(600, seed=21, topics=8)
    base = torch.empty(n, 100, dtype=torch.int8, device="cuda")
    mix.fill(base)
    say(f"building the L2 index: {n} x 100 int8, 16 x 32 leaves")
    arr = build_l2_large(base, coarse=16, fine=32, iters=4, sample=100_000, chunk=1 << 16)
    torch.cuda.synchronize()
    return dict(arr=arr, base=base.cpu().numpy(), n=n)


def test_l2_training_is_the_oracles(l2built):
    """bigann::train on the GPU: the coarse cells and every cell's leaves bit for bit."""
    arr = l2built["arr"]
    coarse, leaves = O.l2_train(l2built["base"], 16, 32, 100_000, 4, threads=16)
    bad_c = int(np.sum(np.any(arr.coarse.cpu().numpy().view(np.uint32) != coarse.view(np.uint32), axis=1)))
    bad_l = int(np.sum(np.any(arr.leaves.cpu().numpy().view(np.uint32) != leaves.view(np.uint32), axis=1)))
    say(f"l2 train: {bad_c} of 16 coarse cells and {bad_l} of 512 leaves differ from the oracle's")
    assert bad_c == 0 and bad_l == 0


def test_l2_assignment_is_the_oracles(l2built):
    """Model::assign on the GPU: every row's leaf bit for bit (through the built layout)."""
    arr, n = l2built["arr"], l2built["n"]
    ids = arr.ids.cpu().numpy().view(np.uint32)
    bo = arr.block_off.cpu().numpy()
    leaf_of = np.full(n, -1, np.int64)
    for c in range(len(bo) - 1):
        seg = ids[bo[c] * 32:bo[c + 1] * 32]
        leaf_of[seg[seg != PAD]] = c
    want = O.l2_assign(l2built["base"].astype(np.float32), arr.coarse.cpu().numpy(), arr.leaves.cpu().numpy(), 32,
                       threads=16)
    bad = int(np.sum(want != leaf_of))
    say(f"l2 assign: {bad} of {n} rows in another leaf than the oracle's")
    assert bad == 0
