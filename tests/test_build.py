"""The index writer (sieve_amd.build) against the oracle's restatement of the reference's
ingest and build steps, on the small seeded workload. Runs on the CPU."""
import numpy as np
import pytest

from oracle import oracle as O

PAD = 0xFFFFFFFF


def test_ingest_matches_reference_encoding(small):
    h = small.host
    rows = small.mix.rows(0, 300).numpy()
    scale = O.int8_fit(np.stack([O.normalize(r)[0] for r in small.mix.rows(0, 8192).numpy()]))
    assert np.array_equal(scale.view(np.uint32), h["quant"].view(np.uint32))
    R = h["rerank"].reshape(-1, 768)
    for i in range(0, 300, 7):
        v, _ = O.normalize(rows[i])
        assert np.array_equal(O.int8_encode(v, h["quant"]), R[i])
        _, n256 = O.normalize(v[:256])
        assert np.float32(n256).view(np.uint32) == h["norm256"][i].view(np.uint32)


def _prefix(h, r):
    x = (h["rerank"].reshape(-1, 768)[r, :256].astype(np.float32) / h["quant"][:256]).astype(np.float32)
    return O.normalize(x)[0]


def test_slots_carry_reference_codes_and_factors(small):
    # Every stored (row, cluster) slot holds rabitq::encode of the rotated prefix against the
    # rotated centroid, bit for bit (src/ivf.rs:323-341, src/rabitq.rs:86-108).
    h = small.host
    bo = h["block_off"].astype(np.int64)
    C = h["centroids"].reshape(-1, 256)
    signs = h["rotation"]
    assert np.array_equal(signs, O.rotation_signs(11))
    f = h["factors"].reshape(-1, 2)
    rs = np.random.default_rng(1)
    for c in rs.choice(small.k, 8, replace=False):
        cr = O.rotation_apply(signs, C[c])
        for s in range(bo[c] * 32, bo[c + 1] * 32, 11):
            row = h["ids"][s]
            blk = h["blocks"][s // 32]
            code = O.unpack_slot(blk, s % 32)
            if row == PAD:
                assert not code.any() and f[s, 0] == 0 and f[s, 1] == 0
                continue
            want_code, want_f = O.rabitq_encode(O.rotation_apply(signs, _prefix(h, row)), cr)
            assert np.array_equal(code, want_code)
            assert np.array_equal(f[s].view(np.uint32), want_f.view(np.uint32))


def test_cluster_layout(small):
    h = small.host
    bo = h["block_off"].astype(np.int64)
    ids = h["ids"]
    seen = np.zeros(small.n, np.int64)
    for c in range(small.k):
        s = ids[bo[c] * 32:bo[c + 1] * 32]
        real = s[s != PAD]
        assert np.all(s[:len(real)] != PAD) and np.all(s[len(real):] == PAD), "padding only at the end"
        assert len(s) - len(real) < 32
        assert np.all(np.diff(real.astype(np.int64)) > 0), "members in row order"
        seen[real] += 1
    assert np.all(seen == 2), "SOAR: every row in exactly two clusters"


def test_primary_and_soar_assignment_agree_with_reference(small):
    h = small.host
    C = h["centroids"].reshape(-1, 256)
    rows = np.arange(0, small.n, 97)
    X = np.stack([_prefix(h, r) for r in rows])
    top = O.assign_top(X, C, 16)
    bo = h["block_off"].astype(np.int64)
    member = {}
    for c in range(small.k):
        for r in h["ids"][bo[c] * 32:bo[c + 1] * 32]:
            if r != PAD:
                member.setdefault(int(r), set()).add(c)
    agree = 0
    for x, r, t in zip(X, rows, top):
        want = {int(t[0]), int(O.soar_secondary(x, C, t, 1.0))}
        agree += member[int(r)] == want
    assert agree >= 0.98 * len(rows)


def test_tier2_matches_reference(small):
    h = small.host
    for r in range(0, small.n, 501):
        code, norm = O.tier2_encode(h["tier2_sigma"], _prefix(h, r))
        assert np.array_equal(code, h["tier2_codes"].reshape(-1, 64)[r])
        assert np.float32(norm).view(np.uint32) == h["tier2_norm"][r].view(np.uint32)
