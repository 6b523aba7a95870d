"""Pins the CPU oracle's small kernels: the reference's own unit tests (ported with their
fixtures and seeds), published known answers, and closed forms computed independently in
float64 / integer numpy. Runs on the CPU."""
import numpy as np
import pytest
from scipy.linalg import hadamard

from oracle import oracle as O

PDIM, DIM, BLOCK = 256, 768, 32


def nibbles_np(code, G):
    c = np.asarray(code, np.uint8)
    return np.stack([(c >> 0) & 15, (c >> 4) & 15], axis=-1).reshape(c.shape[:-1] + (G,))


def bits_np(code, d):
    c = np.asarray(code, np.uint8)
    return np.unpackbits(c, bitorder="little")[:d].astype(bool)


# ----------------------------------------------------------------------- Rng / gemm ----
def test_splitmix64_published_vector():
    # Published splitmix64 outputs for seed 1234567 (Vigna's reference generator).
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert O.rng_stream(1234567, 5).tolist() == want


def test_gemm_small_product_known_answer():
    # blas.rs:58-65
    a = np.array([[1, 2, 3], [4, 5, 6]], np.float32)
    b = np.array([[1, 0, 1], [0, 1, 0]], np.float32)
    assert O.gemm_abt(a, b).ravel().tolist() == [4.0, 2.0, 10.0, 5.0]


def test_gemm_is_sequential_f32_and_within_f64_bound(rng):
    a = rng.standard_normal((5, 300)).astype(np.float32)
    b = rng.standard_normal((7, 300)).astype(np.float32)
    c = O.gemm_abt(a, b, threads=3)
    seq = np.empty((5, 7), np.float32)
    for i in range(5):
        for j in range(7):
            acc = np.float32(-0.0)
            for t in range(300):
                acc = np.float32(acc + np.float32(a[i, t] * b[j, t]))
            seq[i, j] = acc
    assert np.array_equal(c.view(np.uint32), seq.view(np.uint32))
    c64 = a.astype(np.float64) @ b.astype(np.float64).T
    bound = 300 * 2.0 ** -24 * (np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64).T)
    assert np.all(np.abs(c - c64) <= bound)


# ----------------------------------------------------------------------- LUT -----------
def test_lut_kernel_matches_reference(rng):
    # lut.rs:179-196 (kernel_matches_reference), plus an independent bit-level sum.
    for _ in range(20):
        q = rng.uniform(-1, 1, PDIM).astype(np.float32)
        codes = rng.integers(0, 256, (BLOCK, 32), dtype=np.uint8)
        tables, scale, bias = O.lut_new(q)
        block = O.pack_block(codes, 64)
        got = O.scan(tables, block)
        nib = nibbles_np(codes, 64)
        brute = tables[np.arange(64)[None, :], nib].astype(np.int64).sum(1)
        for k in range(BLOCK):
            assert got[k] == O.score_code(tables, codes[k])
        assert np.array_equal(got.astype(np.int64), brute)


def test_lut_quantized_score_tracks_float(rng):
    # lut.rs:198-209 (quantized_score_tracks_float); float_score against float64.
    for _ in range(200):
        q = rng.uniform(-0.2, 0.2, PDIM).astype(np.float32)
        code = rng.integers(0, 256, 32, dtype=np.uint8)
        tables, scale, bias = O.lut_new(q)
        exact = O.float_score(q, code)
        assert abs(exact - q.astype(np.float64)[bits_np(code, PDIM)].sum()) < 1e-5
        approx = np.float32(O.score_code(tables, code)) / np.float32(scale) + np.float32(bias)
        bound = 64 * 0.5 / scale + 1e-4
        assert abs(exact - approx) <= bound


def test_lut_table_invariants(rng):
    # Each group's smallest entry quantises to 0 and the widest group spans 0..255; the
    # dequantised table reproduces the float table within half a step.
    for _ in range(50):
        q = rng.standard_normal(PDIM).astype(np.float32)
        tables, scale, bias = O.lut_new(q)
        assert np.all(tables.min(1) == 0)
        assert tables.max() == 255
        f = np.zeros((64, 16))
        for c in range(16):
            for j in range(4):
                if (c >> j) & 1:
                    f[:, c] += q.astype(np.float64)[np.arange(64) * 4 + j]
        deq = tables / scale + f.min(1, keepdims=True)
        assert np.all(np.abs(deq - f) <= 0.5 / scale + 1e-5)
        assert abs(bias - f.min(1).sum()) < 1e-4


def test_lut_narrow_codes_fixture():
    # lut.rs:212-228: G = 32, q_i = cos(0.91 i), code byte i of doc k = (37k + 11i) mod 256.
    q = np.cos(np.arange(128, dtype=np.float32) * np.float32(0.91)).astype(np.float32)
    tables, _, _ = O.lut_new(q)
    codes = np.array([[(k * 37 + i * 11) % 256 for i in range(16)] for k in range(BLOCK)], np.uint8)
    block = O.pack_block(codes, 32)
    got = O.scan(tables, block)
    for k in range(BLOCK):
        assert got[k] == O.score_code(tables, codes[k])
        assert np.array_equal(O.unpack_slot(block, k), codes[k])


def test_lut_partial_block_pads_with_zero_codes():
    # lut.rs:230-241
    q = np.sin(np.arange(PDIM, dtype=np.float32) * np.float32(0.37)).astype(np.float32)
    tables, _, _ = O.lut_new(q)
    c = np.full((1, 32), 0xA5, np.uint8)
    got = O.scan(tables, O.pack_block(c, 64))
    assert got[0] == O.score_code(tables, c[0])
    assert got[1] == O.score_code(tables, np.zeros(32, np.uint8))


def test_pack_block_layout_bitwise(rng):
    # Byte g*16+k holds doc k's nibble g (low) and doc k+16's (high) (lut.rs:7-9).
    codes = rng.integers(0, 256, (BLOCK, 32), dtype=np.uint8)
    block = O.pack_block(codes, 64)
    nib = nibbles_np(codes, 64)
    for g in range(64):
        for k in range(16):
            assert block[g * 16 + k] == nib[k, g] | (nib[k + 16, g] << 4)


def test_hamming_counts_bits(rng):
    # lut.rs:243-248, plus numpy popcount.
    assert O.hamming(np.full(32, 0xFF, np.uint8), np.full(32, 0x0F, np.uint8)) == 4 * 32
    a = rng.integers(0, 256, 32, dtype=np.uint8)
    b = rng.integers(0, 256, 32, dtype=np.uint8)
    assert O.hamming(a, b) == int(np.unpackbits(a ^ b).sum())


# ----------------------------------------------------------------------- rotation ------
def test_fwht_is_the_sylvester_hadamard_product(rng):
    for d in (4, 32, 128, 256):
        v = rng.standard_normal(d).astype(np.float32)
        want = hadamard(d).astype(np.float64) @ v.astype(np.float64)
        assert np.allclose(O.fwht(v), want, rtol=1e-5, atol=1e-4)


def test_rotation_is_three_sign_flip_hadamard_rounds(rng):
    for d, seed in ((256, 11), (128, 11), (256, 3)):
        signs = O.rotation_signs(seed, d)
        assert set(np.unique(signs)) <= {-1.0, 1.0}
        bits = (O.rng_stream(seed, 3 * d) & np.uint64(1)).astype(np.float32) * 2 - 1
        assert np.array_equal(signs, bits)
        H = hadamard(d).astype(np.float64)
        P = np.eye(d)
        for r in range(3):
            P = H @ np.diag(signs[r * d:(r + 1) * d].astype(np.float64)) @ P
        P *= float(d) ** -1.5
        v = rng.standard_normal(d).astype(np.float32)
        assert np.allclose(O.rotation_apply(signs, v), P @ v.astype(np.float64), rtol=1e-5, atol=1e-6)
        assert np.allclose(P @ P.T, np.eye(d), atol=1e-9)


def _ref_unit(state):
    """The reference test helper `unit` (rabitq.rs:138-145) with the oracle's Rng."""
    v = np.empty(PDIM, np.float32)
    for i in range(PDIM):
        s = np.float32(-0.0)
        for _ in range(4):
            x = O.lib().or_rng_next(O._p(state))
            u = np.float32(np.float32(x >> 11) / np.float32(2.0 ** 53)) - np.float32(0.5)
            s = np.float32(s + np.float32(u))
        v[i] = s
    return O.normalize(v)[0]


def test_rotation_is_orthogonal_reference_seeds():
    # rabitq.rs:147-157: Rng(1) units, Rotation(3).
    st = np.array([1], np.uint64)
    a, b = _ref_unit(st), _ref_unit(st)
    signs = O.rotation_signs(3)
    ra, rb = O.rotation_apply(signs, a), O.rotation_apply(signs, b)
    assert abs(O.dot_f32(ra, ra) - 1.0) < 1e-4
    assert abs(O.dot_f32(ra, rb) - O.dot_f32(a, b)) < 1e-4


def test_rabitq_encode_closed_form(rng):
    for d in (256, 128):
        for _ in range(50):
            xr = rng.standard_normal(d).astype(np.float32) * 0.1
            cr = rng.standard_normal(d).astype(np.float32) * 0.1
            code, (mul, err) = O.rabitq_encode(xr, cr)
            r = xr.astype(np.float64) - cr.astype(np.float64)
            assert np.array_equal(bits_np(code, d), (xr - cr) > 0)
            nr, l1 = np.sqrt((r * r).sum()), np.abs(r).sum()
            od = l1 / (nr * np.sqrt(d))
            assert abs(mul - nr * nr / l1) <= 1e-5 * abs(nr * nr / l1)
            e = nr * 1.9 * np.sqrt(max((1 - od * od) / (od * od * (d - 1)), 0.0))
            assert abs(err - e) <= 1e-4 * e + 1e-7
    code, me = O.rabitq_encode(np.ones(256, np.float32), np.ones(256, np.float32))
    assert me.tolist() == [0.0, 0.0] and not code.any()


def test_rabitq_estimate_is_within_bound_reference_seeds():
    # rabitq.rs:159-187: Rng(2), Rotation(5), 2000 trials; >95% inside 1.1*err, RMSE < 0.05.
    st = np.array([2], np.uint64)
    signs = O.rotation_signs(5)
    inside, sq = 0, 0.0
    trials = 2000
    for _ in range(trials):
        x = _ref_unit(st)
        u = _ref_unit(st)
        c = (x + np.float32(0.8) * u).astype(np.float32)
        c = O.normalize(c)[0]
        q = _ref_unit(st)
        xr, cr = O.rotation_apply(signs, x), O.rotation_apply(signs, c)
        code, (mul, err) = O.rabitq_encode(xr, cr)
        tables, scale, bias, a, b = O.query_new(signs, q)
        est = O.estimate(O.dot_f32(c, q), mul, O.score_code(tables, code), a, b)
        truth = O.dot_f32(x, q)
        sq += (est - truth) ** 2
        inside += abs(est - truth) <= err * 1.1
    assert inside / trials > 0.95
    assert np.sqrt(sq / trials) < 0.05


def test_query_affine_map_recovers_signed_sum(rng):
    # a*lut + b == sum_i sign_i (Pq)_i up to the LUT's rounding (rabitq.rs:14-16, 119-125).
    signs = O.rotation_signs(11)
    for _ in range(50):
        q = O.normalize(rng.standard_normal(PDIM).astype(np.float32))[0]
        code = rng.integers(0, 256, 32, dtype=np.uint8)
        tables, scale, bias, a, b = O.query_new(signs, q)
        pq = O.rotation_apply(signs, q).astype(np.float64)
        sgn = np.where(bits_np(code, PDIM), 1.0, -1.0)
        got = a * O.score_code(tables, code) + b
        assert abs(got - (sgn * pq).sum()) <= 2 * 64 * 0.5 / scale + 1e-4
        assert abs(a - 2.0 / scale) < 1e-12 * abs(a) + 1e-30


# ----------------------------------------------------------------------- quant --------
def test_sign_code_bits():
    # quant.rs:147-159
    p = np.full(PDIM, -1.0, np.float32)
    p[0], p[9], p[255] = 0.5, 0.1, 2.0
    c = O.sign_code(p)
    assert c[0] == 1 and c[1] == 2 and c[31] == 128
    assert int(np.unpackbits(c).sum()) == 3


def test_int8_roundtrip_and_dot():
    # quant.rs:161-171
    v = np.array([((i * 37 % 101) - 50) / 50 for i in range(DIM)], np.float32)
    v = O.normalize(v)[0]
    scale = O.int8_fit(v[None, :])
    c = O.int8_encode(v, scale)
    w = (v / scale).astype(np.float32)
    assert abs(O.dot_i8(w, c) - 1.0) < 0.01


def test_sdot_matches_integer_dot(rng):
    # quant.rs:173-178, against an int64 dot.
    for _ in range(50):
        a = rng.integers(-128, 128, DIM).astype(np.int8)
        b = rng.integers(-128, 128, DIM).astype(np.int8)
        assert O.dot_i8i8(a, b) == int(a.astype(np.int64) @ b.astype(np.int64))
        assert O.dot_i8i8(a[:PDIM], b[:PDIM]) == int(a[:PDIM].astype(np.int64) @ b[:PDIM].astype(np.int64))


def test_query_i8_closed_form(rng):
    for _ in range(50):
        w = (rng.standard_normal(DIM) * rng.uniform(0.01, 10)).astype(np.float32)
        wq, s = O.query_i8(w)
        m = np.float32(max(np.abs(w).max(), np.float32(1e-12)))
        assert s == np.float32(np.float32(127.0) / m)
        x = (w * np.float32(s)).astype(np.float32)
        t = np.trunc(x)
        want = t + np.where(np.abs(x - t) >= 0.5, np.sign(x), 0)
        assert np.array_equal(wq.astype(np.float64), want)
        assert np.abs(wq).max() == 127


# ----------------------------------------------------------------------- TopK ---------
def test_topk_matches_full_sort(rng):
    # topk.rs:64-82
    for _ in range(100):
        n = int(rng.integers(0, 2000))
        k = int(rng.integers(1, 64))
        scores = rng.integers(0, 500, n).astype(np.uint16)
        got, _ = O.topk(scores, k)
        want = np.sort(scores)[::-1][:k]
        assert np.array_equal(got, want)


def test_topk_f32_order_and_ids(rng):
    s = rng.standard_normal(3000).astype(np.float32)
    got_s, got_i = O.topk(s, 37)
    order = np.argsort(-s, kind="stable")[:37]
    assert np.array_equal(got_i, order) and np.array_equal(got_s, s[order])
    # ties sort by id ascending
    s = np.array([1, 2, 2, 3, 2], np.float32)
    gs, gi = O.topk(s, 5)
    assert gi.tolist() == [3, 1, 2, 4, 0]


# ----------------------------------------------------------------------- tier2 --------
def test_tier2_table_score_matches_reconstruction():
    # tier2.rs:81-96 with its xorshift fixture.
    x = [0x2545F4914F6CDD1D]

    def rnd():
        v = x[0]
        v ^= (v << 13) & 0xFFFFFFFFFFFFFFFF
        v ^= v >> 7
        v ^= (v << 17) & 0xFFFFFFFFFFFFFFFF
        x[0] = v
        return np.float32(np.float32(v >> 11) / np.float32(2.0 ** 53)) - np.float32(0.5)

    rows = np.array([rnd() for _ in range(PDIM * 64)], np.float32).reshape(64, PDIM)
    sigma = O.tier2_fit(rows)
    q = np.array([rnd() for _ in range(PDIM)], np.float32)
    tables = O.tier2_tables(sigma, q)
    levels = np.array([-1.510, -0.4528, 0.4528, 1.510], np.float32)
    for r in rows[:8]:
        code, norm = O.tier2_encode(sigma, r)
        lv = (code[np.arange(PDIM) >> 2] >> ((np.arange(PDIM) & 3) * 2)) & 3
        recon = levels[lv] * sigma
        want = float((q.astype(np.float64) * recon).sum()) / norm
        assert abs(O.tier2_score(tables, code, norm) - want) < 1e-4
        assert abs(norm - np.sqrt((recon.astype(np.float64) ** 2).sum())) < 1e-5


# ----------------------------------------------------------------------- prune / eval -
def test_prune_threshold_is_mth_lower_bound():
    # ivf.rs:856-868
    est = np.array([0.9, 0.7, 0.7, 0.3], np.float32)
    err = np.array([0.1, 0.2, 0.05, 0.2], np.float32)
    tau = O.prune_threshold(est, err, 2, 1.0)
    assert tau == np.float32(0.65)
    assert O.prune_threshold(est, err, 5, 1.0) == -np.inf
    kept = [i for i in range(4) if est[i] + err[i] >= tau]
    assert kept == [0, 1, 2]
    assert O.prune_threshold(est, err, 2, 0.5) > tau


def test_recall_counts_overlap():
    # eval.rs:61-64
    from sieve_amd.evaluate import recall_at
    assert recall_at(2, [1, 2, 3], [2, 9, 1]) == 0.5
