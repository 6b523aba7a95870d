"""The oracle's multi-bit refinement of the 1B path (src/bigann.rs:38-113, 537-602, 683-729)
against the reference's own test and what the mathematics fixes (CPU only).

* refine_encode + Refiner: the reference's test (bigann.rs:797-816, seeded splitmix64 Gaussians):
  the estimate of <r, r> from r's own code is exact to 1e-3, and 3-bit codes err less than half
  as much as 2-bit ones.
* refine_encode against float64: the planes and the stored sign are a level vector u in range
  whose top bit is the sign; its step is the best of the 24 tried by cos(r, u - h); f is
  |r|^2 / <r, u - h>.
* Refiner: its table estimate equals f (<q, u> - h sum q) computed directly in float64.
* candidates with refinement: with a shortlist at least as long as the probed slots nothing is
  ever cut, so the candidates are the best `cand` of all probed slots by the refined estimate (a
  full sort); refine 0 is the 1-bit path; results are the top-k of exact L2 over the candidates.
* hot_leaves: probe counts from a float64 route, heat per record byte, budget prefix.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sieve_amd.build import Rng

PD, CODE = 128, 16


def gauss(rng):
    # bigann.rs:782-784: each value a sum of 12 uniform thousandths minus 6, in f32.
    out = np.zeros(PD, np.float32)
    for i in range(PD):
        s = np.float32(-0.0)
        for _ in range(12):
            s = np.float32(s + np.float32(np.float32(rng.below(1000)) / np.float32(1000.0)))
        out[i] = np.float32(s - np.float32(6.0))
    return out


def sign_code(r):
    c = np.zeros(CODE, np.uint8)
    for i in range(PD):
        if r[i] > 0:
            c[i >> 3] |= 1 << (i & 7)
    return c


def levels(sign, planes, bits):
    """u_i = 2^(bits-1) * sign bit + sum_p 2^p * plane_p bit."""
    bit = lambda b, i: (int(b[i >> 3]) >> (i & 7)) & 1
    return np.array([(1 << (bits - 1)) * bit(sign, i) + sum((1 << p) * bit(planes[p * CODE:], i)
                                                          for p in range(bits - 1)) for i in range(PD)])


def test_refined_estimate_is_exact_on_its_own_residual_and_sharpens_with_bits():
    # Port of bigann.rs:797-816 with its seed.
    rng = Rng(9)
    err = {2: 0.0, 3: 0.0}
    for _ in range(200):
        r, q = gauss(rng), gauss(rng)
        truth = np.float32(-0.0)
        for a, b in zip(r, q):
            truth = np.float32(truth + np.float32(a * b))
        for bits in (2, 3):
            planes, f = O.refine_encode(r, bits)
            np_ = (bits - 1) * CODE
            self_est = O.refine_estimate(r, bits, sign_code(r), planes[:np_], f)
            rr = float(np.sum(r.astype(np.float64) ** 2))
            assert abs(self_est - rr) < 1e-3 * rr, (bits, self_est, rr)
            est = O.refine_estimate(q, bits, sign_code(r), planes[:np_], f)
            err[bits] += (est - float(truth)) ** 2
    assert err[3] < 0.5 * err[2], err


@pytest.mark.parametrize("bits", [2, 3])
def test_refine_encode_against_float64(bits):
    rng = np.random.default_rng(bits)
    h = ((1 << bits) - 1) / 2.0
    for trial in range(60):
        r = (rng.standard_normal(PD) * rng.uniform(0.01, 50)).astype(np.float32)
        if trial == 0:
            r[:7] = 0.0  # r_i = 0: the top bit is the sign (clear)
        planes, f = O.refine_encode(r, bits)
        u = levels(sign_code(r), planes, bits)
        assert u.min() >= 0 and u.max() <= (1 << bits) - 1
        top = 1 << (bits - 1)
        assert np.all((u >= top) == (r > 0))
        assert not np.any(planes[(bits - 1) * CODE:])
        r64 = r.astype(np.float64)
        w = u - h
        cos_chosen = r64 @ w / np.linalg.norm(w)
        # every step of the 24-point grid gives a level vector no better than the chosen one
        amax = np.abs(r64).max()
        for st in range(24):
            delta = amax / h * (0.3 + 0.035 * st)
            uu = np.clip(np.floor(r64 / delta + h + 0.5), 0, (1 << bits) - 1).astype(np.int64)
            uu = np.where(r64 > 0, uu | top, uu & (top - 1))
            ww = uu - h
            assert r64 @ ww / np.linalg.norm(ww) <= cos_chosen + 1e-5 * abs(cos_chosen)
        assert f == pytest.approx((r64 @ r64) / (r64 @ w), rel=1e-5)


def test_refine_encode_zero_residual():
    planes, f = O.refine_encode(np.zeros(PD, np.float32), 3)
    assert f == 0.0 and not planes.any()


@pytest.mark.parametrize("bits", [2, 3])
def test_refiner_estimate_is_the_level_inner_product(bits):
    rng = np.random.default_rng(10 + bits)
    h = ((1 << bits) - 1) / 2.0
    for _ in range(50):
        q = rng.standard_normal(PD).astype(np.float32)
        sign = rng.integers(0, 256, CODE, dtype=np.uint8)
        planes = rng.integers(0, 256, (bits - 1) * CODE, dtype=np.uint8)
        f = float(rng.uniform(0.1, 3))
        u = levels(sign, planes, bits)
        want = f * (q.astype(np.float64) @ u - h * q.astype(np.float64).sum())
        got = O.refine_estimate(q, bits, sign, planes, np.float32(f))
        assert got == pytest.approx(want, rel=1e-4, abs=1e-3)


# ------------------------------------------------------------ the refined search ------
@pytest.fixture(scope="module")
def refined():
    """A small L2 index with a refinement section (bits 2 and 3) encoded by the oracle."""
    import torch
    from sieve_amd.build import rotate
    from sieve_amd.l2build import build_l2, int8_mixture, pad
    base = int8_mixture(6000, 24, seed=8)
    arr = build_l2(base, 24, iters=5)
    h = arr.numpy()
    ids = h["ids"].astype(np.int64)
    leaf_of_slot = np.repeat(np.arange(24), np.diff(h["block_off"].astype(np.int64)) * 32)
    signs = torch.as_tensor(h["rotation"])
    xr = rotate(pad(base.to(torch.float32)), signs).numpy()
    cr = rotate(pad(torch.as_tensor(h["leaves"]).reshape(-1, 100)), signs).numpy()
    sections = {}
    for bits in (2, 3):
        nb = O.refine_bytes(bits)
        ext = np.zeros((ids.shape[0], nb), np.uint8)
        for s in np.nonzero(ids != 0xFFFFFFFF)[0]:
            r = (xr[ids[s]] - cr[leaf_of_slot[s]]).astype(np.float32)
            planes, f = O.refine_encode(r, bits)
            ext[s, :nb - 4] = planes[:nb - 4]
            ext[s, nb - 4:] = np.frombuffer(np.float32(f).tobytes(), np.uint8)
        sections[bits] = ext
    q = int8_mixture(16, 24, seed=8)
    q = np.clip(q.numpy().astype(np.int32) + np.random.default_rng(2).integers(-6, 7, q.shape), -127, 127)
    return h, xr, cr, leaf_of_slot, sections, q.astype(np.int8)


def _index(h):
    return O.L2Index(h["leaves"], h["block_off"], h["blocks"], h["factors"], h["ids"], h["rotation"], h["base"])


@pytest.mark.parametrize("bits", [2, 3])
def test_refined_candidates_with_an_unbounded_shortlist_are_a_full_sort(refined, bits):
    import torch
    from sieve_amd.build import rotate
    from sieve_amd.l2build import pad
    h, xr, cr, leaf_of_slot, sections, qs = refined
    ix = _index(h)
    ix.set_refine(bits, sections[bits])
    ids = h["ids"].astype(np.int64)
    nx = h["factors"].reshape(-1, 2)[:, 1]
    ext = sections[bits]
    nb = O.refine_bytes(bits)
    hh = ((1 << bits) - 1) / 2.0
    for q in qs:
        probe, cids, cest, s, r = ix.search(q, 4, 50, 10, refine=1 << 20)
        qr = rotate(pad(torch.as_tensor(q.astype(np.float32))[None, :]), torch.as_tensor(h["rotation"])).numpy()[0]
        qr64 = qr.astype(np.float64)
        leaves = h["leaves"].reshape(-1, 100).astype(np.float64)
        est, sid = [], []
        for c in probe:
            b0, b1 = int(h["block_off"][c]) * 32, int(h["block_off"][c + 1]) * 32
            cent = leaves[c] @ q.astype(np.float64)
            for slot in range(b0, b1):
                if ids[slot] == 0xFFFFFFFF:
                    continue
                sign = sign_code(xr[ids[slot]] - cr[c])
                u = levels(sign, ext[slot, :nb - 4], bits)
                f = np.frombuffer(ext[slot, nb - 4:].tobytes(), np.float32)[0]
                est.append(2 * (cent + f * (qr64 @ u - hh * qr64.sum())) - nx[slot])
                sid.append(ids[slot])
        est, sid = np.array(est), np.array(sid)
        order = np.argsort(-est, kind="stable")
        # the best 50 by the refined estimate, up to float32 resolution at the boundary
        want = set(sid[order[:50]].tolist())
        got = set(cids.tolist())
        cut = est[order[49]]
        near = set(sid[np.abs(est - cut) <= 1e-3 * (1 + abs(cut))].tolist())
        assert (got ^ want) <= near, (got ^ want)
        # final: exact L2 over the candidates in f32 (2 dot as f32 - nx, src/bigann.rs:736-747),
        # best 10; equal scores at the 10th place may come in any order (TopK's boundary)
        sc = np.array([np.float32(2.0 * float(h["base"].reshape(-1, 100)[i].astype(np.int64) @ q.astype(np.int64)))
                       - nx[np.nonzero(ids == i)[0][0]] for i in cids], np.float32)
        o2 = sorted(range(len(cids)), key=lambda j: (-sc[j], cids[j]))[:10]
        pos = {int(c): j for j, c in enumerate(cids)}
        assert [sc[pos[i]] for i in r.tolist()] == [sc[j] for j in o2]
        edge = sc[o2[-1]]
        assert [i for i in r.tolist() if sc[pos[i]] > edge] == [int(cids[j]) for j in o2 if sc[j] > edge]


def test_refine_zero_is_the_one_bit_path(refined):
    h, _, _, _, sections, qs = refined
    plain, ref = _index(h), _index(h)
    ref.set_refine(3, sections[3])
    for q in qs:
        a = plain.search(q, 6, 80, 10)
        b = ref.search(q, 6, 80, 10, refine=0)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_refined_shortlist_contains_the_final_one_bit_top(refined):
    """With a short list the refined candidates are drawn from the slots the 1-bit shortlist
    admitted, which include its final top: raising refine never loses a candidate that a
    full-length shortlist keeps and that ranks above every admitted-only entry."""
    h, _, _, _, sections, qs = refined
    ix = _index(h)
    ix.set_refine(2, sections[2])
    for q in qs:
        full = ix.search(q, 6, 30, 10, refine=1 << 20)
        short = ix.search(q, 6, 30, 10, refine=40)
        assert len(short[1]) == len(full[1]) == 30
        # the refined estimates of candidates the two runs share are identical
        fe = dict(zip(full[1].tolist(), full[2].tolist()))
        for i, e in zip(short[1].tolist(), short[2].tolist()):
            if i in fe:
                assert np.float32(e) == np.float32(fe[i])


def test_hot_leaves_against_float64_routing(refined):
    h, _, _, _, _, qs = refined
    ix = _index(h)
    leaves = h["leaves"].reshape(-1, 100).astype(np.float64)
    rb = (np.diff(h["block_off"].astype(np.int64)) * (896 + 32 * 4) + 4095) // 4096 * 4096
    hits = np.zeros(leaves.shape[0], np.int64)
    for q in qs:
        s = 2 * leaves @ q.astype(np.float64) - (leaves ** 2).sum(1)
        o = np.argsort(-s, kind="stable")
        assert s[o[4]] - s[o[5]] > 1e-2  # no near-tie at the cut in this fixture
        hits[o[:5]] += 1
    heat = hits / np.maximum(rb, 1)
    order = [c for c in sorted(range(len(hits)), key=lambda c: -heat[c]) if hits[c] > 0]
    for budget in (0, int(rb.max()), int(rb.sum() // 2), int(rb.sum()) * 2):
        want, used = [], 0
        for c in order:
            used += int(rb[c])
            if used > budget:
                break
            want.append(c)
        got = ix.hot_leaves(qs, 5, 0.0, rb.astype(np.uint64), budget)
        assert got.tolist() == want, budget
