"""The multi-bit refinement of the 1B path on the GPU against the oracle (src/bigann.rs:38-113,
683-729): the device refine_encode (the builder's refinement section) is bit-identical with the
oracle's, and refined searches return the oracle's probes, candidates (the best cand by refined
estimate among the slots the 1-bit shortlist admitted, in probe order) and top-k, for 2- and 3-bit
codes, shortlists from 1 to longer than every probed slot, cand 1 to 400, batches of one and many.
Candidates are compared as lists with bit-identical refined estimates up to ties at the cut."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def refined_l2():
    import torch
    from sieve_amd.l2build import build_l2, int8_mixture, refine_section
    print("[gpu] refinement: building a 30k-row L2 index with 64 leaves", flush=True)
    base = int8_mixture(30000, 80, seed=5)
    arr = build_l2(base, 64, iters=6)
    h = arr.numpy()
    ext = {b: refine_section(arr, b).numpy() for b in (2, 3)}
    q = int8_mixture(40, 80, seed=5)
    g = torch.Generator().manual_seed(7)
    q = torch.clamp(q.to(torch.float32) + torch.randint(-6, 7, q.shape, generator=g), -127, 127).to(torch.int8)
    return arr, h, ext, q.numpy()


def _ties_equal(g_ids, g_sc, w_ids, w_sc):
    g_sc, w_sc = np.asarray(g_sc, np.float32), np.asarray(w_sc, np.float32)
    if len(g_ids) == 0 or len(w_ids) == 0:
        return len(g_ids) == len(w_ids)
    b = max(g_sc.min(), w_sc.min())
    g = {int(i): s for i, s in zip(g_ids, g_sc) if s > b}
    w = {int(i): s for i, s in zip(w_ids, w_sc) if s > b}
    return g.keys() == w.keys() and all(np.float32(g[i]).view(np.uint32) == np.float32(w[i]).view(np.uint32)
                                        for i in g)


@pytest.mark.parametrize("bits", [2, 3])
def test_device_refine_encode_equals_the_oracle(refined_l2, bits):
    import torch
    from sieve_amd.build import rotate
    from sieve_amd.l2build import pad
    arr, h, ext, _ = refined_l2
    ids = h["ids"].astype(np.int64)
    nb = np.diff(h["block_off"].astype(np.int64))
    leaf_of_slot = np.repeat(np.arange(len(nb)), nb * 32)
    leaves_rot = rotate(pad(torch.as_tensor(h["leaves"]).reshape(-1, 100)), h["rotation"]).numpy()
    rb = O.refine_bytes(bits)
    slots = np.nonzero(ids != 0xFFFFFFFF)[0]
    for s in slots[:: max(1, len(slots) // 1500)]:
        x = torch.as_tensor(h["base"].reshape(-1, 100)[ids[s]].astype(np.float32))[None, :]
        r = (rotate(pad(x), h["rotation"]).numpy()[0] - leaves_rot[leaf_of_slot[s]]).astype(np.float32)
        planes, f = O.refine_encode(r, bits)
        assert np.array_equal(ext[bits][s, :rb - 4], planes[:rb - 4]), s
        assert ext[bits][s, rb - 4:].tobytes() == np.float32(f).tobytes(), s
    pad_slots = np.nonzero(ids == 0xFFFFFFFF)[0]
    assert not ext[bits][pad_slots].any()


@pytest.mark.parametrize("bits", [2, 3])
@pytest.mark.parametrize("nprobe,cand,refine,margin", [(8, 50, 300, 0.0), (16, 100, 1000, 0.0), (6, 30, 1, 0.0),
                                                        (12, 400, 0, 0.0), (10, 60, 1 << 16, 0.0),
                                                        (32, 100, 500, 0.4), (4, 1, 5, 0.0)])
def test_refined_search_matches_oracle(refined_l2, bits, nprobe, cand, refine, margin):
    from sieve_amd.index import L2Index
    arr, h, ext, qs = refined_l2
    hh = dict(h, refine_bits=bits, refine=ext[bits])
    ix = L2Index(hh)
    oix = O.L2Index(h["leaves"], h["block_off"], h["blocks"], h["factors"], h["ids"], h["rotation"], h["base"])
    oix.set_refine(bits, ext[bits])
    s, r, n, probes, pn, cr, ce, cn = ix.search(qs, nprobe, cand, 10, margin=margin, details=True, refine=refine)
    for i, q in enumerate(qs):
        wp, wc, we, ws, wr = oix.search(q, nprobe, cand, 10, margin=margin, refine=refine)
        assert pn[i] == len(wp) and set(probes[i, :pn[i]].tolist()) == set(wp.tolist())
        assert cn[i] == len(wc), (i, cn[i], len(wc))
        assert _ties_equal(cr[i, :cn[i]], ce[i, :cn[i]], wc, we), i
        assert _ties_equal(r[i, :n[i]], s[i, :n[i]], wr, ws), i
    # single queries give the batch's results
    for i in (0, 17):
        s1, r1, n1 = ix.search(qs[i:i + 1], nprobe, cand, 10, margin=margin, refine=refine)
        assert np.array_equal(r1[0, :n1[0]], r[i, :n[i]])
        assert np.array_equal(s1[0, :n1[0]].view(np.uint32), s[i, :n[i]].view(np.uint32))


def test_refine_needs_a_multibit_index(refined_l2):
    """On an index without a refinement section refine is ignored (src/bigann.rs:688)."""
    from sieve_amd.index import L2Index
    _, h, _, qs = refined_l2
    ix = L2Index(h)
    a = ix.search(qs, 8, 100, 10)
    b = ix.search(qs, 8, 100, 10, refine=500)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("bits", [1, 2])
def test_l2_index_open_from_the_reference_files(refined_l2, bits, tmp_path):
    """sieve_l2_index_open over the files write_l2_index writes (postings records, Model files,
    base chunks): searches equal the index created from the same arrays, bit for bit."""
    from sieve_amd.index import L2Index
    from sieve_amd.l2build import write_l2_index
    arr, h, ext, qs = refined_l2
    write_l2_index(arr, str(tmp_path / "ix"), str(tmp_path / "base"), refine_bits=bits,
                   refine=ext[bits] if bits > 1 else None)
    fix = L2Index.open(str(tmp_path / "ix"), str(tmp_path / "base"))
    aix = L2Index(dict(h, refine_bits=bits, refine=ext[bits]) if bits > 1 else h)
    assert fix.k == aix.k and fix.bits == bits
    for nprobe, cand, refine in ((8, 300, 0), (16, 100, 500)):
        a = aix.search(qs, nprobe, cand, 10, details=True, refine=refine)
        b = fix.search(qs, nprobe, cand, 10, details=True, refine=refine)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))


def test_l2_open_rejects_a_mismatched_postings_offset_file(refined_l2, tmp_path):
    from sieve_amd.abi import SieveError
    from sieve_amd.index import L2Index
    from sieve_amd.l2build import write_l2_index
    arr, _, _, _ = refined_l2
    write_l2_index(arr, str(tmp_path / "ix"), str(tmp_path / "base"))
    off = np.fromfile(str(tmp_path / "ix" / "postings.off"), np.uint64)
    off[:-1].tofile(str(tmp_path / "ix" / "postings.off"))
    with pytest.raises(SieveError, match="postings.off"):
        L2Index.open(str(tmp_path / "ix"), str(tmp_path / "base"))


@pytest.mark.parametrize("bits,margin", [(1, 0.0), (2, 0.3)])
def test_hot_leaves_match_the_oracle(refined_l2, bits, margin):
    from sieve_amd.index import L2Index
    arr, h, ext, qs = refined_l2
    ix = L2Index(dict(h, refine_bits=bits, refine=ext[bits]) if bits > 1 else h)
    oix = O.L2Index(h["leaves"], h["block_off"], h["blocks"], h["factors"], h["ids"], h["rotation"], h["base"])
    rec = 512 + 32 * 8 + 32 * 4 + 32 * O.refine_bytes(bits)
    rb = (np.diff(h["block_off"].astype(np.int64)) * rec + 4095) // 4096 * 4096
    for budget in (0, int(rb.max()), int(rb.sum() // 3), int(rb.sum()) * 2):
        got = ix.hot_leaves(qs, 8, margin, budget)
        want = oix.hot_leaves(qs, 8, margin, rb.astype(np.uint64), budget)
        assert got.tolist() == want.tolist(), budget
