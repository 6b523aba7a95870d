"""The large-table L2 builder (two-level leaves, chunked assignment, encode and pack) on the CPU:
a valid index in the reference's record layout (src/bigann.rs:386-460), the reference's training
recipe and assignment rule (the oracle's train and Model::assign, up to the CPU fixture's f32
GEMM rounding), codes equal to the encoder's own output, and the oracle's recall against exact
L2 search. The GPU build's bit-exactness is tests/test_gpu_build.py's."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from sieve_amd.evaluate import recall10_ties
from sieve_amd.l2build import (L2Mixture, build_l2_large, exact_l2_topk, pad, rabitq_encode, rotate,
                               rotation_signs)


@pytest.fixture(scope="module")
def built():
    mix = L2Mixture(300, seed=9, device="cpu", topics=8)
    n = 120_000
    base = torch.empty(n, 100, dtype=torch.int8)
    mix.fill(base)
    arr = build_l2_large(base, coarse=8, fine=16, iters=4, sample=40_000, chunk=1 << 15)
    return mix, base, arr


def test_chunked_rows_are_reproducible():
    mix = L2Mixture(50, seed=4, device="cpu", topics=4)
    mix.CHUNK = 1000
    whole = torch.empty(3500, 100, dtype=torch.int8)
    mix.fill(whole)
    part = torch.empty(1700, 100, dtype=torch.int8)
    mix.fill(part, start=1234)
    assert torch.equal(part, whole[1234:2934])


def test_slots_hold_every_row_once_in_leaf_order(built):
    _, base, arr = built
    ids = arr.ids.numpy().view(np.uint32)
    bo = arr.block_off.numpy()
    assert bo[-1] * 32 == ids.shape[0] and arr.blocks.shape == (bo[-1], 512)
    valid = ids != 0xFFFFFFFF
    assert np.array_equal(np.sort(ids[valid]), np.arange(base.shape[0]))
    for c in range(len(bo) - 1):
        seg = ids[bo[c] * 32:bo[c + 1] * 32]
        real = seg[seg != 0xFFFFFFFF]
        assert np.all(np.diff(real.astype(np.int64)) > 0)        # rows ascending within a leaf
        assert np.all(seg[len(real):] == 0xFFFFFFFF)               # padding only at the end


def test_training_and_assignment_follow_the_reference(built):
    _, base, arr = built
    coarse, leaves = O.l2_train(base.numpy(), 8, 16, 40_000, 4)
    # f32 GEMM against the reference's sequential dots: the centroids drift by rounding only
    assert np.allclose(arr.coarse.numpy(), coarse, rtol=0, atol=1e-3)
    assert np.allclose(arr.leaves.numpy(), leaves, rtol=0, atol=1e-3)
    ids = arr.ids.numpy().view(np.uint32)
    bo = arr.block_off.numpy()
    leaf_of = np.empty(base.shape[0], np.int64)
    for c in range(len(bo) - 1):
        seg = ids[bo[c] * 32:bo[c + 1] * 32]
        leaf_of[seg[seg != 0xFFFFFFFF]] = c
    want = O.l2_assign(base[:3000].numpy().astype(np.float32), arr.coarse.numpy(), arr.leaves.numpy(), 16)
    agree = np.mean(want == leaf_of[:3000])
    assert agree > 0.995, agree


def test_codes_and_factors_are_the_encoder_output(built):
    _, base, arr = built
    ids = arr.ids.numpy().view(np.uint32)
    bo = arr.block_off.numpy()
    c = int(np.argmax(np.diff(bo)))
    slots = np.arange(bo[c] * 32, bo[c + 1] * 32)
    rows = ids[slots]
    keep = rows != 0xFFFFFFFF
    slots, rows = slots[keep], rows[keep]
    signs = rotation_signs(11, 128)
    x = base[torch.as_tensor(rows.astype(np.int64))].to(torch.float32)
    cr = rotate(pad(arr.leaves[c:c + 1]), signs).repeat(x.shape[0], 1)
    code, mul, _ = rabitq_encode(rotate(pad(x), signs), cr)
    # unpack the slots' codes from the reference's nibble-transposed blocks (src/lut.rs:73-83)
    blk = arr.blocks.numpy()[slots // 32]
    k = slots % 32
    v = np.take_along_axis(blk, np.arange(32)[None, :] * 16 + (k % 16)[:, None], axis=1)  # group g of slot k
    nib = np.where(k[:, None] < 16, v & 0x0F, v >> 4)
    packed = (nib[:, 0::2] | (nib[:, 1::2] << 4)).astype(np.uint8)
    assert np.array_equal(packed, code.numpy())
    f = arr.factors.numpy()[slots]
    assert np.array_equal(f[:, 0].view(np.uint32), mul.numpy().view(np.uint32))
    x64 = x.numpy().astype(np.float64)
    assert np.allclose(f[:, 1], (x64 * x64).sum(1))


def test_oracle_recall_on_the_built_index(built):
    mix, base, arr = built
    q = mix.queries(40)
    gi, gs = exact_l2_topk(base, q, k=32, chunk=1 << 15)
    x = base.numpy().astype(np.int64)
    qq = q.numpy().astype(np.int64)
    for i in range(3):
        top = np.sort(2 * qq[i] @ x.T - (x * x).sum(1))[::-1][:32]
        assert np.array_equal(top.astype(np.float32), gs[i].numpy())
    oix = O.L2Index(arr.leaves.numpy(), arr.block_off.numpy().astype(np.uint64), arr.blocks.numpy(),
                    arr.factors.numpy(), arr.ids.numpy().view(np.uint32), arr.rotation, base.numpy())
    rec = [recall10_ties(oix.search(q.numpy()[i], 16, 1000, 10)[4], gi[i].numpy(), gs[i].numpy())
           for i in range(40)]
    assert np.mean(rec) > 0.9, np.mean(rec)


def test_l2_index_files_have_the_reference_layout(tmp_path):
    """write_l2_index: the postings records of src/bigann.rs:386-460 (4 KiB-aligned, code blocks,
    [mul, |x|^2], ids, refinement section), the Model files and the base chunks, read back by an
    independent parse."""
    import os
    import numpy as np
    from sieve_amd.l2build import build_l2, int8_mixture, write_l2_index
    base = int8_mixture(3000, 12, seed=3)
    arr = build_l2(base, 12, iters=4)
    h = arr.numpy()
    S = h["ids"].shape[0]
    rb = 16 + 4
    ext = np.random.default_rng(1).integers(0, 256, (S, rb), dtype=np.uint8)
    write_l2_index(arr, str(tmp_path / "ix"), str(tmp_path / "base"), refine_bits=2, refine=ext, chunk_rows=1000)
    rd = lambda name, dt: np.fromfile(str(tmp_path / "ix" / name), dt)
    assert np.array_equal(rd("centroids.f32", np.float32), h["leaves"].reshape(-1))
    assert np.array_equal(rd("coarse.f32", np.float32), h["leaves"].reshape(-1)) and rd("k2.u64", np.uint64)[0] == 1
    assert rd("bits.u64", np.uint64)[0] == 2 and rd("rotation.f32", np.float32).shape == (384,)
    bo = rd("block_off.u64", np.uint64).astype(np.int64)
    off = rd("postings.off", np.uint64).astype(np.int64)
    post = np.fromfile(str(tmp_path / "ix" / "postings.bin"), np.uint8)
    assert np.array_equal(bo, h["block_off"].astype(np.int64)) and off[-1] == post.shape[0]
    rec = 512 + 32 * 8 + 32 * 4 + 32 * rb
    for c in range(12):
        nb = int(bo[c + 1] - bo[c])
        assert off[c] % 4096 == 0 and off[c + 1] - off[c] == -(-nb * rec // 4096) * 4096
        r = post[off[c]:off[c] + nb * rec]
        b0 = int(bo[c])
        assert np.array_equal(r[:nb * 512], h["blocks"].reshape(-1, 512)[b0:b0 + nb].reshape(-1))
        assert np.array_equal(r[nb * 512:nb * 768].view(np.float32), h["factors"].reshape(-1)[b0 * 64:(b0 + nb) * 64])
        assert np.array_equal(r[nb * 768:nb * 896].view(np.uint32), h["ids"][b0 * 32:(b0 + nb) * 32].astype(np.uint32))
        assert np.array_equal(r[nb * 896:].reshape(-1, rb), ext[b0 * 32:(b0 + nb) * 32])
    rows = [np.fromfile(str(tmp_path / "base" / f"base.{i:02d}"), np.int8) for i in range(3)]
    assert all(os.path.exists(str(tmp_path / "base" / f"base.{i:02d}.done")) for i in range(3))
    assert np.array_equal(np.concatenate(rows), h["base"].reshape(-1))
