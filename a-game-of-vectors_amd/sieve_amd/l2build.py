"""Index writer for the L2 path over 100-d int8 vectors (the SPACEV-1B shape, src/bigann.rs).

Follows bigann's encoding (src/bigann.rs:331-374, :386-460): each row is assigned to its
nearest leaf (max 2 <x, c> - |c|^2), padded to 128 dims and rotated with Rotation::with_dim(11,
128); the leaf is padded and rotated the same way; the slot keeps the RaBitQ code of the
residual, [mul, |x|^2] and the row id; slots are sorted by (leaf, id) and packed 32 per
512-byte block. The leaves come from a flat L2 k-means here (bigann trains them per coarse
cell, src/bigann.rs:175-226; routing searches all leaves either way).
"""
from dataclasses import dataclass

import numpy as np
import torch

from .build import BLOCK, PAD, f32_div, pack_blocks, rabitq_encode, rotate, rotation_signs
from .synth import sub_seed

D, PD = 100, 128


@dataclass
class L2Arrays:
    leaves: torch.Tensor      # k x 100 f32
    block_off: torch.Tensor   # k + 1 int64
    blocks: torch.Tensor      # nblocks x 512 u8
    ids: torch.Tensor         # nblocks*32 int64 (PAD)
    factors: torch.Tensor     # nblocks*32 x 2 f32: [mul, |x|^2]
    rotation: np.ndarray      # 384 f32
    base: torch.Tensor        # n x 100 int8

    def numpy(self):
        return dict(leaves=self.leaves.cpu().numpy(), block_off=self.block_off.cpu().numpy().astype(np.uint64),
                    blocks=self.blocks.cpu().numpy(), factors=self.factors.cpu().numpy(),
                    ids=self.ids.cpu().numpy().astype(np.uint32), rotation=self.rotation,
                    base=self.base.cpu().numpy())


def int8_mixture(n, components, seed, device="cpu", spread=12.0):
    """n x 100 int8 rows from a seeded mixture, clipped to [-127, 127]."""
    g = torch.Generator(device=device)
    g.manual_seed(sub_seed(seed, 11))
    means = torch.randn(components, D, generator=g, device=device) * 40.0
    comp = torch.randint(0, components, (n,), generator=g, device=device)
    x = means[comp] + spread * torch.randn(n, D, generator=g, device=device)
    return torch.clamp(torch.round(x), -127, 127).to(torch.int8)


def kmeans_l2(x, k, iters, seed):
    """Lloyd's k-means in L2 (src/bigann.rs:139-161), float rows."""
    n = x.shape[0]
    g = torch.Generator(device=x.device)
    g.manual_seed(seed)
    cents = x[torch.randint(0, n, (k,), generator=g, device=x.device)].clone()
    for _ in range(iters):
        a = (2.0 * x @ cents.T - (cents * cents).sum(1)).argmax(1)
        sums = torch.zeros_like(cents).index_add_(0, a, x)
        cnt = torch.bincount(a, minlength=k).to(x.dtype)
        empty = cnt == 0
        cents = torch.where(empty[:, None], cents, sums / cnt.clamp(min=1)[:, None])
    return cents


def seq_norm2(x):
    """sum x_i^2 left to right from -0.0 (src/bigann.rs:350)."""
    xt = x.t().contiguous()
    acc = torch.full((x.shape[0],), -0.0, dtype=torch.float32, device=x.device)
    for i in range(xt.shape[0]):
        acc = acc + xt[i] * xt[i]
    return acc


def pad(x):
    return torch.nn.functional.pad(x, (0, PD - D))


def build_l2(base, k, iters=8, seed=7, chunk=1 << 18):
    dev = base.device
    n = base.shape[0]
    xf = base.to(torch.float32)
    sample = xf[torch.randperm(n, device=dev)[:min(n, 64 * k)]]
    leaves = kmeans_l2(sample, k, iters, seed).contiguous()
    signs = rotation_signs(11, PD)
    leaves_rot = rotate(pad(leaves), signs)
    lnorm = (leaves * leaves).sum(1)
    cl, rw, codes, muls, nxs = [], [], [], [], []
    for s in range(0, n, chunk):
        x = xf[s:s + chunk]
        leaf = (2.0 * x @ leaves.T - lnorm).argmax(1)
        code, mul, _ = rabitq_encode(rotate(pad(x), signs), leaves_rot[leaf])
        cl.append(leaf)
        rw.append(torch.arange(s, s + x.shape[0], device=dev))
        codes.append(code)
        muls.append(mul)
        nxs.append(seq_norm2(x))
    cl, rw = torch.cat(cl), torch.cat(rw)
    order = torch.argsort(cl * n + rw)
    cl, rw = cl[order], rw[order]
    code, mul, nx = torch.cat(codes)[order], torch.cat(muls)[order], torch.cat(nxs)[order]
    sizes = torch.bincount(cl, minlength=k)
    nbl = (sizes + BLOCK - 1) // BLOCK
    block_off = torch.zeros(k + 1, dtype=torch.int64, device=dev)
    block_off[1:] = torch.cumsum(nbl, 0)
    first = torch.zeros(k + 1, dtype=torch.int64, device=dev)
    first[1:] = torch.cumsum(sizes, 0)
    pos = block_off[cl] * BLOCK + torch.arange(cl.shape[0], device=dev) - first[cl]
    S = int(block_off[-1]) * BLOCK
    ids = torch.full((S,), PAD, dtype=torch.int64, device=dev)
    ids[pos] = rw
    slot_codes = torch.zeros(S, PD // 8, dtype=torch.uint8, device=dev)
    slot_codes[pos] = code
    factors = torch.zeros(S, 2, dtype=torch.float32, device=dev)
    factors[pos, 0] = mul
    factors[pos, 1] = nx
    return L2Arrays(leaves=leaves, block_off=block_off, blocks=pack_blocks(slot_codes, PD // 4), ids=ids,
                    factors=factors, rotation=signs, base=base)


__all__ = ["L2Arrays", "int8_mixture", "build_l2", "f32_div"]
