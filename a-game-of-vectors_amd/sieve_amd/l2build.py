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
    coarse: torch.Tensor = None  # k1 x 100 f32: the coarse cells of a two-level build (Model::coarse)

    def numpy(self):
        return dict(leaves=self.leaves.cpu().numpy(), block_off=self.block_off.cpu().numpy().astype(np.uint64),
                    blocks=self.blocks.cpu().numpy(), factors=self.factors.cpu().numpy(),
                    ids=self.ids.cpu().numpy().astype(np.uint32), rotation=self.rotation,
                    base=self.base.cpu().numpy())


def int8_mixture(n, components, seed, device="cpu", spread=12.0):
    """n x 100 int8 rows from a seeded mixture, clipped to [-127, 127] (the 100-d int8 L2 workload)."""
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


def refine_section(arr, bits, chunk=1 << 22):
    """The per-slot extra section of a multi-bit index (src/bigann.rs:331-345, 430-436): for every
    slot, refine_encode (src/bigann.rs:70-108, on the device through sieve_l2_refine_encode) of
    its rotated residual, the row padded and rotated minus its leaf padded and rotated; stored
    as the low bits - 1 planes (16 bytes each) followed by the scale f as a little-endian f32,
    refine_bytes(bits) bytes per slot in slot order, zero for padding slots."""
    from . import abi
    assert bits in (2, 3)
    dev = arr.base.device
    ids = arr.ids.to(torch.int64)
    S = ids.shape[0]
    rb = (bits - 1) * 16 + 4
    out = torch.zeros(S, rb, dtype=torch.uint8, device=dev)
    k = arr.leaves.shape[0]
    nb = torch.diff(arr.block_off.to(torch.int64))
    leaf_of_slot = torch.repeat_interleave(torch.arange(k, device=dev), nb * BLOCK)
    leaves_rot = rotate(pad(arr.leaves.reshape(k, D).to(torch.float32)), arr.rotation)
    for s0 in range(0, S, chunk):
        sid = ids[s0:s0 + chunk]
        at = torch.nonzero((sid >= 0) & (sid != PAD)).squeeze(1)
        if at.numel() == 0:
            continue
        x = arr.base[sid[at]].to(torch.float32)
        r = (rotate(pad(x), arr.rotation) - leaves_rot[leaf_of_slot[s0 + at]]).contiguous()
        m = r.shape[0]
        on_dev = r.is_cuda
        planes = torch.empty(m, 32, dtype=torch.uint8, device=dev)
        f = torch.empty(m, dtype=torch.float32, device=dev)
        if on_dev:
            abi.check(abi.lib().sieve_l2_refine_encode(abi.ptr(r), m, bits, abi.ptr(planes), abi.ptr(f)))
        else:
            rn, pn, fn = r.numpy(), planes.numpy(), f.numpy()
            abi.check(abi.lib().sieve_l2_refine_encode(abi.ptr(rn), m, bits, abi.ptr(pn), abi.ptr(fn)))
        out[s0 + at] = torch.cat([planes[:, :(bits - 1) * 16], f.view(torch.uint8).reshape(m, 4)], 1)
    return out


CHUNK_ROWS = 100_000_000  # rows per base.NN file (src/bigann.rs:57-58)
RECORD_ALIGN = 4096       # postings records (src/postings.rs:38)


def write_l2_index(arr, index_dir, base_dir, refine_bits=1, refine=None, coarse=None, k2=None,
                   chunk_rows=CHUNK_ROWS):
    """Writes an L2 index in the reference's on-disk form: the Model files (src/bigann.rs:257-266:
    centroids.f32, coarse.f32, k2.u64, rotation.f32, bits.u64 when refine_bits > 1), the postings
    (src/bigann.rs:386-460: postings.bin with one 4 KiB-aligned record per leaf, its code blocks,
    [mul, |x|^2] per slot, ids and the refinement section, then postings.off and block_off.u64),
    and the base rows as base.NN chunks of chunk_rows rows with base.NN.done markers
    (src/bigann.rs:465-489). A flat index (no coarse cells given) is written as one cell per
    leaf: coarse = the leaves, k2 = 1."""
    import os
    os.makedirs(index_dir, exist_ok=True)
    os.makedirs(base_dir, exist_ok=True)
    h = arr.numpy()
    k = h["leaves"].reshape(-1, D).shape[0]
    leaves = np.ascontiguousarray(h["leaves"], np.float32).reshape(-1)
    coarse = leaves if coarse is None else np.ascontiguousarray(coarse, np.float32).reshape(-1)
    k2 = 1 if k2 is None else int(k2)
    leaves.tofile(os.path.join(index_dir, "centroids.f32"))
    coarse.tofile(os.path.join(index_dir, "coarse.f32"))
    np.array([k2], np.uint64).tofile(os.path.join(index_dir, "k2.u64"))
    np.ascontiguousarray(h["rotation"], np.float32).tofile(os.path.join(index_dir, "rotation.f32"))
    if refine_bits > 1:
        np.array([refine_bits], np.uint64).tofile(os.path.join(index_dir, "bits.u64"))
        ext = refine.cpu().numpy() if hasattr(refine, "cpu") else np.asarray(refine, np.uint8)
        rb = (refine_bits - 1) * 16 + 4
        ext = ext.reshape(-1, rb)
    bo = h["block_off"].astype(np.uint64)
    blocks = h["blocks"].reshape(-1, 512)
    fac = h["factors"].reshape(-1, 2).astype(np.float32)
    ids = h["ids"].astype(np.uint32)
    off = np.zeros(k + 1, np.uint64)
    pos = 0
    with open(os.path.join(index_dir, "postings.bin"), "wb") as f:
        for c in range(k):
            b0, b1 = int(bo[c]), int(bo[c + 1])
            off[c] = pos
            rec = [blocks[b0:b1].tobytes(), fac[b0 * 32:b1 * 32].tobytes(), ids[b0 * 32:b1 * 32].tobytes()]
            if refine_bits > 1:
                rec.append(ext[b0 * 32:b1 * 32].tobytes())
            body = b"".join(rec)
            padn = (-len(body)) % RECORD_ALIGN
            f.write(body + b"\0" * padn)
            pos += len(body) + padn
    off[k] = pos
    off.tofile(os.path.join(index_dir, "postings.off"))
    bo.tofile(os.path.join(index_dir, "block_off.u64"))
    base = h["base"].reshape(-1, D)
    for c, r0 in enumerate(range(0, base.shape[0], chunk_rows)):
        name = os.path.join(base_dir, f"base.{c:02d}")
        np.ascontiguousarray(base[r0:r0 + chunk_rows], np.int8).tofile(name)
        open(name + ".done", "w").close()


class L2Mixture:
    """The 100-d int8 mixture of the 1B-row L2 workload at any size: component means N(0, 40^2) per
    dimension; each component holds `topics` sub-centres at N(0, spread^2) around its mean, and
    a row is a uniformly drawn topic's centre + N(0, topic_spread^2), rounded and clipped to
    [-127, 127] (topics = 0: rows around the component mean at `spread`). Rows come in 2^22-row
    chunks, chunk c from its own seed, so any range can be regenerated; queries are further
    draws plus uniform integer noise in [-noise, noise]."""

    CHUNK = 1 << 22

    def __init__(self, components, seed, device="cuda", spread=12.0, topics=0, topic_spread=4.0):
        self.components, self.seed, self.device, self.spread = components, seed, device, spread
        self.topics, self.topic_spread = topics, topic_spread
        g = torch.Generator(device=device)
        g.manual_seed(sub_seed(seed, 11))
        self.means = torch.randn(components, D, generator=g, device=device) * 40.0
        self.centres = None
        if topics:
            tc = torch.empty(components * topics, D, device=device)
            for c0 in range(0, components, 8192):
                m = self.means[c0:c0 + 8192].repeat_interleave(topics, 0)
                tc[c0 * topics:c0 * topics + m.shape[0]] = m + spread * torch.randn(m.shape, generator=g,
                                                                                      device=device)
            self.centres = tc

    def _draw(self, n, g):
        if self.centres is None:
            comp = torch.randint(0, self.components, (n,), generator=g, device=self.device)
            x = self.means[comp] + self.spread * torch.randn(n, D, generator=g, device=self.device)
        else:
            t = torch.randint(0, self.centres.shape[0], (n,), generator=g, device=self.device)
            x = self.centres[t] + self.topic_spread * torch.randn(n, D, generator=g, device=self.device)
        return torch.clamp(torch.round(x), -127, 127).to(torch.int8)

    def fill(self, out, start=0):
        """Rows [start, start + len(out)) into the int8 tensor `out` (n x 100)."""
        n = out.shape[0]
        c0, c1 = start // self.CHUNK, (start + n + self.CHUNK - 1) // self.CHUNK
        g = torch.Generator(device=self.device)
        for c in range(c0, c1):
            g.manual_seed(sub_seed(self.seed, 12, c))
            x = self._draw(self.CHUNK, g)
            lo, hi = max(start - c * self.CHUNK, 0), min(start + n - c * self.CHUNK, self.CHUNK)
            out[c * self.CHUNK + lo - start:c * self.CHUNK + hi - start] = x[lo:hi]
        return out

    def queries(self, nq, noise=6):
        g = torch.Generator(device=self.device)
        g.manual_seed(sub_seed(self.seed, 13))
        q = self._draw(nq, g).to(torch.int32)
        q = q + torch.randint(-noise, noise + 1, q.shape, generator=g, device=self.device, dtype=torch.int32)
        return torch.clamp(q, -127, 127).to(torch.int8)


def rng_below(seed, count, n):
    """`count` draws of Rng(seed).below(n) (splitmix64, src/ivf.rs:120-135), vectorised: the i-th
    state is seed + i * 0x9E3779B97F4A7C15 (mod 2^64)."""
    i = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z % np.uint64(n)).astype(np.int64)


def _kmeans_l2_torch(x, k, iters, seed):
    """kmeans_l2 (src/bigann.rs:139-161) with torch's f32 GEMM for the scores (CPU fixtures):
    f64 sums in sample order, means as f32, empty clusters reseeded in cluster order."""
    from .build import Rng
    n = x.shape[0]
    rng = Rng(seed)
    cents = x[[rng.below(n) for _ in range(k)]].clone()
    for _ in range(iters):
        a = (2.0 * x @ cents.T - (cents * cents).sum(1)).argmax(1)
        sums = torch.zeros(k, D, dtype=torch.float64).index_add_(0, a, x.double())
        cnt = torch.bincount(a, minlength=k)
        upd = (sums / cnt.clamp(min=1)[:, None].double()).float()
        cents = torch.where((cnt > 0)[:, None], upd, cents)
        for c in (cnt == 0).nonzero().flatten().tolist():
            cents[c] = x[rng.below(n)]
    return cents


def train_l2(base, coarse, fine, iters=8, sample=1 << 24):
    """bigann::train (src/bigann.rs:175-227): the Rng(42) sample (indices sorted), coarse cells
    over rows at an even stride through it, every sample row to its nearest cell, and `fine`
    leaves per cell. On the GPU through sieve_build_l2_train, bit-exact with the reference; on
    the CPU (test fixtures) with torch's f32 GEMM. Returns (coarse k1 x 100, leaves k1 k2 x 100)."""
    n = base.shape[0]
    if not base.is_cuda:
        sample_n = min(n, sample)
        idx = np.sort(rng_below(42, sample_n, n))
        xs = base[torch.as_tensor(idx)].to(torch.float32)
        coarse_n = min(sample_n, coarse * 512)
        crow = xs[::sample_n // coarse_n][:coarse_n]
        cells = _kmeans_l2_torch(crow, coarse, iters, 7)
        a = (2.0 * xs @ cells.T - (cells * cells).sum(1)).argmax(1)
        leaves = torch.empty(coarse * fine, D, dtype=torch.float32)
        for c in range(coarse):
            xc = xs[a == c]
            leaves[c * fine:(c + 1) * fine] = (_kmeans_l2_torch(xc, fine, iters, c) if xc.shape[0]
                                               else cells[c:c + 1].repeat(fine, 1))
        return cells, leaves
    from . import abi
    cells = torch.empty(coarse, D, dtype=torch.float32, device=base.device)
    leaves = torch.empty(coarse * fine, D, dtype=torch.float32, device=base.device)
    abi.check(abi.lib().sieve_build_l2_train(base.data_ptr(), n, coarse, fine, min(n, sample), iters,
                                             cells.data_ptr(), leaves.data_ptr()))
    return cells, leaves


def assign_l2(base, cells, leaves, fine, chunk=1 << 16):
    """Model::assign (src/bigann.rs:277-302): each row's leaf, the best child of its 4 nearest
    coarse cells (a later cell in ascending order wins only if strictly better). On the GPU
    through sieve_build_l2_assign, bit-exact with the reference; on the CPU with torch's GEMM."""
    if not base.is_cuda:
        cn = (cells * cells).sum(1)
        ln = (leaves * leaves).sum(1)
        out = torch.empty(base.shape[0], dtype=torch.int32)
        for s0 in range(0, base.shape[0], chunk):
            x = base[s0:s0 + chunk].to(torch.float32)
            top = torch.topk(2.0 * x @ cells.T - cn, 4, dim=1).indices.sort(dim=1).values
            best = torch.full((x.shape[0],), -float("inf"))
            bl = torch.zeros(x.shape[0], dtype=torch.int64)
            for t in range(4):
                c = top[:, t]
                lc = leaves.view(-1, fine, D)[c]                                   # m x fine x 100
                sc = 2.0 * torch.einsum("md,mfd->mf", x, lc) - ln.view(-1, fine)[c]
                j = sc.argmax(1)
                s = sc.gather(1, j[:, None])[:, 0]
                win = s > best
                best = torch.where(win, s, best)
                bl = torch.where(win, c * fine + j, bl)
            out[s0:s0 + x.shape[0]] = bl.to(torch.int32)
        return out
    from . import abi
    leaf = torch.empty(base.shape[0], dtype=torch.int32, device=base.device)
    abi.check(abi.lib().sieve_build_l2_assign(base.data_ptr(), base.shape[0], cells.data_ptr(), cells.shape[0],
                                              leaves.data_ptr(), fine, leaf.data_ptr()))
    return leaf


def build_l2_large(base, coarse=512, fine=256, iters=8, sample=1 << 24, chunk=1 << 22, log=None):
    """The 1B build (src/bigann.rs:175-374) for a device-resident base: train_l2 (two-level
    k-means on an Rng(42) sample), assign_l2 (each row to the best leaf of its 4 nearest cells),
    then codes, mul and |x|^2 per chunk, slots sorted by (leaf, row) and packed into 512-byte
    blocks a chunk of blocks at a time. Training and assignment run in the library and equal the
    reference's bit for bit; the element-wise encode runs in the reference's f32 order. Peak
    device memory is about the base plus 90 bytes per row."""
    say = log or (lambda m: None)
    dev = base.device
    n = base.shape[0]
    k = coarse * fine
    base = base.contiguous()
    cents, leaves = train_l2(base, coarse, fine, iters, sample)
    say(f"l2 build: {coarse} x {fine} leaves")
    leaf = assign_l2(base, cents, leaves, fine)
    say("l2 build: rows assigned")
    signs = rotation_signs(11, PD)
    leaves_rot = rotate(pad(leaves), signs)
    codes = torch.empty(n, PD // 8, dtype=torch.uint8, device=dev)
    mul = torch.empty(n, dtype=torch.float32, device=dev)
    nx = torch.empty(n, dtype=torch.float32, device=dev)
    for s0 in range(0, n, chunk):
        x = base[s0:s0 + chunk].to(torch.float32)
        lf = leaf[s0:s0 + x.shape[0]].to(torch.int64)
        c_, m_, _ = rabitq_encode(rotate(pad(x), signs), leaves_rot[lf])
        codes[s0:s0 + x.shape[0]], mul[s0:s0 + x.shape[0]], nx[s0:s0 + x.shape[0]] = c_, m_, seq_norm2(x)
    say("l2 build: encoded")
    order = torch.argsort(leaf, stable=True)            # rows of each leaf ascending
    sizes = torch.bincount(leaf, minlength=k)
    del leaf
    nbl = (sizes + BLOCK - 1) // BLOCK
    block_off = torch.zeros(k + 1, dtype=torch.int64, device=dev)
    block_off[1:] = torch.cumsum(nbl, 0)
    first = torch.zeros(k + 1, dtype=torch.int64, device=dev)
    first[1:] = torch.cumsum(sizes, 0)
    S = int(block_off[-1]) * BLOCK
    ids = torch.full((S,), -1, dtype=torch.int32, device=dev)   # PAD
    slot_codes = torch.zeros(S, PD // 8, dtype=torch.uint8, device=dev)
    factors = torch.zeros(S, 2, dtype=torch.float32, device=dev)
    lsorted = torch.repeat_interleave(torch.arange(k, device=dev), sizes)
    for s0 in range(0, n, chunk):
        rows = order[s0:s0 + chunk]
        lf = lsorted[s0:s0 + rows.shape[0]]
        pos = block_off[lf] * BLOCK + torch.arange(s0, s0 + rows.shape[0], device=dev) - first[lf]
        ids[pos] = rows.to(torch.int32)
        slot_codes[pos] = codes[rows]
        factors[pos, 0] = mul[rows]
        factors[pos, 1] = nx[rows]
    del order, lsorted, codes, mul, nx
    nb = S // BLOCK
    blocks = torch.empty(nb, 512, dtype=torch.uint8, device=dev)
    step = chunk // BLOCK
    for b0 in range(0, nb, step):
        blocks[b0:b0 + step] = pack_blocks(slot_codes[b0 * BLOCK:(b0 + step) * BLOCK], PD // 4)
    del slot_codes
    say("l2 build: packed")
    return L2Arrays(leaves=leaves, block_off=block_off, blocks=blocks, ids=ids, factors=factors, rotation=signs,
                    base=base, coarse=cents)


def exact_l2_topk(base, q, k=32, chunk=1 << 22):
    """Exact top-k by 2 <q, x> - |x|^2 (nearest in L2) over every row: integer-valued products and
    sums below 2^24 are exact in f32. Returns (ids, scores) nq x k, best first."""
    qf = q.to(torch.float32)
    best_s = torch.full((q.shape[0], k), -float("inf"), device=q.device)
    best_i = torch.zeros((q.shape[0], k), dtype=torch.int64, device=q.device)
    for s0 in range(0, base.shape[0], chunk):
        x = base[s0:s0 + chunk].to(torch.float32)
        sc = 2.0 * qf @ x.T - (x * x).sum(1)
        v, i = torch.topk(sc, min(k, x.shape[0]), dim=1)
        cs = torch.cat([best_s, v], 1)
        ci = torch.cat([best_i, i + s0], 1)
        # best first, ties by row id (stable over the concatenation order)
        o = torch.argsort(cs, dim=1, descending=True, stable=True)[:, :k]
        best_s, best_i = torch.gather(cs, 1, o), torch.gather(ci, 1, o)
    return best_i, best_s


__all__ = ["L2Arrays", "L2Mixture", "int8_mixture", "build_l2", "build_l2_large", "exact_l2_topk", "f32_div",
           "refine_section"]
