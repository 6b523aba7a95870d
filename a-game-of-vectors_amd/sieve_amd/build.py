"""Ingest and IVF build: writes the reference's corpus and index files (or keeps them in memory).

This is the build side of the path, run once per workload; it uses torch on whatever device
is given (the MI355X for large corpora). It follows the reference's recipe:

  ingest   (src/bin/ingest.rs:117-140): normalise, 256-d prefix + its norm, sign codes,
           per-dimension int8 quantiser fitted on the first batch, int8 rows.
  build    (src/ivf.rs:268-395): sample Rng(42) prefixes reconstructed from the int8 rows
           (Corpus::prefix_f32, src/ivf.rs:110-116), spherical k-means seed 7
           (src/ivf.rs:188-225), 2-bit tier fit/encode (src/tier2.rs:23-50), assign_top(16) +
           SOAR secondary (src/ivf.rs:137-186), Rotation(11) + RaBitQ residual codes per
           (row, cluster) (src/rabitq.rs:86-108), cluster-major 32-slot blocks
           (src/ivf.rs:352-391, src/lut.rs:73-83).

Element-wise steps reproduce the reference's f32 arithmetic exactly (sums run sequentially in
the reference's order; no fused multiply-add), so the codes and [mul, err] factors written
for a given (row, cluster) are bit-identical to the reference's. On the GPU, k-means, the
assignment (assign_top + SOAR) and the tier-2 fit run in the library (sieve_build_kmeans,
sieve_build_assign, sieve_build_tier2_fit): every score is the reference's sequential f32 dot,
so the whole index is the reference's. On the CPU (test fixtures) they run on torch's f32 GEMM,
whose summation order differs; there a row within rounding of a tie between two centroids may
be assigned differently.
"""
import json
import os
from dataclasses import dataclass, field

import numpy as np
import torch

from . import synth

DIM = 768
PDIM = 256
BLOCK = 32
PAD = 0xFFFFFFFF
EPS0 = 1.9
T2_T = 0.9816
T2_LEVELS = (-1.510, -0.4528, 0.4528, 1.510)


# ----------------------------------------------------------------------- f32 helpers ----
def seq_sum_sq(x):
    """Σ x_i^2 left to right from -0.0, per row (Rust `map(|x| x * x).sum()`)."""
    xt = x.t().contiguous()  # one contiguous row per dimension
    acc = torch.full((x.shape[0],), -0.0, dtype=torch.float32, device=x.device)
    for i in range(xt.shape[0]):
        c = xt[i]
        acc = acc + c * c
    return acc


def f32_div(a, b):
    """Correctly rounded f32 a / b: the f64 quotient rounded once to f32 (f64 carries more than
    2 * 24 + 2 bits, so the double rounding is exact). torch's own f32 paths evaluate
    `scalar / tensor` as reciprocal * scalar and may use approximate device division."""
    a = a if torch.is_tensor(a) else torch.tensor(a, dtype=torch.float64, device=b.device)
    return (a.double() / b.double()).float()


def f32_sqrt(a):
    """Correctly rounded f32 sqrt (torch's vectorised f32 sqrt is not always)."""
    return a.double().sqrt().float()


def normalize_rows(x):
    """quant::normalize (src/quant.rs:5-13) per row; returns (unit rows, norms)."""
    n = f32_sqrt(seq_sum_sq(x))
    inv = torch.where(n > 0, f32_div(1.0, n), torch.ones_like(n))
    return x * inv[:, None], n


def dot_f32_rows(a, b):
    """quant::dot_f32 (src/quant.rs:86-100) row-wise for 256-d rows: 16 lane sums, then the
    lanes left to right from -0.0."""
    d = a.shape[-1]
    acc = torch.zeros(a.shape[:-1] + (16,), dtype=torch.float32, device=a.device)
    for i in range(0, d, 16):
        acc = acc + a[..., i:i + 16] * b[..., i:i + 16]
    s = torch.full(a.shape[:-1], -0.0, dtype=torch.float32, device=a.device)
    for j in range(16):
        s = s + acc[..., j]
    return s


def round_half_away(x):
    """f32::round: half away from zero. x - trunc(x) is exact, so the tie test is too."""
    t = torch.trunc(x)
    return t + torch.where(torch.abs(x - t) >= 0.5, torch.sign(x), torch.zeros_like(x))


def pack_bits(bits):
    """bits (n x d bool), bit i -> byte i/8 position i%8 (src/quant.rs:24-34)."""
    n, d = bits.shape
    w = (1 << torch.arange(8, device=bits.device, dtype=torch.int32))
    return (bits.view(n, d // 8, 8).to(torch.int32) * w).sum(-1).to(torch.uint8)


def np_f32_pow(d, e):
    return float(np.power(np.float32(d), np.float32(e)))


# ----------------------------------------------------------------------- rotation -------
def rotation_signs(seed=11, d=PDIM):
    """Rotation::with_dim (src/rabitq.rs:39-44): 3 rounds of d signs from splitmix64."""
    st = seed & ((1 << 64) - 1)
    out = np.empty(3 * d, np.float32)
    for i in range(3 * d):
        st = (st + 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
        z ^= z >> 31
        out[i] = 1.0 if z & 1 else -1.0
    return out


def rotate(x, signs):
    """Rotation::apply (src/rabitq.rs:59-69) on every row: sign flip + unnormalised FWHT,
    three rounds, then scale by d^-1.5."""
    n, d = x.shape
    s = torch.as_tensor(signs, device=x.device).view(3, d)
    v = x.clone()
    for r in range(3):
        v = v * s[r]
        h = 1
        while h < d:
            v4 = v.view(n, d // (2 * h), 2, h)
            a, b = v4[:, :, 0, :], v4[:, :, 1, :]
            v = torch.stack((a + b, a - b), dim=2).view(n, d)
            h *= 2
    return v * np_f32_pow(d, -1.5)


def rabitq_encode(xr, cr):
    """rabitq::encode (src/rabitq.rs:86-108) row-wise: (codes n x d/8, mul n, err n)."""
    r = xr - cr
    n, d = r.shape
    rt = r.t().contiguous()
    nr2 = torch.zeros(n, dtype=torch.float32, device=r.device)
    l1 = torch.zeros(n, dtype=torch.float32, device=r.device)
    for i in range(d):
        c = rt[i]
        nr2 = nr2 + c * c
        l1 = l1 + torch.abs(c)
    code = pack_bits(r > 0)
    df = float(d)
    nr = f32_sqrt(nr2)
    od = f32_div(l1, nr * np.float32(np.sqrt(np.float32(df))))
    inner = f32_div(1.0 - od * od, od * od * np.float32(df - 1.0))
    inner = torch.where(inner > 0, inner, torch.zeros_like(inner))
    err = nr * np.float32(EPS0) * f32_sqrt(inner)
    mul = f32_div(nr2, l1)
    zero = l1 < 1e-12
    mul = torch.where(zero, torch.zeros_like(mul), mul)
    err = torch.where(zero, torch.zeros_like(err), err)
    return code, mul, err


# ----------------------------------------------------------------------- pack ----------
def pack_blocks(codes, G):
    """lut::pack_block (src/lut.rs:73-83) for slots laid out 32 per block: codes is
    (nblocks*32) x (G/2) bytes, returns nblocks x (G*16) bytes."""
    S = codes.shape[0]
    nb = S // BLOCK
    c = codes.to(torch.int32)
    nib = torch.stack((c & 0x0F, c >> 4), dim=2).view(S, G)            # nibble g of each slot
    nib = nib.view(nb, 2, 16, G)                                       # [block, hi?, k&15, g]
    out = nib[:, 0] | (nib[:, 1] << 4)                                 # [block, k, g]
    return out.permute(0, 2, 1).contiguous().to(torch.uint8).view(nb, G * 16)


# ----------------------------------------------------------------------- ingest --------
@dataclass
class Corpus:
    """Ingest output (src/bin/ingest.rs:99-166), as tensors."""
    rerank: torch.Tensor      # n x 768 int8
    quant: torch.Tensor       # 768 f32
    norm256: torch.Tensor     # n f32
    codes: torch.Tensor       # n x 32 u8 (global sign codes of the prefix)

    @property
    def n(self):
        return self.rerank.shape[0]


def ingest_batch(emb, quant):
    """normalize -> prefix -> sign code -> int8 (src/bin/ingest.rs:125-132)."""
    v, _ = normalize_rows(emb)
    p, norm = normalize_rows(v[:, :PDIM].contiguous())
    codes = pack_bits(p > 0)
    if quant is None:
        maxabs = torch.clamp(v.abs().amax(0), min=1e-6)                 # Int8Quantizer::fit
        quant = f32_div(127.0, maxabs)
    q = round_half_away(v * quant).clamp(-127, 127).to(torch.int8)
    return q, quant, norm, codes


def ingest(mix: synth.Mixture, n, batch=65536):
    dev = mix.device
    rerank = torch.empty(n, DIM, dtype=torch.int8, device=dev)
    norm = torch.empty(n, dtype=torch.float32, device=dev)
    codes = torch.empty(n, PDIM // 8, dtype=torch.uint8, device=dev)
    quant = None
    for s in range(0, n, batch):
        m = min(batch, n - s)
        q, quant, nr, cd = ingest_batch(mix.rows(s, m), quant)
        rerank[s:s + m], norm[s:s + m], codes[s:s + m] = q, nr, cd
    return Corpus(rerank, quant, norm, codes)


def prefix_f32(corpus: Corpus, rows):
    """Corpus::prefix_f32 (src/ivf.rs:110-116): the unit 256-d prefix rebuilt from int8 rows."""
    x = f32_div(corpus.rerank[rows, :PDIM].to(torch.float32), corpus.quant[:PDIM])
    return normalize_rows(x)[0]


# ----------------------------------------------------------------------- k-means -------
class Rng:
    """splitmix64 (src/ivf.rs:120-135)."""

    def __init__(self, seed):
        self.s = seed & ((1 << 64) - 1)

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
        return z ^ (z >> 31)

    def below(self, n):
        return self.next() % n


def _scores(rows, cents):
    return rows @ cents.T


def assign_top(rows, cents, m, block=65536):
    """The m best centroids of each row by inner product, best first (src/ivf.rs:137-158)."""
    out = []
    for s in range(0, rows.shape[0], block):
        sc = _scores(rows[s:s + block], cents)
        out.append(torch.topk(sc, m, dim=1).indices if m > 1 else sc.argmax(1, keepdim=True))
    return torch.cat(out).to(torch.int64)


def _lib():
    from . import abi
    return abi


def kmeans(sample, k, iters, seed=7):
    """Spherical k-means (src/ivf.rs:188-225, balance off). On the GPU through
    sieve_build_kmeans, bit-exact with the reference; on the CPU with torch's f32 GEMM and
    scatter sums (a row within rounding of a tie may be assigned differently)."""
    if sample.is_cuda:
        abi = _lib()
        x = sample.contiguous()
        cents = torch.empty(k, PDIM, dtype=torch.float32, device=sample.device)
        abi.check(abi.lib().sieve_build_kmeans(x.data_ptr(), x.shape[0], k, iters, seed, cents.data_ptr()))
        return cents
    n = sample.shape[0]
    rng = Rng(seed)
    idx = [rng.below(n) for _ in range(k)]
    cents = sample[torch.tensor(idx, device=sample.device)].clone()
    for _ in range(iters):
        a = assign_top(sample, cents, 1)[:, 0]
        sums = torch.zeros_like(cents).index_add_(0, a, sample)
        counts = torch.bincount(a, minlength=k)
        empty = (counts == 0).nonzero().flatten().tolist()
        cents = normalize_rows(sums)[0]
        for c in empty:
            cents[c] = sample[rng.below(n)]
    return cents


def assign_pairs(x, cents, soar_lambda, soar_cand=16):
    """The clusters of each row as ivf::build assigns them (src/ivf.rs:313-321): (n, 2) primary and
    SOAR secondary, or (n, 1) without SOAR. On the GPU through sieve_build_assign (bit-exact with
    the reference up to exact f32 ties); on the CPU with torch's f32 GEMM."""
    if x.is_cuda:
        abi = _lib()
        xc, cc = x.contiguous(), cents.contiguous()
        pairs = torch.empty(xc.shape[0], 2, dtype=torch.int32, device=x.device)
        lam = float("nan") if soar_lambda is None else float(soar_lambda)
        abi.check(abi.lib().sieve_build_assign(xc.data_ptr(), xc.shape[0], cc.data_ptr(), cc.shape[0], lam,
                                               pairs.data_ptr()))
        pairs = pairs.to(torch.int64)
        return pairs[:, :1] if soar_lambda is None else pairs
    if soar_lambda is None:
        return assign_top(x, cents, 1)
    top = assign_top(x, cents, soar_cand)
    return torch.stack((top[:, 0], soar_secondary(x, cents, top, soar_lambda)), dim=1)


def soar_secondary(x, cents, cand, lam):
    """SOAR secondary cluster (src/ivf.rs:164-186) for every row of x given its candidates."""
    c1 = cents[cand[:, 0]]
    x_c1 = dot_f32_rows(x, c1)
    r2 = torch.clamp(2.0 - 2.0 * x_c1, min=1e-12)
    r_x = 1.0 - x_c1
    cv = cents[cand[:, 1:]]                                             # n x (m-1) x 256
    x_c = dot_f32_rows(x[:, None, :].expand_as(cv), cv)
    r_c = x_c - dot_f32_rows(c1[:, None, :].expand_as(cv), cv)
    along = r_x[:, None] - r_c
    loss = (2.0 - 2.0 * x_c) + f32_div(np.float32(lam) * along * along, r2[:, None].expand_as(along))
    best = torch.argmin(loss, dim=1)
    return cand[:, 1:].gather(1, best[:, None])[:, 0]


# ----------------------------------------------------------------------- tier2 ---------
def tier2_fit(sample):
    """Tier2::fit (src/tier2.rs:20-27). On the GPU through sieve_build_tier2_fit (sequential f64
    sums, as the reference's); on the CPU with torch's f64 sum."""
    if sample.is_cuda:
        abi = _lib()
        x = sample.contiguous()
        sigma = torch.empty(PDIM, dtype=torch.float32, device=sample.device)
        abi.check(abi.lib().sieve_build_tier2_fit(x.data_ptr(), x.shape[0], sigma.data_ptr()))
        return sigma
    n = float(np.float32(max(sample.shape[0], 1)))
    ss = (sample * sample).double().sum(0)
    return torch.clamp((ss / torch.full_like(ss, n)).sqrt().float(), min=1e-6)


def tier2_encode(prefix, sigma):
    """Tier2::encode (src/tier2.rs:36-47): (codes n x 64, reconstruction norms n)."""
    t = np.float32(T2_T) * sigma
    lv = torch.where(prefix < -t, 0, torch.where(prefix < 0, 1, torch.where(prefix < t, 2, 3)))
    levels = torch.tensor(T2_LEVELS, dtype=torch.float32, device=prefix.device)
    v = levels[lv] * sigma
    vt = v.t().contiguous()
    nn = torch.zeros(prefix.shape[0], dtype=torch.float32, device=prefix.device)
    for i in range(PDIM):
        nn = nn + vt[i] * vt[i]
    sh = (2 * torch.arange(4, device=prefix.device, dtype=torch.int32))
    codes = (lv.view(-1, PDIM // 4, 4).to(torch.int32) << sh).sum(-1).to(torch.uint8)
    return codes, f32_sqrt(nn)


# ----------------------------------------------------------------------- build ---------
@dataclass
class IndexArrays:
    """The build output (src/ivf.rs:354-391) as tensors, plus the corpus it indexes."""
    centroids: torch.Tensor
    block_off: torch.Tensor     # int64 (written as u64)
    blocks: torch.Tensor        # nblocks x 1024 u8
    ids: torch.Tensor           # nblocks*32 int64 with PAD (written as u32)
    factors: torch.Tensor       # nblocks*32 x 2 f32
    rotation: np.ndarray        # 768 f32
    corpus: Corpus
    tier2_sigma: torch.Tensor = None
    tier2_codes: torch.Tensor = None
    tier2_norm: torch.Tensor = None
    sizes: list = field(default_factory=list)
    soar_lambda: float = 1.0    # None: built without SOAR spills

    @property
    def k(self):
        return self.centroids.shape[0]

    @property
    def nblocks(self):
        return self.blocks.shape[0]

    def device(self):
        """The arrays as device tensors in the on-disk dtypes (u32 / u64 as same-width ints),
        for Index.create with the rows borrowed in place."""
        c = self.corpus
        ids = torch.where(self.ids == PAD, torch.full_like(self.ids, -1), self.ids).to(torch.int32)
        d = dict(centroids=self.centroids.contiguous(), block_off=self.block_off.contiguous(),
                 blocks=self.blocks.contiguous(), ids=ids.contiguous(), factors=self.factors.contiguous(),
                 rotation=torch.as_tensor(self.rotation, device=c.rerank.device), rerank=c.rerank,
                 quant=c.quant.contiguous(), norm256=c.norm256.contiguous())
        if self.tier2_sigma is not None:
            d.update(tier2_sigma=self.tier2_sigma.contiguous(), tier2_codes=self.tier2_codes.contiguous(),
                     tier2_norm=self.tier2_norm.contiguous())
        return d

    def numpy(self):
        """Host arrays in the on-disk dtypes (the layout both the device index and the oracle
        read)."""
        c = self.corpus
        out = dict(centroids=self.centroids.cpu().numpy(), block_off=self.block_off.cpu().numpy().astype(np.uint64),
                   blocks=self.blocks.cpu().numpy(), ids=self.ids.cpu().numpy().astype(np.uint32),
                   factors=self.factors.cpu().numpy(), rotation=self.rotation, rerank=c.rerank.cpu().numpy(),
                   quant=c.quant.cpu().numpy(), norm256=c.norm256.cpu().numpy())
        if self.tier2_sigma is not None:
            out.update(tier2_sigma=self.tier2_sigma.cpu().numpy(), tier2_codes=self.tier2_codes.cpu().numpy(),
                       tier2_norm=self.tier2_norm.cpu().numpy())
        return out


def build(corpus: Corpus, k, sample_n=None, iters=12, soar_lambda=1.0, rabitq=True, tier2=True, chunk=1 << 18,
          soar_cand=16):
    """ivf::build (src/ivf.rs:264-395) over an in-memory corpus."""
    dev = corpus.rerank.device
    n = corpus.n
    sample_n = min(sample_n or max(k * 64, 1), n)
    rng = Rng(42)
    pick = torch.tensor([rng.below(n) for _ in range(sample_n)], device=dev)
    sample = prefix_f32(corpus, pick)
    cents = kmeans(sample, k, iters, seed=7)
    t2_sigma = tier2_fit(sample) if tier2 else None
    del sample
    signs = rotation_signs(11, PDIM)
    cents_rot = rotate(cents, signs) if rabitq else None

    t2_codes = torch.empty(n, PDIM // 4, dtype=torch.uint8, device=dev) if tier2 else None
    t2_norm = torch.empty(n, dtype=torch.float32, device=dev) if tier2 else None
    slot_cluster, slot_row, slot_code, slot_mul, slot_err = [], [], [], [], []
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        rows = torch.arange(s, s + m, device=dev)
        x = prefix_f32(corpus, rows)
        if tier2:
            t2_codes[s:s + m], t2_norm[s:s + m] = tier2_encode(x, t2_sigma)
        pairs = assign_pairs(x, cents, soar_lambda, soar_cand)
        xr = rotate(x, signs) if rabitq else None
        for h in range(pairs.shape[1]):
            c = pairs[:, h]
            slot_cluster.append(c)
            slot_row.append(rows)
            if rabitq:
                code, mul, err = rabitq_encode(xr, cents_rot[c])
                slot_code.append(code)
                slot_mul.append(mul)
                slot_err.append(err)
            else:
                slot_code.append(corpus.codes[rows])
    cl = torch.cat(slot_cluster)
    rw = torch.cat(slot_row)
    # Members of each cluster in row order (src/ivf.rs:342-349): sort by (cluster, row).
    order = torch.argsort(cl * n + rw)
    cl, rw = cl[order], rw[order]
    code = torch.cat(slot_code)[order]
    sizes = torch.bincount(cl, minlength=k)
    nbl = (sizes + BLOCK - 1) // BLOCK
    block_off = torch.zeros(k + 1, dtype=torch.int64, device=dev)
    block_off[1:] = torch.cumsum(nbl, 0)
    nblocks = int(block_off[-1])
    first = torch.zeros(k + 1, dtype=torch.int64, device=dev)
    first[1:] = torch.cumsum(sizes, 0)
    rank = torch.arange(cl.shape[0], device=dev) - first[cl]
    pos = block_off[cl] * BLOCK + rank
    S = nblocks * BLOCK
    ids = torch.full((S,), PAD, dtype=torch.int64, device=dev)
    ids[pos] = rw
    slot_codes = torch.zeros(S, PDIM // 8, dtype=torch.uint8, device=dev)
    slot_codes[pos] = code
    blocks = pack_blocks(slot_codes, PDIM // 4)
    factors = torch.zeros(S, 2, dtype=torch.float32, device=dev)
    if rabitq:
        factors[pos, 0] = torch.cat(slot_mul)[order]
        factors[pos, 1] = torch.cat(slot_err)[order]
    return IndexArrays(centroids=cents, block_off=block_off, blocks=blocks, ids=ids,
                       factors=factors if rabitq else None, rotation=signs if rabitq else None, corpus=corpus,
                       tier2_sigma=t2_sigma, tier2_codes=t2_codes, tier2_norm=t2_norm,
                       sizes=sizes.cpu().tolist(), soar_lambda=soar_lambda)


# ----------------------------------------------------------------------- ground truth --
def exact_topk(mix: synth.Mixture, n, q768, k=100, block=1 << 18):
    """Exact top-k by f32 inner product over the 768-d unit rows (the reference's Truth,
    src/bin/ingest.rs:42-78), streamed over the corpus. Returns (ids nq x k, scores)."""
    nq = q768.shape[0]
    best_s = torch.full((nq, k), float("-inf"), device=q768.device)
    best_i = torch.full((nq, k), -1, dtype=torch.int64, device=q768.device)
    for s in range(0, n, block):
        m = min(block, n - s)
        sc = q768 @ mix.rows(s, m).T
        kk = min(k, m)
        ts, ti = torch.topk(sc, kk, dim=1)
        cs = torch.cat((best_s, ts), 1)
        ci = torch.cat((best_i, ti + s), 1)
        top = torch.topk(cs, k, dim=1)
        best_s = top.values
        best_i = ci.gather(1, top.indices)
    return best_i, best_s


# ----------------------------------------------------------------------- files ---------
def _w(path, arr):
    np.ascontiguousarray(arr).tofile(path)


def write_index(arrays: IndexArrays, index_dir, ingest_dir, queries=None, gt=None):
    """Writes the reference's files: index dir (src/ivf.rs:354-391) and ingest dir
    (src/bin/ingest.rs:99-166)."""
    os.makedirs(index_dir, exist_ok=True)
    os.makedirs(ingest_dir, exist_ok=True)
    a = arrays.numpy()
    _w(os.path.join(index_dir, "centroids.f32"), a["centroids"])
    _w(os.path.join(index_dir, "block_off.u64"), a["block_off"])
    _w(os.path.join(index_dir, "blocks.bin"), a["blocks"])
    _w(os.path.join(index_dir, "ids.u32"), a["ids"])
    if a.get("factors") is not None and arrays.factors is not None:
        _w(os.path.join(index_dir, "factors.f32"), a["factors"])
        _w(os.path.join(index_dir, "rotation.f32"), a["rotation"])
    if "tier2_sigma" in a:
        _w(os.path.join(index_dir, "tier2.sigma"), a["tier2_sigma"])
        _w(os.path.join(index_dir, "tier2.bin"), a["tier2_codes"])
        _w(os.path.join(index_dir, "tier2.norm"), a["tier2_norm"])
    # meta.json as src/bin/index.rs:53-60 writes it (serde_json's sorted keys, two-space indent);
    # Live::open reads its soar_lambda (src/live.rs:170-171).
    meta = {"k": int(arrays.k), "rabitq": arrays.factors is not None, "rows": int(arrays.corpus.n),
            "soar_lambda": None if arrays.soar_lambda is None else float(arrays.soar_lambda)}
    with open(os.path.join(index_dir, "meta.json"), "w") as f:
        f.write(json.dumps(meta, indent=2, sort_keys=True))
    _w(os.path.join(ingest_dir, "rerank.i8"), a["rerank"])
    _w(os.path.join(ingest_dir, "quant.f32"), a["quant"])
    _w(os.path.join(ingest_dir, "norm256.f32"), a["norm256"])
    _w(os.path.join(ingest_dir, "codes.bin"), arrays.corpus.codes.cpu().numpy())
    n = arrays.corpus.n
    names = "".join(str(i) for i in range(n)).encode()
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(str(i)) for i in range(n)])
    with open(os.path.join(ingest_dir, "docids.bin"), "wb") as f:
        f.write(names)
    _w(os.path.join(ingest_dir, "docids.off"), off)
    if queries is not None:
        _w(os.path.join(ingest_dir, "queries.f32"), np.asarray(queries, np.float32))
    if gt is not None:
        ids, scores = gt
        _w(os.path.join(ingest_dir, "gt768.ids"), np.asarray(ids, np.uint32))
        _w(os.path.join(ingest_dir, "gt768.scores"), np.asarray(scores, np.float32))


def read_index(index_dir, ingest_dir):
    """The reference's files as host arrays (Index::open + Corpus::open)."""
    def r(d, name, dt):
        p = os.path.join(d, name)
        return np.fromfile(p, dtype=dt) if os.path.exists(p) else None
    a = dict(centroids=r(index_dir, "centroids.f32", np.float32), block_off=r(index_dir, "block_off.u64", np.uint64),
             blocks=r(index_dir, "blocks.bin", np.uint8), ids=r(index_dir, "ids.u32", np.uint32),
             factors=r(index_dir, "factors.f32", np.float32), rotation=r(index_dir, "rotation.f32", np.float32),
             tier2_sigma=r(index_dir, "tier2.sigma", np.float32), tier2_codes=r(index_dir, "tier2.bin", np.uint8),
             tier2_norm=r(index_dir, "tier2.norm", np.float32), rerank=r(ingest_dir, "rerank.i8", np.int8),
             quant=r(ingest_dir, "quant.f32", np.float32), norm256=r(ingest_dir, "norm256.f32", np.float32))
    return a
