"""Seeded synthetic corpora: an anisotropic Gaussian mixture on the unit sphere.

The reference ingests real embeddings from parquet (src/data.rs:52-86); the workloads here
are synthetic stand-ins of the same shape: 768-d L2-normalised rows whose energy decays
along the dimensions, so the 256-d Matryoshka prefix carries most of it, drawn from C
components with log-normal (or Zipf) weights and a per-component diagonal spread; queries
come from the same mixture plus N(0, noise^2) per dimension and are renormalised.

Optionally each component is itself a mixture of `topics` sub-components (topic means spread
`topic_spread` around the component mean, topic counts proportional to component weight, rows
spread `spread` around their topic): local structure at two scales, as real embedding corpora
have. A flat Gaussian component of millions of rows has none: its 10th and 11th nearest
neighbours of a query differ by ~5e-4 in cosine, far below what a 1-bit code resolves.

Rows are generated in fixed chunks of CHUNK rows, each from its own generator seeded with
splitmix64(seed, chunk), so any range of rows can be regenerated on its own (exact ground
truth streams over the corpus instead of holding it). The stream depends on the torch
device type (CPU and GPU generators differ); everything built from one run is consistent.
"""
from dataclasses import dataclass

import numpy as np
import torch

DIM = 768
PDIM = 256
CHUNK = 1 << 16
_M64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    """One step of splitmix64 (the reference's Rng, src/ivf.rs:120-135) as a hash."""
    x = (x + 0x9E3779B97F4A7C15) & _M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def sub_seed(seed: int, *keys: int) -> int:
    s = seed & _M64
    for k in keys:
        s = splitmix64(s ^ (k & _M64))
    return s & ((1 << 63) - 1)


@dataclass(frozen=True)
class MixtureSpec:
    components: int = 1024
    seed: int = 20241203
    weights: str = "lognormal"     # "lognormal" (sigma 1) or "zipf" (exponent zipf_s)
    zipf_s: float = 1.1
    spread: float = 0.045          # per-dimension std scale of a component around its mean
    spectrum_knee: float = 32.0    # energy of dim i ~ 1 / (1 + i / knee)
    query_noise: float = 0.05
    topics: int = 0                # sub-components in all (0: flat components)
    topic_spread: float = 0.0      # per-dimension std scale of topic means around the component mean


def _spectrum(spec, device):
    i = torch.arange(DIM, dtype=torch.float32, device=device)
    return (1.0 + i / spec.spectrum_knee).rsqrt()


class Mixture:
    """Component means, diagonal spreads and the weight CDF of a MixtureSpec."""

    def __init__(self, spec: MixtureSpec, device="cpu"):
        self.spec = spec
        self.device = torch.device(device)
        g = torch.Generator(device=self.device)
        g.manual_seed(sub_seed(spec.seed, 1))
        C = spec.components
        s = _spectrum(spec, self.device)
        means = torch.randn(C, DIM, generator=g, device=self.device) * s
        self.means = means / means.norm(dim=1, keepdim=True)
        jitter = torch.exp(0.3 * torch.randn(C, DIM, generator=g, device=self.device))
        self.sig = spec.spread * s * jitter
        if spec.weights == "zipf":
            r = torch.arange(1, C + 1, dtype=torch.float64, device=self.device)
            w = r.pow(-spec.zipf_s)
            w = w[torch.randperm(C, generator=g, device=self.device)]
        else:
            w = torch.exp(torch.randn(C, generator=g, device=self.device, dtype=torch.float64))
        self.cdf = torch.cumsum(w / w.sum(), 0)
        self.cdf[-1] = 1.0
        self.tmeans = None
        if spec.topics > 0:
            # Topics per component ~ its weight (at least one): every topic holds about the same
            # number of rows whatever the component's size.
            cnt = torch.clamp(torch.round(w / w.sum() * spec.topics), min=1).to(torch.int64)
            self.tcount = cnt
            self.toff = torch.cumsum(cnt, 0) - cnt
            owner = torch.repeat_interleave(torch.arange(C, device=self.device), cnt)
            gt = torch.Generator(device=self.device)
            gt.manual_seed(sub_seed(spec.seed, 4))
            tz = torch.randn(int(cnt.sum()), DIM, generator=gt, device=self.device)
            self.tmeans = self.means[owner] + spec.topic_spread * s * jitter[owner] * tz

    def _draw(self, n, g):
        u = torch.rand(n, generator=g, device=self.device, dtype=torch.float64)
        comp = torch.searchsorted(self.cdf, u).clamp_(max=self.spec.components - 1)
        z = torch.randn(n, DIM, generator=g, device=self.device)
        if self.tmeans is None:
            x = self.means[comp] + self.sig[comp] * z
        else:
            v = torch.rand(n, generator=g, device=self.device, dtype=torch.float64)
            tc = self.tcount[comp]
            topic = self.toff[comp] + torch.minimum((v * tc).to(torch.int64), tc - 1)
            x = self.tmeans[topic] + self.sig[comp] * z
        return x / x.norm(dim=1, keepdim=True), comp

    def chunk(self, c):
        """Rows [c*CHUNK, (c+1)*CHUNK) as f32 (CHUNK x 768), unit length."""
        g = torch.Generator(device=self.device)
        g.manual_seed(sub_seed(self.spec.seed, 2, c))
        return self._draw(CHUNK, g)[0]

    def rows(self, start, count):
        out = []
        c0, c1 = start // CHUNK, (start + count + CHUNK - 1) // CHUNK
        for c in range(c0, c1):
            x = self.chunk(c)
            lo = max(start - c * CHUNK, 0)
            hi = min(start + count - c * CHUNK, CHUNK)
            out.append(x[lo:hi])
        return torch.cat(out) if out else torch.empty(0, DIM, device=self.device)

    def queries(self, nq):
        """(q768, q256): nq queries, each a mixture draw plus N(0, noise^2) noise, renormalised;
        q256 is the renormalised 256-d prefix."""
        g = torch.Generator(device=self.device)
        g.manual_seed(sub_seed(self.spec.seed, 3))
        x, _ = self._draw(nq, g)
        x = x + self.spec.query_noise * torch.randn(nq, DIM, generator=g, device=self.device)
        q768 = x / x.norm(dim=1, keepdim=True)
        p = q768[:, :PDIM]
        q256 = p / p.norm(dim=1, keepdim=True)
        return q768.contiguous(), q256.contiguous()


def np_queries(mix: Mixture, nq):
    q768, q256 = mix.queries(nq)
    return q768.cpu().numpy(), q256.cpu().numpy()
