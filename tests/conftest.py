import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "a-game-of-vectors_amd"))
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X")


class SmallWorkload:
    """A small seeded corpus + RaBitQ/SOAR IVF index built on the CPU, shared by tests."""

    def __init__(self, n=24000, k=48, nq=40, components=96, seed=7, tier2=True):
        import torch
        from sieve_amd import build, synth
        torch.manual_seed(0)
        self.mix = synth.Mixture(synth.MixtureSpec(components=components, seed=seed), device="cpu")
        self.corpus = build.ingest(self.mix, n, batch=8192)
        self.arrays = build.build(self.corpus, k, sample_n=min(n, k * 80), iters=6, tier2=tier2, chunk=8192)
        self.host = self.arrays.numpy()
        q768, q256 = self.mix.queries(nq)
        self.q768 = q768.numpy()
        self.q256 = q256.numpy()
        self.n, self.k, self.nq = n, k, nq


_cache = {}


@pytest.fixture(scope="session")
def small():
    if "small" not in _cache:
        _cache["small"] = SmallWorkload()
    return _cache["small"]


@pytest.fixture(scope="session")
def oracle_index(small):
    from oracle import oracle as O
    return O.Index(small.host)


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
