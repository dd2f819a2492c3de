import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the product kernels")


def make_corpus(seed: int, D: int, V: int, mean_len: float, zipf: float = 1.3, max_len: int = 2000,
                empty_docs: int = 0):
    """Small random doc-major corpus (numpy RNG): Zipf-ish words, Poisson lengths."""
    rng = np.random.default_rng(seed)
    lens = np.clip(rng.poisson(mean_len, D), 1, max_len)
    if empty_docs:
        lens[rng.choice(D, empty_docs, replace=False)] = 0
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    words = ((rng.zipf(zipf, int(off[-1])) - 1) % V).astype(np.int32)
    return off, words


@pytest.fixture
def small_corpus():
    return make_corpus(1, 300, 400, 60)
