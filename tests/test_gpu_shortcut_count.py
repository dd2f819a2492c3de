"""GPU: the sampler's shortcut decisions, counted, against the oracle with the
same bound ((n_d - n_dk1) phi_k2, the oracle's shortcut mode 2)."""
import numpy as np
import pytest

from conftest import make_corpus
from lda_sc20_amd import ezlda
from oracle import binding as ob

pytestmark = pytest.mark.gpu

BETA = 0.01


@pytest.mark.parametrize("K", [8, 64, 257])
def test_full_draw_count_matches_the_oracle_with_the_same_bound(K):
    off, words = make_corpus(8, 150, 120, 80, zipf=1.3)
    V, seed = 120, 11
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, shortcut=1)
    ctx.load(ezlda.Corpus(V, off, words))
    orc = ob.Oracle(K, float(np.float32(50.0 / K)), BETA, seed, V, off, words, shortcut=2)
    ctx.reset_stats()
    for s in range(8):
        ctx.sweep(1)
        orc.sweep(1)
        assert np.array_equal(ctx.get_topics(), orc.get_topics()), (K, s)
    assert ctx.stats()[0] == orc.full_draws(), (K, ctx.stats()[0], orc.full_draws())
    assert orc.full_draws() < 8 * len(words), K
