"""ezlda_loglik_noalloc, the entry point that allocates no device memory: its
per-item partials and result live in the library's static device array.
Through the C ABI it must return exactly what ezlda_loglik returns (same
kernels, same summation order) on consecutive calls whose inputs shrink, so
stale contents of the shared array from a larger call never leak into a
smaller one; and it must refuse a corpus with more sampler items than the
array holds."""
import numpy as np
import pytest

from conftest import make_corpus
from lda_sc20_amd import ezlda
from oracle import binding as ob

pytestmark = pytest.mark.gpu

BETA = 0.01


def _ctx(off, words, V, K, seed=5, **kw):
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, **kw)
    ctx.load(ezlda.Corpus(V, off, words))
    return ctx


def test_noalloc_loglik_equals_loglik_on_shrinking_inputs():
    # first call: the larger corpus; second: a smaller one (fewer items, so
    # the array still holds partials of the first call past its end)
    big_off, big_words = make_corpus(31, 600, 900, 80, zipf=1.2)
    small_off, small_words = make_corpus(32, 40, 120, 30, zipf=1.2)
    cases = [(big_off, big_words, 900, 64), (small_off, small_words, 120, 16)]
    items = []
    for off, words, V, K in cases:
        ctx = _ctx(off, words, V, K)
        ctx.sweep(2)
        items.append(ctx.info().num_sample_items)
        ll_na = ctx.loglik_noalloc()
        ll = ctx.loglik()
        assert ll_na == ll, (K, ll_na, ll)
        orc = ob.Oracle(K, float(np.float32(50.0 / K)), BETA, 5, V, off, words)
        orc.sweep(2)
        ll_o = orc.loglik()
        assert abs(ll_na - ll_o) <= 1e-9 * abs(ll_o)
        # the state is unchanged by the call: a sweep after it is the oracle's
        ctx.sweep(1)
        orc.sweep(1)
        assert np.array_equal(ctx.get_topics(), orc.get_topics())
        assert ctx.loglik_noalloc() == ctx.loglik()
        ctx.close()
    assert items[1] < items[0]


def test_noalloc_loglik_wide_rows_and_chunked():
    off, words = make_corpus(33, 200, 400, 70, zipf=1.15)
    for kw, K in (({}, 10000), ({"num_chunks": 3}, 64)):
        ctx = _ctx(off, words, 400, K, **kw)
        ctx.sweep(2)
        a = ctx.loglik_noalloc()
        assert a == ctx.loglik(), (K, kw)
        ctx.close()


def test_noalloc_loglik_refuses_more_items_than_the_array_holds():
    # one token per word: every word is a tail word with one sampler item
    V = ezlda.LOGLIK_NOALLOC_MAX_ITEMS + 5000
    D = V // 100
    off = np.arange(0, 100 * D + 1, 100, dtype=np.int64)
    words = np.random.default_rng(3).permutation(V)[: 100 * D].astype(np.int32)
    ctx = ezlda.Context(K=8, alpha=50.0 / 8, beta=BETA, seed=5)
    ctx.load(ezlda.Corpus(V, off, words))
    assert ctx.info().num_sample_items > ezlda.LOGLIK_NOALLOC_MAX_ITEMS
    with pytest.raises(ezlda.EzldaError) as e:
        ctx.loglik_noalloc()
    assert e.value.code == ezlda.EINVAL and "at most" in str(e.value)
    assert np.isfinite(ctx.loglik())  # the allocating entry point still works
    ctx.close()
