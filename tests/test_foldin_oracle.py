"""The oracle's fold-in (frozen-model inference of new documents).

Fold-in has no counterpart in the reference (EZLDA trains only), so its
oracle is pinned by what the mathematics fixes: each fold-in sweep draws every
token of a new document from p(k) ~ (n_dk + alpha) * phi_wk, with n_dk the
document's counts before the sweep (Jacobi, own token included, as in the
training sweep of EZLDA/SamplingKernel.cu) and phi_wk = (n_wk + beta) /
(n_k + V beta) of the frozen model.  The tests recompute that inverse CDF in
FP64 from the public counts and the Philox stream, and check that results
depend only on (seed, request id, document), not on batch order.
"""
import numpy as np
import pytest

from conftest import make_corpus
from oracle import binding as ob

BETA = 0.01
TAG_FOLDIN, TAG_FOLDIN_INIT = 3, 4


def _model(K=48, V=200, seed=5, sweeps=3):
    off, words = make_corpus(11, 150, V, 50, zipf=1.2)
    alpha = float(np.float32(50.0 / K))
    orc = ob.Oracle(K, alpha, BETA, seed, V, off, words)
    orc.sweep(sweeps)
    orc.build_wt()  # counts of the final topics: the model that fold-in freezes
    return orc, alpha


def _x0(seed, rid, i, sweep, tag):
    ctr = np.array([i, rid, sweep, tag], np.uint32)
    key = np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32)
    return int(ob.philox(ctr, key)[0])


def _u(seed, rid, i, sweep):
    return (_x0(seed, rid, i, sweep, TAG_FOLDIN) >> 8) * 2.0**-24


def _new_docs(V, lens, seed=3):
    rng = np.random.default_rng(seed)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    words = ((rng.zipf(1.2, int(off[-1])) - 1) % V).astype(np.int32)
    return off, words


def _fp64_phi(orc, V, K):
    nwk = orc.get_wt_dense().astype(np.float64)
    nk = orc.get_nk().astype(np.float64)
    return (nwk + BETA) / (nk + V * BETA)[None, :]


def _check_inverse_cdf(z_prev, z_next, words, phi, alpha, K, seed, rid, sweep, tol=1e-4):
    """Number of tokens whose draw is outside its FP64 inverse-CDF interval.

    The mass is laid out as the two buckets the samplers walk: first the
    document part n_dk phi_wk over ascending k, then the smoothing part
    alpha phi_wk over ascending k; topic k owns one interval in each."""
    cnt = np.bincount(z_prev, minlength=K).astype(np.float64)
    bad = 0
    for i, w in enumerate(words):
        cdf = np.cumsum(np.concatenate([cnt * phi[w], alpha * phi[w]]))
        t = _u(seed, rid, i, sweep) * cdf[-1]
        k = int(z_next[i])
        ok = False
        for j in (k, K + k):
            lo = cdf[j - 1] if j > 0 else 0.0
            ok |= cdf[j] > lo and lo - tol * cdf[-1] <= t <= cdf[j] + tol * cdf[-1]
        bad += not ok
    return bad


def test_foldin_initial_topics_follow_init_stream():
    orc, alpha = _model()
    K, V, seed = orc.K, orc.V, 77
    off, words = _new_docs(V, [30])
    rid = 9
    z0 = np.array([_x0(seed, rid, i, 0, TAG_FOLDIN_INIT) % K for i in range(30)], np.uint16)
    phi = _fp64_phi(orc, V, K)
    z1, _ = orc.foldin(off, words, [rid], 1, seed)
    assert _check_inverse_cdf(z0, z1, words, phi, alpha, K, seed, rid, 0) == 0


@pytest.mark.parametrize("sweep", [1, 2, 5])
def test_foldin_sweep_is_the_fp64_inverse_cdf(sweep):
    orc, alpha = _model()
    K, V, seed = orc.K, orc.V, 123
    off, words = _new_docs(V, [1, 7, 64, 65, 300])
    rids = np.array([3, 1 << 20, 5, 0xFFFFFFFF, 17], np.uint32)
    phi = _fp64_phi(orc, V, K)
    za, _ = orc.foldin(off, words, rids, sweep, seed)
    zb, _ = orc.foldin(off, words, rids, sweep + 1, seed)
    for r in range(len(rids)):
        s, e = off[r], off[r + 1]
        assert _check_inverse_cdf(za[s:e], zb[s:e], words[s:e], phi, alpha, K, seed, int(rids[r]), sweep) == 0
    # negative control: the uniforms of another sweep do not explain the draws
    s, e = off[4], off[5]
    assert _check_inverse_cdf(za[s:e], zb[s:e], words[s:e], phi, alpha, K, seed, int(rids[4]), sweep + 1) > 30


def test_foldin_depends_only_on_request_not_on_batch():
    orc, _ = _model()
    V = orc.V
    lens = [40, 0, 5, 120, 1, 77]
    off, words = _new_docs(V, lens, seed=8)
    rids = np.arange(100, 100 + len(lens), dtype=np.uint32)
    z, cnt = orc.foldin(off, words, rids, 6, 42)
    perm = [3, 0, 5, 1]
    sub_lens = [lens[p] for p in perm]
    sub_off = np.concatenate([[0], np.cumsum(sub_lens)]).astype(np.int64)
    sub_words = np.concatenate([words[off[p]: off[p + 1]] for p in perm]).astype(np.int32)
    z2, cnt2 = orc.foldin(sub_off, sub_words, rids[perm], 6, 42)
    for j, p in enumerate(perm):
        assert np.array_equal(z2[sub_off[j]: sub_off[j + 1]], z[off[p]: off[p + 1]])
        assert np.array_equal(cnt2[j], cnt[p])
    # counts are the final topics' histogram; another seed gives other topics
    for r in range(len(lens)):
        assert np.array_equal(cnt[r], np.bincount(z[off[r]: off[r + 1]], minlength=orc.K))
    z3, _ = orc.foldin(off, words, rids, 6, 43)
    assert not np.array_equal(z3, z)


def test_foldin_recovers_a_planted_topic():
    """A document made of one topic's most likely words folds into that topic."""
    orc, _ = _model(K=16, sweeps=30)
    K, V = orc.K, orc.V
    phi = _fp64_phi(orc, V, K)
    k = int(np.argmax(orc.get_nk()))
    top = np.argsort(-phi[:, k])[:5]
    words = np.repeat(top, 40).astype(np.int32)
    off = np.array([0, len(words)], np.int64)
    _, cnt = orc.foldin(off, words, [1], 20, 9)
    assert int(np.argmax(cnt[0])) == k and cnt[0][k] > 0.5 * len(words)


def test_foldin_single_topic():
    orc = ob.Oracle(1, 50.0, BETA, 1, 20, *make_corpus(2, 10, 20, 10))
    orc.sweep(1)
    orc.build_wt()
    off, words = _new_docs(20, [10, 3])
    z, cnt = orc.foldin(off, words, [1, 2], 4, 3)
    assert not z.any() and list(cnt[:, 0]) == [10, 3]
