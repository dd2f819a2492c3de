"""GPU parity on the benchmark workloads themselves, through the C ABI.

Every configuration the benchmark quotes runs here on its own generator and
launch shape (full size where the OpenMP oracle finishes in about a minute,
else a prefix of the same generator's documents), bit-exact in topics and
n_k against the oracle, the FP64 log-likelihood within 1e-6 relative:

* the reference's CPU-sized corpus (V=10k, D=10k, Poisson(100), K=64) for
  all of its 100 sweeps;
* the ~99M-token corpus at K=1024 at full size with the library's default
  launch shape (persistent per-XCD queues, doc tiles, padding items, runs of
  4), two sweeps;
* the 738M-token corpus's first 200,000 documents with every word dense, at
  K=4096 (compact rows, pieces of 15) and K=10000 (32-bit rows);
* the heavy-tailed corpus's first 20,000 documents (Pareto lengths up to
  20,000) at K=2048, under the tuned layout and the plain launch path;
* fold-in serving against a K=1024 model over the full 102,660-word
  vocabulary, 64 log-normal requests in [10, 5000] through 2,048 slots,
  alone, in one mixed batch and on the plain path;
* long documents whose rows pass the sampler's checkpointed blocks (more
  than 384 compact entries at K=1024; more than 512 slots with 32-bit rows).

The oracle here is the OpenMP build (bit-identical to the single-threaded
one; tests/test_oracle.py pins that)."""
import math

import numpy as np
import pytest

from bench import SERVE_GEN, WORKLOADS
from lda_sc20_amd import ezlda
from oracle import binding as ob

pytestmark = pytest.mark.gpu

BETA = 0.01
LL_REL = 1e-6


def _gen(name, **over):
    g = dict(WORKLOADS[name]["gen"])
    g.update(over)
    return ezlda.gen_corpus(**g)


def _dense(name, V):
    f = WORKLOADS[name]["head_frac"]
    return -1 if f is None else int(math.ceil(f * V))


def _alpha(K):
    return float(np.float32(50.0 / K))


def _check(ctx, orc, what):
    assert np.array_equal(ctx.get_topics(), orc.get_topics()), what
    ctx.build_wt()
    orc.build_wt()
    assert np.array_equal(ctx.get_nk(), orc.get_nk()), what


def test_cfg1_all_100_sweeps():
    wl = WORKLOADS["cfg1"]
    c = _gen("cfg1")
    K, seed = wl["K"], wl["seed"]
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed)
    ctx.load(c)
    orc = ob.Oracle(K, _alpha(K), BETA, seed, c.V, c.doc_off, c.words, omp=True)
    done = 0
    for step in (1, 9, 40, 50):
        ctx.sweep_graph(step)
        orc.sweep(step)
        done += step
        _check(ctx, orc, ("cfg1", done))
    assert done == 100 and ctx.info().sweep == 100
    assert ctx.loglik() == pytest.approx(orc.loglik(), rel=LL_REL)


def test_cfg2_full_size_default_launch_shape():
    wl = WORKLOADS["cfg2"]
    c = _gen("cfg2")
    K, seed = wl["K"], wl["seed"]
    assert c.N > 98_000_000 and c.D == 300_000
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, dense_words=_dense("cfg2", c.V), shortcut=1)
    ctx.load(c)
    i = ctx.info()
    # the library's default layout: persistent one-wavefront workgroups on
    # per-XCD queues, doc tiles with padding items, runs of at most 4 tokens
    assert i.grid_dim > 0 and i.block_dim == 64 and i.doc_tiles > 0 and i.run_cap == 4
    assert i.num_head_words == 1027
    orc = ob.Oracle(K, _alpha(K), BETA, seed, c.V, c.doc_off, c.words, omp=True)
    for s in range(2):
        ctx.sweep(1)
        orc.sweep(1)
        _check(ctx, orc, ("cfg2", s))
        if s == 0:
            ctx.build_dt()
            _, nnz = ctx.get_dt_raw()
            assert nnz.max() > 384  # rows past the sampler's checkpointed blocks
    assert ctx.loglik() == pytest.approx(orc.loglik(), rel=LL_REL)


@pytest.mark.parametrize("K", [4096, 10000])
def test_cfg3_prefix_every_word_dense(K):
    wl = WORKLOADS["cfg3"]
    c = _gen("cfg3", D=200_000)
    seed = wl["seed"]
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, dense_words=_dense("cfg3", c.V))
    ctx.load(c)
    assert ctx.info().num_head_words == c.V
    orc = ob.Oracle(K, _alpha(K), BETA, seed, c.V, c.doc_off, c.words, omp=True)
    for s in range(2):
        ctx.sweep(1)
        orc.sweep(1)
        _check(ctx, orc, ("cfg3", K, s))
    assert ctx.loglik() == pytest.approx(orc.loglik(), rel=LL_REL)


def test_cfg4_prefix_tuned_and_plain_launch():
    wl = WORKLOADS["cfg4"]
    c = _gen("cfg4", D=20_000)
    K, seed = wl["K"], wl["seed"]
    lens = np.diff(c.doc_off)
    assert lens.max() > 10_000  # the heavy tail is in the prefix
    # tuned: the autotuner's choice on this workload (head of 4% of the
    # vocabulary, runs of 8); plain: one 256-thread workgroup per item, one
    # row scan per token, no doc tiles (the reference's launch shape)
    tuned = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, dense_words=int(math.ceil(0.04 * c.V)),
                          run_cap=8)
    plain = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, dense_words=_dense("cfg4", c.V),
                          grid_dim=0, block_dim=256, run_cap=1, doc_tiles=-1)
    tuned.load(c)
    plain.load(c)
    orc = ob.Oracle(K, _alpha(K), BETA, seed, c.V, c.doc_off, c.words, omp=True)
    for s in range(2):
        tuned.sweep_graph(1)
        plain.sweep(1)
        orc.sweep(1)
        _check(tuned, orc, ("cfg4 tuned", s))
        assert np.array_equal(plain.get_topics(), tuned.get_topics()), ("cfg4 plain", s)
    assert tuned.loglik() == pytest.approx(orc.loglik(), rel=LL_REL)


def test_cfg5_foldin_full_vocabulary():
    wl = WORKLOADS["cfg2"]
    c = _gen("cfg2", D=20_000)  # a prefix of the training corpus, full vocabulary
    K, seed = wl["K"], wl["seed"]
    assert c.V == 102_660
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, dense_words=_dense("cfg2", c.V))
    ctx.load(c)
    orc = ob.Oracle(K, _alpha(K), BETA, seed, c.V, c.doc_off, c.words, omp=True)
    ctx.sweep(3)
    orc.sweep(3)
    _check(ctx, orc, "cfg5 model")
    # requests: further documents of the same generative model
    g = dict(wl["gen"])
    g.pop("D")
    g.update(SERVE_GEN)
    req = ezlda.gen_corpus(D=64, first_doc=300_000, **g)
    lens = np.diff(req.doc_off)
    assert lens.min() >= 10 and lens.max() <= 5000
    rids = np.arange(70_000, 70_064, dtype=np.uint32)
    sweeps, sseed = 20, wl["seed"] + 7
    zo, co = orc.foldin(req.doc_off, req.words, rids, sweeps, sseed)
    srv = ezlda.Server(ctx, max_slots=2048, sweeps=sweeps, hbm_budget_bytes=1 << 30, seed=sseed)
    docs = [req.words[req.doc_off[r]: req.doc_off[r + 1]] for r in range(64)]
    # alone: one request in flight at a time
    for r in (0, 17, 63):
        srv.submit(int(rids[r]), docs[r])
        z, cnt, _ = srv.wait(int(rids[r]), len(docs[r]), timeout_ms=60000)
        assert np.array_equal(z, zo[req.doc_off[r]: req.doc_off[r + 1]]), ("alone", r)
        assert np.array_equal(cnt, co[r]), ("alone", r)
    # in the mix: all 64 in one batch
    srv.submit_many(rids, req.doc_off, req.words)
    for r in range(64):
        z, cnt, _ = srv.wait(int(rids[r]), len(docs[r]), timeout_ms=60000)
        assert np.array_equal(z, zo[req.doc_off[r]: req.doc_off[r + 1]]), ("mix", r)
        assert np.array_equal(cnt, co[r]), ("mix", r)
    # the plain one-document path
    for r in (5, 40):
        z, cnt = srv.plain(int(rids[r]), docs[r])
        assert np.array_equal(z, zo[req.doc_off[r]: req.doc_off[r + 1]]) and np.array_equal(cnt, co[r])
    srv.close()


@pytest.mark.parametrize("K,lo,hi", [(1024, 1500, 5000), (10000, 600, 3000)])
def test_long_documents_pass_checkpointed_blocks(K, lo, hi):
    rng = np.random.default_rng(K)
    D, V = 60, 3000
    lens = rng.integers(lo, hi, D)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    words = ((rng.zipf(1.2, int(off[-1])) - 1) % V).astype(np.int32)
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=77)
    ctx.load(ezlda.Corpus(V, off, words))
    orc = ob.Oracle(K, _alpha(K), BETA, 77, V, off, words, omp=True)
    for s in range(3):
        ctx.build_wt()
        ctx.topk()
        ctx.build_dt()
        _, nnz = ctx.get_dt_raw()
        if s == 0:
            # compact rows: 12 checkpointed blocks of 32 entries; 32-bit rows:
            # documents of more than 512 slots in the DT rebuild
            assert nnz.max() > (384 if K <= 8192 else 512), nnz.max()
        ctx.sample()
        orc.sweep(1)
        assert np.array_equal(ctx.get_topics(), orc.get_topics()), (K, s)
    assert ctx.loglik() == pytest.approx(orc.loglik(), rel=LL_REL)
