"""The out-of-core path (num_chunks > 1): documents cut into chunks whose
token state and layout live in pinned host memory and pass through two device
buffers per sweep.  Same Jacobi semantics as the resident corpus, so the
topics, counts and draws must be identical (bit-exact), and the FP64
log-likelihood equal within 1e-9 relative (summed per chunk)."""
import os

import numpy as np
import pytest

from conftest import make_corpus
from lda_sc20_amd import ezlda
from oracle import binding as ob

pytestmark = pytest.mark.gpu

BETA = 0.01


def _ctx(off, words, V, K, seed=21, **kw):
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, **kw)
    ctx.load(ezlda.Corpus(V, off, words))
    return ctx


@pytest.mark.parametrize("K,chunks,kw", [(64, 3, {}), (64, 7, {"run_cap": 1}), (1024, 2, {}),
                                         (300, 4, {"grid_dim": 0}), (10000, 3, {})])
def test_chunked_sweeps_equal_resident_and_oracle(K, chunks, kw):
    off, words = make_corpus(11, 260, 500, 70, zipf=1.15)
    V = 500
    res = _ctx(off, words, V, K, **kw)
    chk = _ctx(off, words, V, K, num_chunks=chunks, **kw)
    orc = ob.Oracle(K, float(np.float32(50.0 / K)), BETA, 21, V, off, words)
    assert chk.info().num_chunks == min(chunks, 260) and res.info().num_chunks == 1
    assert np.array_equal(chk.get_topics(), res.get_topics())
    for step in (1, 2, 3):
        res.sweep(step)
        chk.sweep(step)
        orc.sweep(step)
        zc = chk.get_topics()
        assert np.array_equal(zc, res.get_topics()), step
        assert np.array_equal(zc, orc.get_topics()), step
    assert chk.info().sweep == res.info().sweep == 6
    # the sweep's model (built from the topics before the last sweep)
    assert np.array_equal(chk.get_inv().view(np.uint32), res.get_inv().view(np.uint32))
    assert np.array_equal(chk.get_top(), res.get_top())
    # chunked sweeps add up the new topics' counts as they go; the resident
    # context rebuilds them at the next sweep's start (or here)
    res.build_wt()
    assert np.array_equal(chk.get_nk(), res.get_nk())
    assert np.array_equal(chk.get_wt_dense(), res.get_wt_dense())
    lc, lr = chk.loglik(), res.loglik()
    assert abs(lc - lr) <= 1e-9 * abs(lr)
    assert abs(lc - orc.loglik()) <= 1e-9 * abs(lr)
    sc, sr = chk.stats(), res.stats()
    assert sc[0] == sr[0] and sc[1] == sr[1]  # full draws and DT entries scanned


def test_chunked_topics_checkpoint_and_refusals(tmp_path):
    off, words = make_corpus(12, 120, 300, 50)
    V, K = 300, 64
    chk = _ctx(off, words, V, K, num_chunks=4)
    res = _ctx(off, words, V, K)
    rng = np.random.default_rng(3)
    z = rng.integers(0, K, len(words)).astype(np.uint16)
    chk.set_topics(z)
    res.set_topics(z)
    assert np.array_equal(chk.get_topics(), z)
    chk.sweep(2)
    res.sweep(2)
    assert np.array_equal(chk.get_topics(), res.get_topics())
    for ctx, name in ((chk, "chunked"), (res, "resident")):
        os.makedirs(tmp_path / name)
        ctx.export_model(str(tmp_path / name))
    for f in ("phi.f32", "theta.txt", "model.txt"):
        assert open(tmp_path / "chunked" / f, "rb").read() == open(tmp_path / "resident" / f, "rb").read(), f
    path = str(tmp_path / "ck.bin")
    chk.save_checkpoint(path)
    chk.sweep(3)
    z5 = chk.get_topics()
    chk.load_checkpoint(path)
    assert chk.info().sweep == 2
    chk.sweep(3)
    assert np.array_equal(chk.get_topics(), z5)
    total, stages = chk.time_sweeps(2)
    assert total > 0 and all(v == -1.0 for v in stages.values())
    for call in (lambda: chk.sweep_graph(1), chk.get_dt_raw, chk.build_wt, chk.sample,
                 lambda: chk.time_sweeps(1, use_graph=True)):
        with pytest.raises(ezlda.EzldaError) as e:
            call()
        assert e.value.code == ezlda.ESTATE


def test_more_chunks_than_documents():
    off = np.array([0, 5, 9, 30, 31, 60], np.int64)
    words = (np.arange(60) * 7 % 40).astype(np.int32)
    res = _ctx(off, words, 40, 16)
    chk = _ctx(off, words, 40, 16, num_chunks=8)
    assert chk.info().num_chunks == 5
    res.sweep(4)
    chk.sweep(4)
    assert np.array_equal(chk.get_topics(), res.get_topics())


def test_foldin_server_on_a_chunked_model():
    off, words = make_corpus(13, 150, 300, 60)
    V, K = 300, 64
    res = _ctx(off, words, V, K)
    chk = _ctx(off, words, V, K, num_chunks=3)
    res.sweep(3)
    chk.sweep(3)
    qoff, qwords = make_corpus(14, 6, V, 40)
    outs = []
    for ctx in (res, chk):
        srv = ezlda.Server(ctx, max_slots=4, sweeps=5, hbm_budget_bytes=1 << 20, seed=9)
        got = []
        for d in range(6):
            srv.submit(d, qwords[qoff[d]: qoff[d + 1]])
        for d in range(6):
            z, c, _ = srv.wait(d, int(qoff[d + 1] - qoff[d]), timeout_ms=30000)
            got.append((z, c))
        srv.close()
        outs.append(got)
    for (za, ca), (zb, cb) in zip(*outs):
        assert np.array_equal(za, zb) and np.array_equal(ca, cb)
