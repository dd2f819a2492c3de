"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact for every integer and topic output; the FP64 log-likelihood within
1e-9 relative (different summation order).  Sizes are small enough for the
single-threaded oracle to finish in seconds.
"""
import os

import numpy as np
import pytest

from conftest import make_corpus
from lda_sc20_amd import ezlda
from oracle import binding as ob

pytestmark = pytest.mark.gpu

BETA = 0.01


def _pair(off, words, V, K, seed=21, shortcut=True, z0=None, **kw):
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, shortcut=int(shortcut), **kw)
    ctx.load(ezlda.Corpus(V, off, words), z0=z0)
    orc = ob.Oracle(K, float(np.float32(50.0 / K)), BETA, seed, V, off, words, z0=z0, shortcut=shortcut)
    return ctx, orc


# ------------------------------------------------------------ primitives
def test_device_philox_matches_oracle_and_kats():
    rng = np.random.default_rng(1)
    ctr = rng.integers(0, 2**32, (500, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2**32, (500, 2), dtype=np.uint64).astype(np.uint32)
    ctr[0], key[0] = 0, 0
    ctr[1], key[1] = 0xFFFFFFFF, 0xFFFFFFFF
    got = ezlda.selftest_philox(ctr, key)
    exp = np.stack([ob.philox(c, k) for c, k in zip(ctr, key)])
    assert np.array_equal(got, exp)
    assert list(got[0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(got[1]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]


# ---------------------------------------------------------------- stages
@pytest.mark.parametrize("dense_words", [-1, 0, 10**6])
def test_stages_bit_exact(dense_words):
    off, words = make_corpus(2, 250, 300, 60, zipf=1.2)
    V, K = 300, 64
    ctx, orc = _pair(off, words, V, K, dense_words=dense_words, dense_threshold=40)
    for sweep in range(3):
        ctx.build_wt()
        orc.build_wt()
        assert np.array_equal(ctx.get_wt_dense(), orc.get_wt_dense()), sweep
        assert np.array_equal(ctx.get_nk(), orc.get_nk()), sweep
        ctx.topk()
        orc.topk()
        assert np.array_equal(ctx.get_inv().view(np.uint32), orc.get_inv().view(np.uint32)), sweep
        assert np.array_equal(ctx.get_top(), orc.get_top()), sweep
        ctx.build_dt()
        orc.build_dt()
        ge, gn = ctx.get_dt_raw()
        oe, on = orc.get_dt_raw()
        assert np.array_equal(gn, on), sweep
        mask = np.zeros(len(ge), bool)
        doff = ezlda.dt_offsets(off, K)
        for d in range(len(gn)):
            mask[doff[d]: doff[d] + gn[d]] = True
        assert np.array_equal(ge[mask], oe[mask]), sweep
        assert np.array_equal(ctx.get_aux(), orc.get_aux()), sweep
        ctx.sample()
        orc.sample()
        assert np.array_equal(ctx.get_topics(), orc.get_topics()), sweep


# ---------------------------------------------------------- whole sweeps
CASES = [
    # (K, V, D, mean_len, extra context args)
    (64, 400, 300, 60, {}),
    (100, 500, 200, 80, {"dense_threshold": 30}),
    (7, 80, 100, 40, {}),
    (3, 50, 60, 30, {}),
    (2, 50, 60, 30, {}),
    (1, 40, 50, 20, {}),
    (1024, 600, 150, 300, {"dense_threshold": 200}),
    (300, 700, 200, 120, {"token_segment": 64, "dense_words": 25}),
    (64, 400, 300, 60, {"block_dim": 64}),
    (64, 400, 300, 60, {"block_dim": 512}),
    # one DT scan per token; long runs; persistent workgroups on item queues
    (64, 400, 300, 60, {"run_cap": 1}),
    (64, 400, 300, 60, {"run_cap": 16}),
    (64, 400, 300, 60, {"run_cap": 32, "grid_dim": 37}),
    (1024, 600, 150, 300, {"dense_threshold": 200, "run_cap": 3, "grid_dim": 300, "doc_tiles": 16}),
    (100, 500, 200, 80, {"grid_dim": 1, "doc_tiles": -1}),
    (100, 500, 200, 80, {"grid_dim": 0, "block_dim": 256}),
    # large K: the per-block LDS of the sampler, the histograms and the
    # log-likelihood pass 64 KiB (the topic count of the large-K corpus, and
    # the largest K the 16-bit DT entries allow)
    (10000, 300, 120, 200, {"dense_threshold": 150}),
    (16383, 150, 40, 120, {"dense_threshold": 100}),
]


@pytest.mark.parametrize("K,V,D,L,kw", CASES)
def test_sweeps_bit_exact(K, V, D, L, kw):
    off, words = make_corpus(K + V, D, V, L, zipf=1.25)
    ctx, orc = _pair(off, words, V, K, **kw)
    for s in range(3):
        ctx.sweep(1)
        orc.sweep(1)
        assert np.array_equal(ctx.get_topics(), orc.get_topics()), (K, s)
    assert ctx.loglik() == pytest.approx(orc.loglik(), rel=1e-9)


def test_shortcut_and_plain_identical_on_device():
    off, words = make_corpus(3, 400, 500, 100, zipf=1.2)
    a, _ = _pair(off, words, 500, 128, shortcut=True)
    b, _ = _pair(off, words, 500, 128, shortcut=False)
    for _ in range(4):
        a.sweep(1)
        b.sweep(1)
        assert np.array_equal(a.get_topics(), b.get_topics())


def test_head_tail_split_does_not_change_topics():
    off, words = make_corpus(4, 300, 400, 70)
    zs = []
    for dw in (0, 17, 400):
        c, _ = _pair(off, words, 400, 96, dense_words=dw)
        c.sweep(3)
        zs.append(c.get_topics())
    assert np.array_equal(zs[0], zs[1]) and np.array_equal(zs[0], zs[2])


@pytest.mark.parametrize("kw", [{}, {"grid_dim": 40, "doc_tiles": 8}])
def test_graph_replay_matches_eager(kw):
    off, words = make_corpus(5, 300, 400, 70)
    a, orc = _pair(off, words, 400, 64, **kw)
    b, _ = _pair(off, words, 400, 64, **kw)
    for _ in range(3):
        a.sweep_graph(1)
        b.sweep(1)
        orc.sweep(1)
        za = a.get_topics()  # eager host work between replays
        assert np.array_equal(za, b.get_topics())
        assert np.array_equal(za, orc.get_topics())
    a.sweep_graph(2)
    orc.sweep(2)
    assert np.array_equal(a.get_topics(), orc.get_topics())


def test_empty_docs_and_unused_words():
    off, words = make_corpus(6, 200, 300, 40, empty_docs=15)
    V = 350  # words 300..349 never occur
    ctx, orc = _pair(off, words, V, 48)
    ctx.sweep(3)
    orc.sweep(3)
    assert np.array_equal(ctx.get_topics(), orc.get_topics())
    assert ctx.loglik() == pytest.approx(orc.loglik(), rel=1e-9)


def test_given_initial_topics_and_resume(tmp_path):
    off, words = make_corpus(7, 200, 300, 50)
    K = 40
    z0 = np.random.default_rng(3).integers(0, K, len(words)).astype(np.uint16)
    a, orc = _pair(off, words, 300, K, z0=z0)
    a.sweep(2)
    path = str(tmp_path / "ck.bin")
    a.save_checkpoint(path)
    a.sweep(2)
    orc.sweep(4)
    assert np.array_equal(a.get_topics(), orc.get_topics())
    b, _ = _pair(off, words, 300, K)
    b.load_checkpoint(path)
    assert b.info().sweep == 2
    b.sweep(2)
    assert np.array_equal(b.get_topics(), a.get_topics())


def test_chunked_load_matches_oracle(tmp_path):
    off, words = make_corpus(8, 150, 200, 50)
    K = 32
    ezlda.Corpus(200, off, words).write_chunked(str(tmp_path / "c"), K=K, num_chunks=3, seed=4)
    back = ezlda.read_chunked(str(tmp_path / "c"))
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=9)
    ctx.load_chunked(str(tmp_path / "c"))
    orc = ob.Oracle(K, float(np.float32(50.0 / K)), BETA, 9, back.V, back.doc_off, back.words, z0=back.topics)
    assert np.array_equal(ctx.get_topics(), back.topics)
    ctx.sweep(2)
    orc.sweep(2)
    assert np.array_equal(ctx.get_topics(), orc.get_topics())


def test_reference_small_config_parity():
    """The reference's CPU-sized case: V=10k, D=10k, Poisson(100), K=64."""
    c = ezlda.gen_corpus(V=10000, D=10000, K_true=64, len_dist=0, len_mean=100, zipf_s=0.0, phi_conc=0.01,
                         theta_conc=0.1, seed=1001)
    K = 64
    ctx, orc = _pair(c.doc_off, c.words, c.V, K, seed=2001)
    for s in range(3):
        ctx.sweep(1)
        orc.sweep(1)
        assert np.array_equal(ctx.get_topics(), orc.get_topics()), s
    assert ctx.loglik() == pytest.approx(orc.loglik(), rel=1e-9)


def test_bad_inputs_rejected():
    ctx = ezlda.Context(K=16)
    off = np.array([0, 3], np.int64)
    with pytest.raises(ezlda.EzldaError, match="word id"):
        ctx.load(ezlda.Corpus(5, off, np.array([0, 7, 1], np.int32)))
    with pytest.raises(ezlda.EzldaError, match="initial topic"):
        ctx.load(ezlda.Corpus(5, off, np.array([0, 1, 1], np.int32)), z0=np.array([0, 16, 1], np.uint16))
    with pytest.raises(ezlda.EzldaError, match="65535"):
        ctx.load(ezlda.Corpus(5, np.array([0, 70000], np.int64), np.zeros(70000, np.int32)))
    with pytest.raises(ezlda.EzldaError, match="no corpus"):
        ctx.sweep(1)


# ----------------------------------------------------------- layout tuning
def test_set_layout_between_sweeps_keeps_topics():
    off, words = make_corpus(6, 300, 400, 70, zipf=1.2)
    ctx, orc = _pair(off, words, 400, 64)
    ctx.sweep(2)
    ctx.set_layout(run_cap=1, dense_words=0, grid_dim=50)
    ctx.sweep(2)
    orc.sweep(4)
    assert np.array_equal(ctx.get_topics(), orc.get_topics())
    ctx.set_layout(dense_words=10**6, doc_tiles=-1, block_dim=512, run_cap=32, grid_dim=0, token_segment=100)
    ctx.sweep_graph(2)  # the layout change dropped the captured sweep
    orc.sweep(2)
    assert np.array_equal(ctx.get_topics(), orc.get_topics())
    assert ctx.info().num_head_words == 400
    with pytest.raises(ezlda.EzldaError):
        ctx.set_layout(run_cap=33)


def test_autotune_sweeps_are_ordinary_sweeps():
    off, words = make_corpus(7, 400, 500, 80, zipf=1.2)
    ctx, orc = _pair(off, words, 500, 48)
    best, ms, trials = ctx.autotune(sweeps_per_trial=1)
    assert ms > 0 and len(trials) >= 10
    assert min(t["ms_per_sweep"] for t in trials) == pytest.approx(ms)
    n = ctx.info().sweep
    assert n == 2 * len(trials)
    orc.sweep(n)
    assert np.array_equal(ctx.get_topics(), orc.get_topics())
    ctx.sweep(1)
    orc.sweep(1)
    assert np.array_equal(ctx.get_topics(), orc.get_topics())


def test_export_model(tmp_path):
    off, words = make_corpus(8, 120, 200, 90, zipf=1.2)
    ctx, orc = _pair(off, words, 200, 40)
    ctx.sweep(3)
    orc.sweep(3)
    ctx.export_model(str(tmp_path))
    K, V, D = (int(x) for x in open(tmp_path / "model.txt").read().split()[:3])
    assert (K, V, D) == (40, 200, 120)
    phi = np.fromfile(tmp_path / "phi.f32", dtype=np.float32).reshape(V, K)
    orc.build_wt()
    orc.topk()
    wt, inv = orc.get_wt_dense(), orc.get_inv()
    assert np.array_equal(phi, ((wt.astype(np.float32) + np.float32(BETA)) * inv).astype(np.float32))
    orc.build_dt()
    dt = orc.get_dt_dense()
    for d, line in enumerate(open(tmp_path / "theta.txt").read().splitlines()):
        f = line.split()
        assert int(f[0]) == off[d + 1] - off[d]
        row = np.zeros(K, np.int32)
        for kc in f[1:]:
            k, c = kc.split(":")
            row[int(k)] = int(c)
        assert np.array_equal(row, dt[d])


@pytest.mark.parametrize("K", [1024, 2048, 4096, 8192])
def test_compact_rows_with_pieces(K):
    """Counts far above a compact entry's piece (63/31/15/7): rows, topics
    and log-likelihood bit-exact with the oracle through several sweeps."""
    off, words = make_corpus(9, 40, 300, 500, zipf=1.2)
    z0 = (np.arange(len(words)) % 3 == 0).astype(np.uint16) * 5  # topics 0 and 5 only
    ctx, orc = _pair(off, words, 300, K, z0=z0)
    for s in range(3):
        ctx.build_wt()
        orc.build_wt()
        ctx.topk()
        orc.topk()
        ctx.build_dt()
        orc.build_dt()
        ge, gn = ctx.get_dt_raw()
        oe, on = orc.get_dt_raw()
        assert np.array_equal(gn, on), s
        doff = ezlda.dt_offsets(off, K)
        for d in range(len(gn)):
            assert np.array_equal(ge[doff[d]: doff[d] + gn[d]], oe[doff[d]: doff[d] + on[d]]), (s, d)
        if s == 0:
            assert (ge >> 16).max() == ezlda.dt_piece(K)  # pieces are present
        ctx.sample()
        orc.sample()
        assert np.array_equal(ctx.get_topics(), orc.get_topics()), s
    assert ctx.loglik() == pytest.approx(orc.loglik(), rel=1e-9)


# ------------------------------------------- launcher-granularity stages
@pytest.mark.parametrize("order,kw", [
    (("shortcut", "sample_head", "sample_tail"), {}),
    (("sample_tail", "sample_head"), {}),
    (("sample_head", "shortcut", "sample_tail"), {"grid_dim": 0}),
    (("shortcut", "sample_tail", "sample_head"), {"shortcut": False}),
    (("shortcut", "sample_head", "sample_tail"), {"K": 10000}),
])
def test_launcher_stages_equal_fused_sampler(order, kw):
    """ezlda_shortcut / _sample_head / _sample_tail (the reference's
    UpdateProbKernelD, SampleKernelD, SampleKernel) in any order, then
    ezlda_commit_head, give the fused sampler's topics; commit_head leaves
    the head rows of the new topics and ends the sweep; wt_rowsum gives n_k."""
    kw = dict(kw)
    K = kw.pop("K", 64)
    shortcut = kw.pop("shortcut", True)
    V = 400
    off, words = make_corpus(12, 300, V, 80, zipf=1.2)
    ctx, orc = _pair(off, words, V, K, shortcut=shortcut, dense_threshold=30, **kw)
    VD = ctx.info().num_head_words
    assert 0 < VD < V
    cnt = np.bincount(words, minlength=V)
    head = np.lexsort((np.arange(V), -cnt))[:VD]  # ranks 0..VD-1: count desc, id asc
    for s in range(3):
        ctx.build_wt()
        ctx.topk()
        ctx.build_dt()
        for stage in order:
            getattr(ctx, stage)()
        ctx.commit_head()
        orc.sweep(1)
        assert ctx.info().sweep == s + 1
        assert np.array_equal(ctx.get_topics(), orc.get_topics()), (order, s)
        orc.build_wt()
        assert np.array_equal(ctx.get_wt_dense()[head], orc.get_wt_dense()[head]), (order, s)
    ctx.wt_rowsum()
    assert np.array_equal(ctx.get_nk(), orc.get_nk())


@pytest.mark.parametrize("K", [8192, 10000, 16383])
def test_dt_rows_and_aux_of_sorted_short_docs(K):
    """At large K the DT rows of documents of at most 512 tokens are built by
    sorting their topics, longer ones by a K-sized histogram: both give the
    oracle's rows (with count pieces at K=8192) and per-token counts."""
    rng = np.random.default_rng(K)
    D, V = 120, 500
    lens = rng.integers(1, 1200, D)
    lens[:6] = [1, 2, 63, 511, 512, 513]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    words = ((rng.zipf(1.2, int(off[-1])) - 1) % V).astype(np.int32)
    # few topics per doc, so counts exceed a compact entry's piece
    z0 = (rng.integers(0, 6, len(words)) * (K // 7)).astype(np.uint16)
    ctx, orc = _pair(off, words, V, K, dense_threshold=50, z0=z0)
    doff = ezlda.dt_offsets(off, K)
    for s in range(2):
        for stage in ("build_wt", "topk", "build_dt"):
            getattr(ctx, stage)()
            getattr(orc, stage)()
        ge, gn = ctx.get_dt_raw()
        oe, on = orc.get_dt_raw()
        assert np.array_equal(gn, on), (K, s)
        for d in range(D):
            assert np.array_equal(ge[doff[d]: doff[d] + gn[d]], oe[doff[d]: doff[d] + on[d]]), (K, s, d)
        assert np.array_equal(ctx.get_aux(), orc.get_aux()), (K, s)
        ctx.sample()
        orc.sample()
        assert np.array_equal(ctx.get_topics(), orc.get_topics()), (K, s)
