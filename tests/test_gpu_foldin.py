"""GPU fold-in serving against the oracle's fold-in, through the C ABI.

Bit-exact topics and counts per request, whatever the batch a request shares
a launch with, its arrival order, the slot count or the HBM token budget, and
on the one-document-per-launch plain path.
"""
import threading

import numpy as np
import pytest

from conftest import make_corpus
from lda_sc20_amd import ezlda
from oracle import binding as ob

pytestmark = pytest.mark.gpu

BETA = 0.01


def _model(K, V=300, D=200, L=60, sweeps=3, seed=21, **kw):
    off, words = make_corpus(4, D, V, L, zipf=1.2)
    alpha = float(np.float32(50.0 / K))
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed, **kw)
    ctx.load(ezlda.Corpus(V, off, words))
    orc = ob.Oracle(K, alpha, BETA, seed, V, off, words)
    ctx.sweep(sweeps)
    orc.sweep(sweeps)
    assert np.array_equal(ctx.get_topics(), orc.get_topics())
    return ctx, orc


def _docs(V, lens, seed=5):
    rng = np.random.default_rng(seed)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    words = ((rng.zipf(1.2, int(off[-1])) - 1) % V).astype(np.int32)
    return off, words


LENS = [0, 1, 2, 63, 64, 65, 129, 500, 1200, 37, 5, 300]


@pytest.mark.parametrize("K,kw", [(64, {}), (100, {"dense_threshold": 30}), (7, {}), (1024, {}),
                                  (10000, {"dense_threshold": 100})])
def test_server_matches_oracle(K, kw):
    ctx, orc = _model(K, **kw)
    off, words = _docs(300, LENS)
    rids = np.arange(1000, 1000 + len(LENS), dtype=np.uint32)
    sweeps, seed = 12, 99
    zo, co = orc.foldin(off, words, rids, sweeps, seed)
    srv = ezlda.Server(ctx, max_slots=8, sweeps=sweeps, hbm_budget_bytes=1 << 20, seed=seed)
    for r, rid in enumerate(rids):
        srv.submit(int(rid), words[off[r]: off[r + 1]])
    for r, rid in enumerate(rids):
        z, c, lat = srv.wait(int(rid), int(off[r + 1] - off[r]), timeout_ms=60000)
        assert np.array_equal(z, zo[off[r]: off[r + 1]]), (K, r)
        assert np.array_equal(c, co[r]), (K, r)
        assert lat > 0
    st = srv.stats()
    assert st.completed == st.submitted == len(LENS) and st.rejected == 0
    assert st.active == 0 and st.pending == 0 and 1 <= st.max_active <= 8
    assert 0 < st.p50_ms <= st.p99_ms <= st.max_ms
    # plain path: one document per launch, same results
    for r in (0, 3, 7):
        z, c = srv.plain(int(rids[r]), words[off[r]: off[r + 1]])
        assert np.array_equal(z, zo[off[r]: off[r + 1]]) and np.array_equal(c, co[r])
    srv.close()


def test_alone_equals_in_mix_and_any_arrival_order():
    ctx, orc = _model(64)
    rng = np.random.default_rng(3)
    lens = rng.integers(1, 400, 60)
    off, words = _docs(300, lens, seed=9)
    rids = rng.permutation(1 << 20)[:60].astype(np.uint32)
    zo, co = orc.foldin(off, words, rids, 20, 5)
    srv = ezlda.Server(ctx, max_slots=16, sweeps=20, hbm_budget_bytes=ezlda.FOLDIN_TOKEN_BYTES * 3000, seed=5)
    # alone
    srv.submit(int(rids[7]), words[off[7]: off[8]])
    z, c, _ = srv.wait(int(rids[7]), int(lens[7]), timeout_ms=30000)
    assert np.array_equal(z, zo[off[7]: off[8]]) and np.array_equal(c, co[7])
    # in a mix, submitted from two threads in shuffled orders, with a token
    # budget that forces admission to wait for retirements
    order = rng.permutation(60)

    def feed(part):
        for r in part:
            srv.submit(int(rids[r]), words[off[r]: off[r + 1]])

    th = [threading.Thread(target=feed, args=(order[i::2],)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.drain()
    for r in range(60):
        z, c, _ = srv.wait(int(rids[r]), int(lens[r]), timeout_ms=1000)
        assert np.array_equal(z, zo[off[r]: off[r + 1]]), r
        assert np.array_equal(c, co[r]), r
    st = srv.stats()
    assert st.completed == 61 and st.steps >= 20
    srv.close()


def test_budget_rejects_oversized_and_bad_requests():
    ctx, _ = _model(32)
    srv = ezlda.Server(ctx, max_slots=2, sweeps=3, hbm_budget_bytes=ezlda.FOLDIN_TOKEN_BYTES * 100, seed=1)
    with pytest.raises(ezlda.EzldaError) as e:
        srv.submit(1, np.zeros(101, np.int32))
    assert e.value.code == ezlda.ENOMEM
    with pytest.raises(ezlda.EzldaError):
        srv.submit(2, np.array([0, 300], np.int32))  # word id out of range (V = 300)
    with pytest.raises(ezlda.EzldaError):
        srv.submit(3, np.zeros(70000, np.int32))
    srv.submit(4, np.zeros(100, np.int32))  # exactly the budget
    with pytest.raises(ezlda.EzldaError):
        srv.submit(4, np.zeros(3, np.int32))  # the same id while in flight
    z, c, _ = srv.wait(4, 100, timeout_ms=30000)
    assert c.sum() == 100 and len(z) == 100
    assert srv.stats().rejected == 1
    with pytest.raises(ezlda.EzldaError):
        srv.wait(4, 100, timeout_ms=1000)  # already claimed
    with pytest.raises(ezlda.EzldaError):
        srv.wait(12345, 1, timeout_ms=-1)  # never submitted: no wait
    srv.close()


def test_server_survives_model_training_afterwards():
    """The server froze phi: training the context further changes nothing."""
    ctx, orc = _model(48)
    off, words = _docs(300, [80, 20])
    zo, co = orc.foldin(off, words, [1, 2], 8, 4)
    srv = ezlda.Server(ctx, max_slots=4, sweeps=8, seed=4)
    ctx.sweep(2)
    for r in range(2):
        srv.submit(r + 1, words[off[r]: off[r + 1]])
    for r in range(2):
        z, c, _ = srv.wait(r + 1, int(off[r + 1] - off[r]), timeout_ms=30000)
        assert np.array_equal(z, zo[off[r]: off[r + 1]]) and np.array_equal(c, co[r])
    srv.close()


def test_submit_many_all_or_none():
    ctx, orc = _model(64)
    lens = np.array([5, 0, 300, 64, 1, 77])
    off, words = _docs(300, lens, seed=12)
    rids = np.array([9, 3, 77, 1000, 5, 6], np.uint32)
    zo, co = orc.foldin(off, words, rids, 6, 31)
    srv = ezlda.Server(ctx, max_slots=4, sweeps=6, hbm_budget_bytes=1 << 20, seed=31)
    bad = words.copy()
    bad[off[3] + 2] = 300  # word id out of range in the fourth document
    with pytest.raises(ezlda.EzldaError):
        srv.submit_many(rids, off, bad)
    with pytest.raises(ezlda.EzldaError):  # nothing was queued
        srv.wait(int(rids[0]), int(lens[0]), timeout_ms=1000)
    with pytest.raises(ezlda.EzldaError):  # an id twice in one batch
        srv.submit_many(np.array([1, 2, 1], np.uint32), off[:4], words)
    srv.submit_many(rids, off, words)
    srv.drain()
    for d in range(len(lens)):
        z, c, _ = srv.wait(int(rids[d]), int(lens[d]), timeout_ms=30000)
        assert np.array_equal(z, zo[off[d]: off[d + 1]]) and np.array_equal(c, co[d]), d
    assert srv.stats().completed == len(lens)
    srv.close()


@pytest.mark.parametrize("K,lens", [(16383, [3000, 1, 0, 700]), (1024, [20000, 5000, 64, 2])])
def test_plain_one_launch_per_document(K, lens):
    """The plain path runs all of a document's sweeps in one launch (its row
    in LDS): bit-exact with the oracle at the largest K (the largest LDS row)
    and for documents far longer than the workgroup."""
    ctx, orc = _model(K, dense_threshold=100)
    off, words = _docs(300, lens, seed=K)
    rids = np.arange(40, 40 + len(lens), dtype=np.uint32)
    zo, co = orc.foldin(off, words, rids, 7, 13)
    srv = ezlda.Server(ctx, max_slots=4, sweeps=7, hbm_budget_bytes=1 << 20, seed=13)
    for r in range(len(lens)):
        z, c = srv.plain(int(rids[r]), words[off[r]: off[r + 1]])
        assert np.array_equal(z, zo[off[r]: off[r + 1]]), (K, r)
        assert np.array_equal(c, co[r]), (K, r)
    srv.close()


def test_wait_buffer_comes_from_the_submitted_length():
    """The library copies as many topics as the request was submitted with;
    Server.wait sizes its buffer from that, whatever n the caller passes."""
    ctx, orc = _model(32)
    off, words = _docs(300, [150])
    zo, co = orc.foldin(off, words, [5], 4, 2)
    srv = ezlda.Server(ctx, max_slots=2, sweeps=4, seed=2)
    srv.submit(5, words)
    z, c, _ = srv.wait(5, 1, timeout_ms=30000)
    assert len(z) == 150 and np.array_equal(z, zo) and np.array_equal(c, co[0])
    srv.close()
