"""Pins the CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, so each function of
the oracle is pinned here by known-answer vectors (Philox), an independent
implementation (rocRAND's Philox), brute force in numpy / FP64, closed forms
and invariants.  The device reproduces these results bit for bit (tests/test_gpu_parity.py).
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, make_corpus
from lda_sc20_amd import ezlda
from oracle import binding as ob

ALPHA = lambda K: 50.0 / K  # noqa: E731
BETA = 0.01


# ------------------------------------------------------------------ Philox
# Known-answer vectors of Philox4x32-10 from the Random123 distribution
# (kat_vectors, "philox4x32 10" lines).
KATS = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KATS)
def test_philox_kat(ctr, key, out):
    assert list(ob.philox(ctr, key)) == out


ROCRAND_CHECK = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>
struct Probe : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};
int main() {
    uint32_t c[4], k[2];
    Probe e;
    while (std::scanf("%u %u %u %u %u %u", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
        uint4 o = e.ten_rounds(uint4{c[0], c[1], c[2], c[3]}, uint2{k[0], k[1]});
        std::printf("%u %u %u %u\n", o.x, o.y, o.z, o.w);
    }
    return 0;
}
"""


@pytest.mark.skipif(not os.path.exists("/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"),
                    reason="rocRAND headers absent")
def test_philox_matches_rocrand_host():
    """Independent implementation: rocRAND's Philox4x32-10 rounds, run on the host."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    rng = np.random.default_rng(5)
    rows = rng.integers(0, 2**32, size=(200, 6), dtype=np.uint64)
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "chk.cpp")
        exe = os.path.join(td, "chk")
        with open(src, "w") as f:
            f.write(ROCRAND_CHECK)
        r = subprocess.run([hipcc, "-O1", "-x", "hip", "--offload-arch=gfx950", src, "-o", exe],
                           capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip("cannot build the rocRAND host check: " + r.stderr[-400:])
        inp = "\n".join(" ".join(str(int(x)) for x in row) for row in rows) + "\n"
        p = subprocess.run([exe], input=inp, capture_output=True, text=True, timeout=60)
        assert p.returncode == 0, p.stderr
    got = np.array([[int(x) for x in ln.split()] for ln in p.stdout.strip().splitlines()], dtype=np.uint64)
    for row, g in zip(rows, got):
        assert list(ob.philox(row[:4], row[4:])) == list(g)


def test_uniform_grid():
    us = np.array([ob.uniform(3, t, 0) for t in range(4000)])
    assert us.min() >= 0.0 and us.max() < 1.0
    assert np.all(us * 2**24 == np.floor(us * 2**24))
    assert abs(us.mean() - 0.5) < 0.02


# ------------------------------------------------------------- counts
def _oracle(off, words, K, seed=11, shortcut=True, omp=False, z0=None, V=None):
    V = V if V is not None else int(words.max()) + 1 if len(words) else 1
    return ob.Oracle(K, ALPHA(K), BETA, seed, V, off, words, z0=z0, shortcut=shortcut, omp=omp)


def _doc_of_token(off):
    return np.repeat(np.arange(len(off) - 1), np.diff(off))


def test_counts_bruteforce():
    off, words = make_corpus(2, 120, 90, 40)
    K = 24
    o = _oracle(off, words, K, V=90)
    o.sweep(2)
    o.build_wt()
    o.build_dt()
    z = o.get_topics().astype(np.int64)
    doc = _doc_of_token(off)
    wt = np.zeros((90, K), np.int64)
    np.add.at(wt, (words, z), 1)
    dt = np.zeros((len(off) - 1, K), np.int64)
    np.add.at(dt, (doc, z), 1)
    assert np.array_equal(o.get_wt_dense(), wt)
    assert np.array_equal(o.get_nk(), wt.sum(0))
    assert np.array_equal(o.get_dt_dense(), dt)
    _check_rows(o, off, K)


def _check_rows(o, off, K):
    """Rows hold each topic once in ascending order (compact rows: in pieces
    of 63 and a remainder), counts summing to n_d, within the capacity."""
    ent, nnz = o.get_dt_raw()
    doff = ezlda.dt_offsets(off, K)
    dense = o.get_dt_dense()
    for d in range(len(off) - 1):
        row = ent[doff[d]: doff[d] + nnz[d]]
        tk, ct = (row & 0xFFFF).astype(np.int64), (row >> 16).astype(np.int64)
        assert np.all(np.diff(tk) >= 0) and np.all(ct > 0)
        assert nnz[d] <= doff[d + 1] - doff[d]
        assert ct.sum() == off[d + 1] - off[d]
        exp = []
        for k in np.nonzero(dense[d])[0]:
            c = int(dense[d, k])
            if K <= ezlda.COMPACT_K:
                p = ezlda.dt_piece(K)
                exp += [(k, p)] * ((c - 1) // p) + [(k, c - p * ((c - 1) // p))]
            else:
                exp.append((k, c))
        assert list(zip(tk.tolist(), ct.tolist())) == exp


@pytest.mark.parametrize("K", [2, 5, 1025, 2048, 4097, 8193])
def test_dt_rows_split_large_counts(K):
    # long docs over few topics: counts far above 63 (split when K <= 1024)
    off, words = make_corpus(7, 12, 30, 600)
    o = ob.Oracle(K, float(np.float32(ALPHA(K))), BETA, 3, 30, off, words,
                  z0=(np.arange(len(words)) % 2).astype(np.uint16))
    o.build_wt()
    o.build_dt()
    assert o.get_dt_dense().max() > 130
    _check_rows(o, off, K)
    o.sweep(2)
    o.build_wt()
    o.build_dt()
    _check_rows(o, off, K)


def _phi32(wt, nk, V, K):
    vbeta = np.float32(V * BETA)  # (double)V*beta rounded once, as the oracle
    vbeta = np.float32(np.float64(V) * np.float64(np.float32(BETA)))
    inv = (np.float32(1.0) / (nk.astype(np.float32) + vbeta)).astype(np.float32)
    phi = ((wt.astype(np.float32) + np.float32(BETA)) * inv).astype(np.float32)
    return inv, phi


def test_inv_and_top3_bruteforce():
    off, words = make_corpus(3, 150, 120, 50)
    K = 40
    o = _oracle(off, words, K, V=120)
    o.sweep(1)
    o.build_wt()
    o.topk()
    wt, nk = o.get_wt_dense(), o.get_nk()
    inv, phi = _phi32(wt, nk, 120, K)
    assert np.array_equal(o.get_inv(), inv)
    top = o.get_top()
    for w in range(120):
        order = np.lexsort((np.arange(K), -phi[w].astype(np.float64)))
        assert list(top[w]) == list(order[:3]), w


def test_aux_bruteforce():
    off, words = make_corpus(4, 100, 80, 45)
    K = 20
    o = _oracle(off, words, K, V=80)
    o.sweep(1)
    o.build_wt()
    o.topk()
    o.build_dt()
    top, dt = o.get_top(), o.get_dt_dense()
    doc = _doc_of_token(off)
    aux = o.get_aux()
    exp = dt[doc, top[words, 0]] | (dt[doc, top[words, 1]] << 16)
    assert np.array_equal(aux, exp.astype(np.uint32))


# ------------------------------------------------------------- sampler
def _fp64_draw(u, k1, P1, s_terms, q_terms):
    """Inverse CDF in FP64 over [k1] + S entries + Q topics (canonical order)."""
    labels = [k1] + [k for k, _ in s_terms] + [k for k, _ in q_terms]
    mass = np.array([P1] + [m for _, m in s_terms] + [m for _, m in q_terms], np.float64)
    c = np.cumsum(mass)
    x = u * c[-1]
    i = int(np.searchsorted(c, x, side="right"))
    i = min(i, len(c) - 1)
    gap = np.min(np.abs(c - x)) / c[-1]
    return labels[i], gap


@pytest.mark.parametrize("K,shortcut,mean_len", [(7, True, 30), (64, True, 30), (64, False, 30), (100, True, 30),
                                                 (300, True, 400), (300, False, 400)])
def test_sampler_matches_fp64_inverse_cdf(K, shortcut, mean_len):
    """Every draw equals an FP64 inverse CDF over the canonical bucket order
    with the same u, except within rounding distance of a bucket boundary."""
    off, words = make_corpus(5, 80 if mean_len < 100 else 20, 60, mean_len)
    V = 60
    seed = 99
    o = _oracle(off, words, K, seed=seed, shortcut=shortcut, V=V)
    o.sweep(2)
    o.build_wt()
    o.topk()
    o.build_dt()
    wt, nk, top, dt = o.get_wt_dense(), o.get_nk(), o.get_top(), o.get_dt_dense()
    _, phi32 = _phi32(wt, nk, V, K)
    phi = phi32.astype(np.float64)
    sweep = o.sweep_index
    _, _, doc_tok, _ = o.get_layout()
    o.sample()
    znew = o.get_topics()
    doc = _doc_of_token(off)
    alpha = float(np.float32(ALPHA(K)))
    checked = skipped = 0
    for i in range(len(words)):
        w, d = words[i], doc[i]
        k1 = int(top[w, 0])
        u = ob.uniform(seed, int(doc_tok[i]), sweep)
        P1 = (dt[d, k1] + alpha) * phi[w, k1]
        s_terms = [(k, dt[d, k] * phi[w, k]) for k in range(K) if dt[d, k] > 0 and k != k1]
        q_terms = [(k, alpha * phi[w, k]) for k in range(K) if k != k1]
        exp, gap = _fp64_draw(u, k1, P1, s_terms, q_terms)
        if gap < 1e-5:
            skipped += 1
            continue
        checked += 1
        assert znew[i] == exp, (i, w, d, u)
    assert checked > 0.99 * (checked + skipped)


def test_sampler_chi2_fixed_context():
    """All tokens of one (word, doc) draw iid from p(k) ~ (n_dk+a)(n_wk+b)/(n_k+Vb)."""
    from scipy.stats import chi2

    K, V = 12, 4
    rng = np.random.default_rng(8)
    n0 = 40000
    words0 = np.where(rng.random(n0) < 0.8, 0, rng.integers(1, V, n0)).astype(np.int32)
    words1 = rng.integers(0, V, 3000).astype(np.int32)
    off = np.array([0, n0, n0 + 3000], np.int64)
    words = np.concatenate([words0, words1])
    o = _oracle(off, words, K, seed=17, V=V)
    o.sweep(3)
    o.build_wt()
    o.topk()
    o.build_dt()
    wt, nk, dt = o.get_wt_dense().astype(np.float64), o.get_nk().astype(np.float64), o.get_dt_dense()
    alpha = ALPHA(K)
    p = (dt[0] + alpha) * (wt[0] + BETA) / (nk + V * BETA)
    p /= p.sum()
    o.sample()
    z = o.get_topics()[:n0][words0 == 0]
    obs = np.bincount(z, minlength=K).astype(np.float64)
    exp = p * len(z)
    keep = exp > 5
    stat = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum() + ((obs[~keep].sum() - exp[~keep].sum()) ** 2 /
                                                              max(exp[~keep].sum(), 1e-9) if (~keep).any() else 0)
    dof = keep.sum() - 1 + (1 if (~keep).any() else 0)
    assert stat < chi2.ppf(0.9999, dof), (stat, dof)


def test_shortcut_is_exact():
    off, words = make_corpus(6, 200, 150, 70, zipf=1.2)
    for K in (16, 64, 130):
        a = _oracle(off, words, K, shortcut=True, V=150)
        b = _oracle(off, words, K, shortcut=False, V=150)
        for _ in range(4):
            a.sweep(1)
            b.sweep(1)
            assert np.array_equal(a.get_topics(), b.get_topics()), K


def test_shortcut_bounds_give_the_same_topics():
    """The GPU path's looser bound (n_d - n_dk1) phi_k2 (mode 2) and the
    reference's bound (mode 1) both accept only tokens the full draw sends to
    k1: the topics equal those without a shortcut, sweep by sweep; the looser
    bound takes at least as many full draws, and both skip some."""
    off, words = make_corpus(8, 150, 120, 80, zipf=1.3)
    N = len(words)
    for K in (8, 64, 257):
        o = [_oracle(off, words, K, shortcut=m, V=120) for m in (0, 1, 2)]
        for s in range(8):
            for x in o:
                x.sweep(1)
            z = o[0].get_topics()
            assert np.array_equal(o[1].get_topics(), z), (K, s)
            assert np.array_equal(o[2].get_topics(), z), (K, s)
        f0, f1, f2 = (x.full_draws() for x in o)
        assert f0 == 8 * N, K
        assert f1 <= f2 < f0, (K, f1, f2, f0)


def test_openmp_build_is_bit_identical():
    off, words = make_corpus(7, 300, 200, 60)
    a = _oracle(off, words, 48, omp=False, V=200)
    b = _oracle(off, words, 48, omp=True, V=200)
    a.sweep(3)
    b.sweep(3)
    assert np.array_equal(a.get_topics(), b.get_topics())
    assert a.loglik() == pytest.approx(b.loglik(), rel=1e-12)


def test_resume_from_topics_is_identical():
    off, words = make_corpus(8, 100, 90, 50)
    a = _oracle(off, words, 32, V=90)
    a.sweep(4)
    b = _oracle(off, words, 32, V=90)
    b.sweep(2)
    c = _oracle(off, words, 32, V=90, z0=b.get_topics())
    c.set_sweep(2)
    c.sweep(2)
    assert np.array_equal(a.get_topics(), c.get_topics())


def test_k1_always_topic_zero():
    off, words = make_corpus(9, 50, 40, 20)
    o = _oracle(off, words, 1, V=40)
    o.sweep(2)
    assert np.all(o.get_topics() == 0)


def test_loglik_matches_numpy_fp64():
    off, words = make_corpus(10, 90, 70, 35)
    K, V = 20, 70
    o = _oracle(off, words, K, V=V)
    o.sweep(3)
    ll = o.loglik()
    z = o.get_topics()
    doc = _doc_of_token(off)
    wt = np.zeros((V, K))
    np.add.at(wt, (words, z), 1)
    dt = np.zeros((len(off) - 1, K))
    np.add.at(dt, (doc, z), 1)
    a, b = float(np.float32(ALPHA(K))), float(np.float32(BETA))
    phi = (wt + b) / (wt.sum(0) + V * b)
    theta = (dt + a) / (dt.sum(1, keepdims=True) + K * a)
    exp = np.log((theta[doc] * phi[words]).sum(1)).sum()
    assert ll == pytest.approx(exp, rel=1e-10)


def test_loglik_k1_closed_form():
    off, words = make_corpus(11, 60, 50, 25)
    V = 50
    o = _oracle(off, words, 1, V=V)
    ll = o.loglik()
    b = float(np.float32(BETA))
    nw = np.bincount(words, minlength=V).astype(np.float64)
    N = len(words)
    exp = (nw * np.log((nw + b) / (N + V * b)))[nw > 0].sum()
    assert ll == pytest.approx(exp, rel=1e-12)


def test_layout_word_major_order():
    off, words = make_corpus(12, 70, 60, 30)
    o = _oracle(off, words, 8, V=60)
    wo, td, dtk, ro = o.get_layout()
    cnt = np.bincount(words, minlength=60)
    ranks = sorted(range(60), key=lambda w: (-cnt[w], w))
    assert [int(ro[w]) for w in ranks] == list(range(60))
    doc = _doc_of_token(off)
    # token ids sort (word rank, doc length desc, doc, position)
    dlen = np.diff(off)[doc]
    order = np.lexsort((np.arange(len(words)), doc, -dlen, ro[words]))
    assert np.array_equal(dtk[order], np.arange(len(words)))
    assert np.array_equal(td, doc[order])
