"""The C-ABI library on the CPU: it loads, exports every symbol the header
declares, and its host-only parts (corpus generator, docword reader, chunked
format writer/reader, error paths) work without a GPU."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, make_corpus
from lda_sc20_amd import ezlda

HEADER = os.path.join(ROOT, "include", "ezlda.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ezlda_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_expected_symbols():
    assert header_functions() == sorted(ezlda.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ezlda.lib()
    for name in header_functions():
        assert hasattr(L, name), name


def test_header_has_no_gpu_or_torch_types():
    code = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    idents = set(re.findall(r"[A-Za-z_][A-Za-z0-9_]*", code))
    bad = [i for i in idents if re.match(r"(hip|cuda|cu|torch|at|c10)[A-Z_]", i) or i.endswith("Stream_t")]
    assert not bad, bad
    assert "#include <hip" not in code and "#include <cuda" not in code and "torch" not in code


def test_create_fails_loudly_without_device():
    import torch  # noqa: F401  (device presence only)

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(ezlda.EzldaError) as e:
        ezlda.Context(K=16)
    assert "device" in str(e.value).lower()


def test_default_args_mirror_reference_argument_surface():
    a = ezlda.default_args()
    # EZLDA/Argument.cuh:3-9 (K, alpha = 50/K, beta); the launch shape is
    # this build's measured default: workgroup size and persistent grid
    # chosen from K and the device (block_dim 0, grid_dim < 0), runs of up
    # to 4 tokens
    assert a.K == 1024 and a.beta == pytest.approx(0.01)
    assert a.alpha == pytest.approx(50.0 / 1024)
    assert a.block_dim == 0 and a.grid_dim == -1 and a.run_cap == 4 and a.token_segment == 4096


# -------------------------------------------------------------- generator
def test_generator_is_deterministic_and_thread_independent():
    a = ezlda.gen_corpus(V=500, D=200, K_true=10, len_mean=40, seed=5, threads=1)
    b = ezlda.gen_corpus(V=500, D=200, K_true=10, len_mean=40, seed=5, threads=7)
    c = ezlda.gen_corpus(V=500, D=200, K_true=10, len_mean=40, seed=6, threads=3)
    assert np.array_equal(a.doc_off, b.doc_off) and np.array_equal(a.words, b.words)
    assert not np.array_equal(a.words, c.words)


def test_generator_first_doc_continues_the_same_corpus():
    a = ezlda.gen_corpus(V=500, D=200, K_true=10, len_mean=40, seed=5)
    b = ezlda.gen_corpus(V=500, D=50, K_true=10, len_mean=40, seed=5, first_doc=150)
    assert np.array_equal(np.diff(a.doc_off)[150:], np.diff(b.doc_off))
    assert np.array_equal(a.words[a.doc_off[150]:], b.words)


def test_generator_poisson_lengths_and_ranges():
    c = ezlda.gen_corpus(V=1000, D=4000, K_true=16, len_dist=0, len_mean=100, seed=1)
    n = np.diff(c.doc_off)
    assert n.min() >= 1 and abs(n.mean() - 100) < 1.5 and abs(n.var() - 100) < 15
    assert c.words.min() >= 0 and c.words.max() < 1000


def test_generator_lognormal_mean_and_zipf_skew():
    c = ezlda.gen_corpus(V=20000, D=3000, K_true=64, len_dist=1, len_mean=330, len_sigma=0.8, zipf_s=1.07,
                         seed=2)
    n = np.diff(c.doc_off)
    assert abs(n.mean() / 330 - 1) < 0.05
    cnt = np.sort(np.bincount(c.words, minlength=20000))[::-1]
    # Zipf base: the top 1% of words carry a large share of the tokens
    assert cnt[:200].sum() / cnt.sum() > 0.4


def test_generator_topic_structure():
    # theta ~ Dir(0.1): each doc is dominated by a few true topics, so its
    # word distribution is far from the corpus-wide one.
    c = ezlda.gen_corpus(V=2000, D=300, K_true=50, len_mean=200, seed=3)
    glob = np.bincount(c.words, minlength=2000) / len(c.words)
    d = c.words[c.doc_off[0]:c.doc_off[1]]
    loc = np.bincount(d, minlength=2000) / len(d)
    assert 0.5 * np.abs(loc - glob).sum() > 0.3


# ------------------------------------------------------------ file formats
def _canon(c: ezlda.Corpus):
    """Docs as sorted tuples of (word, topic) for order-free comparison."""
    out = []
    for d in range(c.D):
        w = c.words[c.doc_off[d]:c.doc_off[d + 1]]
        t = c.topics[c.doc_off[d]:c.doc_off[d + 1]] if c.topics is not None else np.zeros_like(w)
        out.append(tuple(sorted(zip(w.tolist(), t.tolist()))))
    return sorted(out)


def test_chunked_roundtrip(tmp_path):
    off, words = make_corpus(3, 120, 300, 40)
    V, K = 300, 32
    rng = np.random.default_rng(0)
    z0 = rng.integers(0, K, len(words)).astype(np.uint16)
    src = ezlda.Corpus(V, off, words)
    src.write_chunked(str(tmp_path / "c"), K=K, num_chunks=4, z0=z0)
    back = ezlda.read_chunked(str(tmp_path / "c"))
    assert back.D == src.D and back.N == src.N and back.V == V
    # word ids come back as ranks (count desc, ties by id)
    cnt = np.bincount(words, minlength=V)
    order = sorted(range(V), key=lambda w: (-cnt[w], w))
    rank = np.empty(V, np.int64)
    rank[order] = np.arange(V)
    ranked = ezlda.Corpus(V, off, rank[words].astype(np.int32), z0)
    assert _canon(back) == _canon(ranked)


def test_chunked_files_follow_reference_layout(tmp_path):
    off, words = make_corpus(4, 60, 100, 30)
    K = 16
    ezlda.Corpus(100, off, words).write_chunked(str(tmp_path / "c"), K=K, num_chunks=2)
    root = tmp_path / "c"
    lv = [int(x) for x in open(root / "lengthVec.txt").read().split()]
    assert len(lv) == 8 and lv[4] == 100 and lv[6] + lv[7] == 100
    tll = [int(x) for x in open(root / "TLLength.txt").read().split()]
    assert sum(tll) == len(words) and lv[0] == max(tll)
    cnt = np.bincount(words, minlength=100)
    assert lv[6] == int((cnt >= K).sum())  # head = count >= K (streamTestBalance.cpp:291)
    for ch in range(2):
        d = root / f"chunk{ch}"
        tl = [int(x) for x in open(d / "TL.txt").read().split()]
        assert min(tl) >= 1 and max(tl) <= K and len(tl) == tll[ch]
        wco = np.array(open(d / "wordCountOffset.txt").read().split(), np.int64).reshape(-1, 2)
        assert wco.shape == (100, 2) and wco[:, 0].sum() == tll[ch]
        assert np.array_equal(wco[:, 1], np.concatenate([[0], np.cumsum(wco[:-1, 0])]))
        dtco = np.array(open(d / "DTCountOffset.txt").read().split(), np.int64).reshape(-1, 2)
        dco = np.array(open(d / "docCountOffset.txt").read().split(), np.int64).reshape(-1, 2)
        assert np.array_equal(dtco[:, 0], np.minimum(((dco[:, 0] - 1) // 32 + 1) * 32, K))
        w2d = [int(x) for x in open(d / "word2DocMap.txt").read().split()]
        assert sorted(w2d) == list(range(tll[ch]))


def test_chunked_reader_errors(tmp_path):
    with pytest.raises(ezlda.EzldaError, match="lengthVec"):
        ezlda.read_chunked(str(tmp_path / "missing"))
    off, words = make_corpus(5, 30, 50, 20)
    ezlda.Corpus(50, off, words).write_chunked(str(tmp_path / "c"), K=8, num_chunks=2)
    tl = tmp_path / "c" / "chunk1" / "TL.txt"
    lines = open(tl).read().splitlines()
    open(tl, "w").write("\n".join(lines[:-3]) + "\n")
    with pytest.raises(ezlda.EzldaError, match="TL.txt"):
        ezlda.read_chunked(str(tmp_path / "c"))
    with pytest.raises(ezlda.EzldaError, match="chunks"):
        ezlda.read_chunked(str(tmp_path / "c"), 3)


def test_docword_reader(tmp_path):
    p = tmp_path / "docword.txt"
    p.write_text("3 5 5\n1 1 2\n1 4 1\n2 2 3\n3 5 1\n3 1 1\n")
    c = ezlda.read_docword(str(p))
    assert c.D == 3 and c.V == 5
    assert list(np.diff(c.doc_off)) == [3, 3, 2]
    assert sorted(c.words[0:3].tolist()) == [0, 0, 3]
    assert c.words[3:6].tolist() == [1, 1, 1]
    bad = tmp_path / "bad.txt"
    bad.write_text("2 5 3\n1 1 2\n2 9 1\n")
    with pytest.raises(ezlda.EzldaError):
        ezlda.read_docword(str(bad))


def test_chunked_reader_rejects_negative_counts(tmp_path):
    """A negative chunk token or document count is an error code, not an
    exception escaping the C ABI (which would abort the host process)."""
    off, words = make_corpus(6, 30, 50, 20)
    ezlda.Corpus(50, off, words).write_chunked(str(tmp_path / "c"), K=8, num_chunks=2)
    root = tmp_path / "c"
    good_tll = open(root / "TLLength.txt").read()
    vals = good_tll.split()
    open(root / "TLLength.txt", "w").write(f"-1\n{vals[1]}\n")
    with pytest.raises(ezlda.EzldaError, match="negative or oversized"):
        ezlda.read_chunked(str(root))
    open(root / "TLLength.txt", "w").write(good_tll)
    dl = open(root / "docLength.txt").read().split()
    open(root / "docLength.txt", "w").write(f"{dl[0]}\n-5\n")
    with pytest.raises(ezlda.EzldaError, match="negative or oversized"):
        ezlda.read_chunked(str(root))



# ------------------------------------------------------ binary sidecar
def _same(a: ezlda.Corpus, b: ezlda.Corpus):
    return (a.V == b.V and np.array_equal(a.doc_off, b.doc_off) and np.array_equal(a.words, b.words)
            and np.array_equal(a.topics, b.topics))


def test_sidecar_equals_text_reader(tmp_path):
    off, words = make_corpus(7, 200, 400, 50)
    root = tmp_path / "c"
    ezlda.Corpus(400, off, words).write_chunked(str(root), K=24, num_chunks=3, seed=5)
    assert (root / "corpus.ezb").exists()
    fast = ezlda.read_chunked(str(root))
    os.remove(root / "corpus.ezb")
    text = ezlda.read_chunked(str(root))
    assert _same(fast, text)


def test_sidecar_ignored_once_a_text_file_changes(tmp_path):
    off, words = make_corpus(8, 100, 300, 40)
    root = tmp_path / "c"
    ezlda.Corpus(300, off, words).write_chunked(str(root), K=24, num_chunks=2, seed=5)
    before = ezlda.read_chunked(str(root))
    tl = root / "chunk1" / "TL.txt"
    vals = open(tl).read().split()
    vals[0] = str(int(vals[0]) % 24 + 1)  # another topic for the first token of chunk 1's TL
    open(tl, "w").write("\n".join(vals) + "\n")
    after = ezlda.read_chunked(str(root))
    assert not np.array_equal(before.topics, after.topics)
    os.remove(root / "corpus.ezb")
    assert _same(after, ezlda.read_chunked(str(root)))


def test_sidecar_rejects_a_foreign_or_truncated_file(tmp_path):
    off, words = make_corpus(9, 80, 200, 30)
    root = tmp_path / "c"
    ezlda.Corpus(200, off, words).write_chunked(str(root), K=16, num_chunks=2, seed=5)
    good = ezlda.read_chunked(str(root))
    side = root / "corpus.ezb"
    data = side.read_bytes()
    side.write_bytes(data[: len(data) // 2])  # truncated: the text path is taken
    assert _same(good, ezlda.read_chunked(str(root)))
    side.write_bytes(b"NOTEZLDA" + data[8:])
    assert _same(good, ezlda.read_chunked(str(root)))


# ------------------------------------------- the integration example
def test_integration_example_compiles_and_links(tmp_path):
    """The C++ driver shown in INTEGRATION.md builds against include/ezlda.h
    and links with libezlda.so (run on the GPU by tests/test_gpu_cli.py)."""
    import re
    import shutil
    import subprocess

    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = [b for b in re.findall(r"```cpp\n(.*?)```", txt, flags=re.S) if "int main" in b]
    assert blocks, "no C++ driver in INTEGRATION.md"
    src = tmp_path / "driver.cpp"
    src.write_text(blocks[0])
    gxx = shutil.which("g++")
    lib = os.path.join(ROOT, "lda_sc20_amd")
    r = subprocess.run([gxx, "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o",
                        str(tmp_path / "driver"), f"-L{lib}", "-lezlda", f"-Wl,-rpath,{lib}"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
