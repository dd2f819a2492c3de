"""The `ezlda` command and the integration example on the GPU, against the
oracle: preprocess a docword file into the reference's chunked folder, train
with a checkpoint in both forms (the private binary file and the folder's own
TL.txt plus seed and sweep), resume from the folder, fold in new documents;
then the C++ driver of INTEGRATION.md, built and run on a generated folder."""
import collections
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, make_corpus
from lda_sc20_amd import ezlda
from oracle import binding as ob

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "lda_sc20_amd", "ezlda")
BETA = 0.01


def _run(*args):
    r = subprocess.run([CLI, *[str(a) for a in args]], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stdout
    return r.stdout


def _write_docword(path, off, words, V):
    """UCI bag of words: header "D W NNZ", then "doc word count", 1-based."""
    lines = []
    for d in range(len(off) - 1):
        for w, c in sorted(collections.Counter(words[off[d]:off[d + 1]].tolist()).items()):
            lines.append(f"{d + 1} {w + 1} {c}")
    with open(path, "w") as f:
        f.write(f"{len(off) - 1} {V} {len(lines)}\n" + "\n".join(lines) + "\n")


def test_cli_preprocess_train_checkpoints_resume_foldin(tmp_path):
    V, K, seed = 300, 32, 9
    off, words = make_corpus(21, 150, V, 60, zipf=1.2)
    dw = tmp_path / "docword.txt"
    _write_docword(dw, off, words, V)
    folder = tmp_path / "corpus"
    _run("preprocess", "--docword", dw, "--out", folder, "--K", K, "--chunks", 2, "--seed", 3)
    init = ezlda.read_chunked(str(folder))
    assert init.N == len(words) and init.D == 150
    # train 3 sweeps, both checkpoint forms
    ck = tmp_path / "ck.bin"
    _run("train", "--prefix", folder, "--K", K, "--iters", 3, "--seed", seed, "--graph", "--checkpoint", ck,
         "--checkpoint-tl")
    orc = ob.Oracle(K, float(np.float32(50.0 / K)), BETA, seed, init.V, init.doc_off, init.words, z0=init.topics)
    orc.sweep(3)
    after = ezlda.read_chunked(str(folder))  # TL.txt now holds sweep 3's topics
    assert np.array_equal(after.words, init.words)
    assert np.array_equal(after.topics, orc.get_topics())
    state = open(folder / "ezlda_state.txt").read().split()
    assert state == ["seed", str(seed), "sweep", "3", "K", str(K)]
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=seed)
    ctx.load_chunked(str(folder))
    ctx.load_checkpoint(str(ck))
    assert ctx.info().sweep == 3 and np.array_equal(ctx.get_topics(), orc.get_topics())
    ctx.close()
    # resume from the folder: two more sweeps continue the same chain
    _run("train", "--prefix", folder, "--K", K, "--iters", 2, "--seed", seed, "--resume-tl", "--checkpoint-tl")
    orc.sweep(2)
    assert np.array_equal(ezlda.read_chunked(str(folder)).topics, orc.get_topics())
    assert open(folder / "ezlda_state.txt").read().split()[3] == "5"
    # fold in new documents against the model of sweep 5
    roff, rwords = make_corpus(22, 12, V, 40, zipf=1.2)
    rq = tmp_path / "requests.txt"
    _write_docword(rq, roff, rwords, V)
    mix = tmp_path / "mix.txt"
    _run("foldin", "--prefix", folder, "--K", K, "--seed", seed, "--resume-tl", "--iters", 0, "--requests", rq,
         "--foldin-sweeps", 5, "--slots", 4, "--out", mix)
    req = ezlda.read_docword(str(rq))
    _, co = orc.foldin(req.doc_off, req.words, np.arange(req.D, dtype=np.uint32), 5, seed + 7)
    lines = open(mix).read().splitlines()
    assert len(lines) == req.D
    for d, line in enumerate(lines):
        f = line.split()
        assert int(f[0]) == d and int(f[1]) == req.doc_off[d + 1] - req.doc_off[d]
        got = np.zeros(K, np.int64)
        for kc in f[2:]:
            k, c = kc.split(":")
            got[int(k)] = int(c)
        assert np.array_equal(got, co[d]), d


def test_tl_checkpoint_api_and_mismatches(tmp_path):
    V, K = 250, 40
    off, words = make_corpus(23, 120, V, 50, zipf=1.2)
    folder = str(tmp_path / "c")
    ezlda.Corpus(V, off, words).write_chunked(folder, K=K, num_chunks=3, seed=4)
    a = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=11)
    a.load_chunked(folder)
    a.sweep(2)
    a.save_tl(folder)
    z2 = a.get_topics()
    a.sweep(3)
    b = ezlda.Context(K=K, alpha=50.0 / K, beta=BETA, seed=11)
    b.load_tl(folder)
    assert b.info().sweep == 2 and np.array_equal(b.get_topics(), z2)
    b.sweep(3)
    assert np.array_equal(b.get_topics(), a.get_topics())
    # the sidecar was rewritten with the same topics as the text
    fast = ezlda.read_chunked(folder)
    os.remove(os.path.join(folder, "corpus.ezb"))
    assert np.array_equal(fast.topics, ezlda.read_chunked(folder).topics)
    # another seed or K cannot resume the chain
    for kw in ({"seed": 12}, {"K": 41}):
        args = dict(K=K, alpha=50.0 / K, beta=BETA, seed=11)
        args.update(kw)
        c = ezlda.Context(**args)
        with pytest.raises(ezlda.EzldaError) as e:
            c.load_tl(folder)
        assert e.value.code == ezlda.EINVAL
    # a folder of another corpus is refused
    off2, words2 = make_corpus(24, 120, V, 50, zipf=1.2)
    other = str(tmp_path / "o")
    ezlda.Corpus(V, off2, words2).write_chunked(other, K=K, num_chunks=3, seed=4)
    with pytest.raises(ezlda.EzldaError):
        a.save_tl(other)


def test_integration_example_runs(tmp_path):
    """INTEGRATION.md's C++ driver: built, run on an 8-chunk folder with the
    library's defaults (K=1024, seed 2002, 100 graph sweeps), and its
    log-likelihood equal to the oracle's."""
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    src = [b for b in re.findall(r"```cpp\n(.*?)```", txt, flags=re.S) if "int main" in b][0]
    V = 300
    off, words = make_corpus(25, 160, V, 60, zipf=1.2)
    folder = tmp_path / "pubmed"
    ezlda.Corpus(V, off, words).write_chunked(str(folder), K=1024, num_chunks=8, seed=6)
    assert "/path/to/pubmed" in src
    (tmp_path / "driver.cpp").write_text(src.replace("/path/to/pubmed", str(folder)))
    lib = os.path.join(ROOT, "lda_sc20_amd")
    r = subprocess.run([shutil.which("g++"), "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
                        str(tmp_path / "driver.cpp"), "-o", str(tmp_path / "driver"), f"-L{lib}", "-lezlda",
                        f"-Wl,-rpath,{lib}"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    run = subprocess.run([str(tmp_path / "driver")], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=180, cwd=str(tmp_path))
    assert run.returncode == 0, run.stdout
    out = run.stdout.splitlines()
    assert len(out) == 101 and all(float(x) > 0 for x in out[:100])
    ll = float(out[100].split()[-1])
    c = ezlda.read_chunked(str(folder))
    orc = ob.Oracle(1024, float(np.float32(50.0 / 1024)), BETA, 2002, c.V, c.doc_off, c.words, z0=c.topics)
    orc.sweep(100)
    assert ll == pytest.approx(orc.loglik(), rel=1e-6)
