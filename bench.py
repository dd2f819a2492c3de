#!/usr/bin/env python3
"""Sampled tokens/s per Gibbs sweep at K=1024 on the ~100M-token synthetic
corpus (V=102,660, D=300,000, log-normal doc lengths mean 330 / sigma 0.8,
Zipf(1.07) vocabulary, dense WT head = top 1% of words), on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]     (defaults: K=20, W=5)

A step is one whole sweep: WT/n_k rebuild, top-3, DT rebuild, sampling.  The
corpus and model are resident in HBM before the timed region; the K timed
sweeps run back to back on the library's stream with HIP events around every
stage (no host synchronisation inside the region); they give the value, the
stage split and the sampler's roofline.  K more sweeps replayed from one
captured hipGraph are reported beside them.  Prints one JSON line.

Other modes (each prints one JSON line and exits): --copy-calibration and
--gather-calibration run known-size kernels for scaling traffic counters;
--cpu-protocol runs the CPU-baseline protocol on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # north-star corpus: ~100M tokens, K=1024
    "cfg2": dict(gen=dict(V=102660, D=300000, K_true=1024, len_dist=1, len_mean=330.0, len_sigma=0.8,
                          zipf_s=1.07, phi_conc=0.01, theta_conc=0.1, seed=1002),
                 K=1024, seed=2002, head_frac=0.01,
                 desc="V=102660, D=300000, lognormal(mean 330, sigma 0.8), Zipf(1.07) base, ~99M tokens, "
                      "K=1024, alpha=50/K, beta=0.01, dense head = top 1% of words"),
    # PubMed-sized corpus, fully dense WT (K in {1024, 4096, 10000} via --K)
    "cfg3": dict(gen=dict(V=141043, D=8200000, K_true=1024, len_dist=1, len_mean=90.0, len_sigma=0.8,
                          zipf_s=1.07, phi_conc=0.01, theta_conc=0.1, seed=1003),
                 K=1024, seed=2003, head_frac=1.0,
                 desc="V=141043, D=8.2M, lognormal(mean 90, sigma 0.8), Zipf(1.07) base, ~738M tokens, "
                      "alpha=50/K, beta=0.01, dense WT for every word"),
    # heavy-tailed document lengths: Pareto(1.2) with x_m = 48, clamped to [1, 20000]
    "cfg4": dict(gen=dict(V=100000, D=2000000, K_true=1024, len_dist=2, pareto_shape=1.2, pareto_xm=48.0,
                          len_min=1, len_max=20000, zipf_s=1.1, phi_conc=0.01, theta_conc=0.1, seed=1004),
                 K=2048, seed=2004, head_frac=0.01,
                 desc="V=100000, D=2M, Pareto(shape 1.2, x_m 48) lengths in [1, 20000], Zipf(1.1) base, "
                      "~400M tokens, K=2048, alpha=50/K, beta=0.01, dense head = top 1% of words"),
    # the reference's CPU-sized case
    "cfg1": dict(gen=dict(V=10000, D=10000, K_true=64, len_dist=0, len_mean=100.0, len_sigma=0.0,
                          zipf_s=0.0, phi_conc=0.01, theta_conc=0.1, seed=1001),
                 K=64, seed=2001, head_frac=None,
                 desc="V=10000, D=10000, Poisson(100), ~1M tokens, K=64, alpha=50/K, beta=0.01"),
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_sync() -> None:
    try:
        ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()
    except OSError:
        pass


def measured_traffic(workload: str, run: dict):
    """Bytes per k_sample launch that crossed the L2 boundary, measured by
    rocprofv3 on this workload (FETCH_SIZE then WRITE_SIZE, each scaled by a
    calibration kernel of known bytes), with where it came from -- or None
    when the committed measurement was taken with another configuration or
    sweep window than this run's (its `config` must match `run` key by key)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    try:
        with open(path) as f:
            m = json.load(f)
        prof = m["config"]
        if any(prof.get(k) != v for k, v in run.items()):
            return None
        return m
    except (OSError, KeyError, ValueError, TypeError):
        return None


def host_cpu() -> dict:
    """Model name, sockets and cores of this host (lscpu)."""
    out = {"model": None, "sockets": None, "cores_per_socket": None, "threads": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                out["model"] = v
            elif k == "Socket(s)":
                out["sockets"] = int(v)
            elif k == "Core(s) per socket":
                out["cores_per_socket"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return out


def oracle_sweep_rates(corpus, K: int, seed: int, max_docs: int, sweeps: int, omp: bool, label: str):
    """The CPU oracle (OpenMP build or single-threaded) on the first max_docs
    docs of the corpus, full vocabulary: tokens/s of each of `sweeps` sweeps
    (each a whole sweep: counts, top-3, DT rows, sampling), timed one by one."""
    from oracle import binding as ob

    D = min(max_docs, corpus.D)
    off = corpus.doc_off[: D + 1].copy()
    words = corpus.words[: int(off[-1])].copy()
    orc = ob.Oracle(K, float(np.float32(50.0 / K)), 0.01, seed, corpus.V, off, words, omp=omp)
    n = int(off[-1])
    rates = []
    for i in range(sweeps):
        t = time.perf_counter()
        orc.sweep(1)
        dt = time.perf_counter() - t
        rates.append(n / dt)
        log(f"{label}: sweep {i + 1}/{sweeps} {dt:.2f}s ({n / dt:.3e} tokens/s)")
    cores = int(ob.lib(omp).orc_num_threads())
    orc.close()
    return D, n, cores, rates


def cpu_baseline(corpus, K: int, seed: int, max_docs: int, sweeps: int = 11):
    """The OpenMP build of the CPU oracle (same sweep semantics, bit-exact) on
    the first max_docs docs of the same corpus: the median tokens/s over
    sweeps 2..sweeps (sweep 1 is warm-up)."""
    D, n, cores, rates = oracle_sweep_rates(corpus, K, seed, max_docs, sweeps, True, "cpu baseline")
    cpu = host_cpu()
    return {"value": float(np.median(rates[1:])), "unit": "tokens/s", "cores": cores, "kind": "port",
            "sample": f"first {D} docs ({n} tokens) of the same corpus, full vocabulary, K={K}; median tokens/s of "
                      f"sweeps 2-{sweeps} (sweep 1 is warm-up), OpenMP oracle ({cores} threads)",
            "host_cpu": cpu, "per_sweep_tokens_per_s": rates}


def cpu_protocol(parts) -> dict:
    """The CPU-baseline protocol: the single-threaded oracle on the reference's
    CPU-sized corpus (all 100 sweeps) and on a 3-sweep sample of the ~99M-token
    corpus; the OpenMP oracle on those and on the 738M-token corpus at K=1024.
    Median tokens/s over sweeps 2-11 (sweep 1 is warm-up; fewer where noted).
    Prints progress per sweep; returns one dict per part."""
    from lda_sc20_amd import ezlda

    out = {"host_cpu": host_cpu()}
    plan = {
        "cfg1_1t": ("cfg1", None, 100, False),
        "cfg2_1t": ("cfg2", 30000, 3, False),
        "cfg1_omp": ("cfg1", None, 11, True),
        "cfg2_omp": ("cfg2", None, 11, True),
        "cfg3_omp": ("cfg3", None, 4, True),
    }
    corpora = {}
    for part in parts:
        name, docs, sweeps, omp = plan[part]
        wl = WORKLOADS[name]
        if name not in corpora:
            t = time.perf_counter()
            corpora[name] = ezlda.gen_corpus(**wl["gen"])
            log(f"{name}: generated {corpora[name].N} tokens in {time.perf_counter() - t:.1f}s")
        c = corpora[name]
        D, n, cores, rates = oracle_sweep_rates(c, wl["K"], wl["seed"], docs or c.D, sweeps, omp, part)
        out[part] = {"workload": f"{name}: {wl['desc']}", "docs": D, "tokens": n, "K": wl["K"], "threads": cores,
                     "sweeps": sweeps, "median_tokens_per_s_after_first": float(np.median(rates[1:])),
                     "median_window": f"sweeps 2-{min(sweeps, 11)}" if sweeps <= 11 else "sweeps 2-11",
                     "per_sweep_tokens_per_s": rates}
        if sweeps > 11:
            out[part]["median_tokens_per_s_after_first"] = float(np.median(rates[1:11]))
            out[part]["all_sweeps_tokens_per_s"] = n * sweeps / sum(n / r for r in rates)
    return out


# fold-in requests: further documents of the training corpus's generative
# model (same seed and topics, documents numbered after the training ones)
SERVE_GEN = dict(len_dist=1, len_mean=300.0, len_sigma=0.8, len_min=10, len_max=5000)
SERVE_SWEEPS = 20


def _wait_until(t: float) -> None:
    while True:
        d = t - time.perf_counter()
        if d <= 0:
            return
        if d > 2e-3:
            time.sleep(d - 1e-3)


def foldin_perplexity(wt, nk, beta: float, alpha: float, docs, counts) -> float:
    """FP64 perplexity of fold-in documents: p(w|d) = sum_k theta_dk phi_wk with
    theta_dk = (n_dk + alpha) / (n_d + K alpha) from the folded-in counts."""
    V, K = wt.shape
    inv = 1.0 / (nk.astype(np.float64) + V * beta)
    ll, n = 0.0, 0
    for w, c in zip(docs, counts):
        theta = (c + alpha) / (c.sum() + K * alpha)
        ll += float(np.log(((wt[w].astype(np.float64) + beta) * inv) @ theta).sum())
        n += len(w)
    return math.exp(-ll / n)


def serving_bench(ctx, ezlda, wl: dict, args) -> dict:
    """Fold-in serving against the trained model frozen: an open-loop burst
    measures batched capacity, then Poisson arrivals at fractions of it give
    p50/p99 latency; the plain one-document-per-launch path runs beside it."""
    fr = [float(x) for x in args.serve_levels.split(",") if x]
    n_burst, secs = args.serve_docs, args.serve_seconds
    K = ctx.args.K
    alpha = float(np.float32(50.0 / K))
    g = dict(wl["gen"])
    base_docs = g.pop("D")
    g.update(SERVE_GEN)
    t = time.perf_counter()
    # enough documents for the burst plus the paced levels at up to 1.5x the
    # burst's rate; generated once, reused round-robin if a level needs more
    req = ezlda.gen_corpus(D=max(4 * n_burst, 1), first_doc=base_docs, **g)
    docs = [req.words[req.doc_off[i]: req.doc_off[i + 1]] for i in range(req.D)]
    log(f"serving: {req.D} request docs ({req.N} tokens) in {time.perf_counter() - t:.1f}s")
    srv = ezlda.Server(ctx, max_slots=args.serve_slots, sweeps=SERVE_SWEEPS,
                       hbm_budget_bytes=args.serve_budget, seed=wl["seed"] + 7)
    rid = 0

    def run(n: int, qps):
        nonlocal rid
        srv.reset_stats()
        first = rid
        rng = np.random.default_rng(rid + 1)
        gaps = rng.exponential(1.0 / qps, n) if qps else np.zeros(n)
        t0 = time.perf_counter()
        at = t0 + np.cumsum(gaps)
        if qps:  # Poisson arrivals, one document per call
            for j in range(n):
                _wait_until(at[j])
                srv.submit(rid, docs[rid % len(docs)])
                rid += 1
        else:  # the burst: one batch call
            batch = [docs[(rid + j) % len(docs)] for j in range(n)]
            boff = np.concatenate([[0], np.cumsum([len(b) for b in batch])]).astype(np.int64)
            srv.submit_many(np.arange(rid, rid + n, dtype=np.uint32), boff, np.concatenate(batch))
            rid += n
        submit_s = time.perf_counter() - t0
        srv.drain()
        wall = time.perf_counter() - t0
        st = srv.stats()
        out = {}
        for r in range(first, rid):
            _, c, _ = srv.wait(r, len(docs[r % len(docs)]), timeout_ms=60000)
            out[r] = c
        lvl = {"offered_qps": qps, "docs": n, "docs_per_s": st.docs_per_s, "p50_ms": st.p50_ms,
               "p99_ms": st.p99_ms, "max_ms": st.max_ms, "mean_ms": st.mean_ms, "max_active": st.max_active,
               "launches": st.steps, "wall_s": wall, "submit_s": submit_s}
        return lvl, out

    run(min(256, n_burst), None)  # warm-up
    burst, burst_counts = run(n_burst, None)
    cap = burst["docs_per_s"]
    log(f"serving burst: {n_burst} docs, {cap:.0f} docs/s, p50 {burst['p50_ms']:.1f} ms, "
        f"p99 {burst['p99_ms']:.1f} ms, max active {burst['max_active']}, submitted in {burst['submit_s'] * 1e3:.1f} ms, "
        f"{burst['launches']} steps")
    levels = []
    for f in fr:
        qps = f * cap
        n = max(16, int(qps * secs))
        lvl, _ = run(n, qps)
        lvl["fraction_of_burst_rate"] = f
        levels.append(lvl)
        log(f"serving at {qps:.0f} qps: {lvl['docs_per_s']:.0f} docs/s, p50 {lvl['p50_ms']:.2f} ms, "
            f"p99 {lvl['p99_ms']:.2f} ms")
    # plain path on the first documents of the burst (same request ids)
    first_burst = rid - n_burst - sum(l["docs"] for l in levels)
    n_plain = min(args.serve_plain_docs, n_burst)
    t = time.perf_counter()
    same = True
    plain_counts = []
    for r in range(first_burst, first_burst + n_plain):
        _, c = srv.plain(r, docs[r % len(docs)])
        same &= bool(np.array_equal(c, burst_counts[r]))
        plain_counts.append(c)
    plain_rate = n_plain / (time.perf_counter() - t)
    log(f"plain path: {plain_rate:.1f} docs/s over {n_plain} docs; batched/plain = {cap / plain_rate:.1f}x")
    srv.close()
    wt, nk = ctx.get_wt_dense(), ctx.get_nk()
    sample = list(range(first_burst, first_burst + min(256, n_burst)))
    ppl_b = foldin_perplexity(wt, nk, 0.01, alpha, [docs[r % len(docs)] for r in sample],
                              [burst_counts[r] for r in sample])
    pdocs = list(range(first_burst, first_burst + n_plain))
    ppl_p = foldin_perplexity(wt, nk, 0.01, alpha, [docs[r % len(docs)] for r in pdocs], plain_counts)
    ppl_pb = foldin_perplexity(wt, nk, 0.01, alpha, [docs[r % len(docs)] for r in pdocs],
                               [burst_counts[r] for r in pdocs])
    return {
        "workload": f"fold-in against the frozen K={K} model just trained (V={wl['gen']['V']}); requests are "
                    f"further documents of the same generative model, lognormal lengths mean 300 in [10, 5000]; "
                    f"{SERVE_SWEEPS} fold-in sweeps, alpha=50/K; Poisson arrivals at fractions of the burst rate",
        "sweeps": SERVE_SWEEPS, "max_slots": args.serve_slots, "hbm_budget_bytes": args.serve_budget,
        "burst": burst, "levels": levels,
        "plain_docs_per_s": plain_rate, "plain_docs": n_plain,
        "batched_vs_plain": cap / plain_rate,
        "plain_equals_batched": same,
        "perplexity_batched_sample": ppl_b,
        "perplexity_plain_docs": {"plain": ppl_p, "batched": ppl_pb},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-docs", type=int, default=30000)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--K", type=int, default=0, help="topics (0 = the workload's)")
    ap.add_argument("--autotune", action="store_true",
                    help="time the default layout, then tune the layout (its sweeps count as warm-up) and time "
                         "the tuned one")
    ap.add_argument("--run-cap", type=int, default=0, help="tokens per shared DT scan (0 = library default)")
    ap.add_argument("--grid-dim", type=int, default=-1,
                    help="persistent sampler workgroups (-1 = library default, 0 = one per item)")
    ap.add_argument("--block-dim", type=int, default=0, help="sampler workgroup size (0 = library default)")
    ap.add_argument("--doc-tiles", type=int, default=0,
                    help="doc tiles for head-word items (0 auto, <0 off); topics do not depend on it")
    ap.add_argument("--no-serve", action="store_true", help="skip the fold-in serving measurement")
    ap.add_argument("--serve-docs", type=int, default=8192, help="open-loop burst size")
    ap.add_argument("--serve-levels", default="0.25,0.5,0.8",
                    help="Poisson arrival rates as fractions of the burst's docs/s")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-slots", type=int, default=2048)
    ap.add_argument("--chunks", type=int, default=1,
                    help="> 1: out-of-core sweeps over that many chunks in pinned host memory (times whole sweeps, "
                         "PCIe transfers included)")
    ap.add_argument("--serve-budget", type=int, default=1 << 30, help="HBM bytes for resident requests")
    ap.add_argument("--serve-plain-docs", type=int, default=256)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (to scale HBM counters) and exit")
    ap.add_argument("--gather-calibration", type=int, default=0, metavar="BYTES",
                    help="only run scattered 64-byte block reads (the sampler's row pattern), one per distinct "
                         "block of a BYTES buffer, and exit")
    ap.add_argument("--cpu-protocol", default="", metavar="PARTS",
                    help="only run the CPU-baseline protocol parts (comma list of cfg1_1t, cfg2_1t, cfg1_omp, "
                         "cfg2_omp, cfg3_omp) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the Gibbs sweep here is single-GPU (--gpus {args.gpus} is not "
              f"supported; see DESIGN.md for how the path would shard)", file=sys.stderr)
        return 2

    from lda_sc20_amd import ezlda

    if args.copy_calibration:
        ms = ezlda.selftest_copy(args.copy_calibration, 10)
        print(json.dumps({"copy_bytes": args.copy_calibration, "copies": 11, "ms_per_copy": ms,
                          "GBs": 2 * args.copy_calibration / ms / 1e6}), flush=True)
        return 0
    if args.gather_calibration:
        nblk = 1
        while nblk * 2 * 64 <= args.gather_calibration:
            nblk *= 2
        ms = ezlda.selftest_gather(args.gather_calibration, nblk, 10)
        print(json.dumps({"buffer_bytes": args.gather_calibration, "reads": nblk, "read_bytes": 64 * nblk,
                          "launches": 11, "ms_per_launch": ms, "GBs": 64 * nblk / ms / 1e6}), flush=True)
        return 0
    if args.cpu_protocol:
        print(json.dumps(cpu_protocol([p for p in args.cpu_protocol.split(",") if p])), flush=True)
        return 0

    wl = WORKLOADS[args.workload]
    t = time.perf_counter()
    corpus = ezlda.gen_corpus(**wl["gen"])
    log(f"generated {corpus.N} tokens in {time.perf_counter() - t:.1f}s")
    K = args.K or wl["K"]
    dense = -1 if wl["head_frac"] is None else int(math.ceil(wl["head_frac"] * corpus.V))
    ctx = ezlda.Context(K=K, alpha=50.0 / K, beta=0.01, seed=wl["seed"], dense_words=dense, shortcut=1,
                        doc_tiles=args.doc_tiles, run_cap=args.run_cap, grid_dim=args.grid_dim,
                        block_dim=args.block_dim, num_chunks=args.chunks)
    chunked = args.chunks > 1
    if chunked:  # whole sweeps only (no stage times, graphs or layout tuning)
        args.no_graph, args.autotune = True, False
    t = time.perf_counter()
    ctx.load(corpus)
    info = ctx.info()
    log(f"loaded in {time.perf_counter() - t:.1f}s: N={info.num_tokens} V={info.num_words} D={info.num_docs} "
        f"V_D={info.num_head_words} items={info.num_sample_items} doc_tiles={info.doc_tiles} xcd={info.num_xcd} "
        f"device={info.device_bytes / 2**30:.2f} GiB")

    t = time.perf_counter()
    if args.no_graph or args.warmup < 1:
        ctx.sweep(args.warmup)
    else:  # the last warm-up sweep captures and first launches the graph
        ctx.sweep(args.warmup - 1)
        ctx.sweep_graph(1)
    log(f"{args.warmup} warm-up sweeps in {time.perf_counter() - t:.2f}s")
    tuning = None
    if args.autotune:
        n_cmp = max(5, min(20, args.steps // 4))
        # the plain launch path first: one workgroup of 256 per item, one DT
        # scan per token, no doc tiles (the reference's launch shape)
        ctx.set_layout(grid_dim=0, block_dim=256, run_cap=1, doc_tiles=-1)
        plain_ms, _ = ctx.time_sweeps(n_cmp, use_graph=False)
        ctx.set_layout(grid_dim=-1, block_dim=0, run_cap=0, doc_tiles=0)
        default_ms, _ = ctx.time_sweeps(n_cmp, use_graph=False)
        t = time.perf_counter()
        best, tuned_ms, trials = ctx.autotune(sweeps_per_trial=3)
        tuning = {"plain_launch_ms_per_sweep": plain_ms / n_cmp,
                  "default_layout_ms_per_sweep": default_ms / n_cmp, "default_layout_sweeps": n_cmp,
                  "tuned_ms_per_sweep_in_tuner": tuned_ms, "trials": trials, "tuning_s": time.perf_counter() - t,
                  "chosen": {k: getattr(best, k) for k in ("dense_words", "run_cap", "doc_tiles", "block_dim",
                                                           "token_segment", "grid_dim")}}
        log(f"autotune: plain launch path {plain_ms / n_cmp:.3f}, default layout {default_ms / n_cmp:.3f} ms/sweep "
            f"-> {tuned_ms:.3f} ms/sweep with "
            f"{tuning['chosen']} ({len(trials)} trials, {tuning['tuning_s']:.1f}s)")
        info = ctx.info()

    device_sync()
    sweeps_before = int(ctx.info().sweep)
    K_steps = args.steps
    N = int(info.num_tokens)
    ctx.reset_stats()
    t0 = time.perf_counter()
    total_ms, stage_ms = ctx.time_sweeps(K_steps, use_graph=False)
    device_sync()
    wall = time.perf_counter() - t0
    n_full, n_scanned, tail_nnz, dt_nnz = ctx.stats()
    eager_ms = ms_per_step = total_ms / K_steps
    value = N * K_steps / (total_ms / 1e3)
    log(f"timed {K_steps} sweeps: {total_ms:.2f} ms (wall {wall * 1e3:.1f} ms), stages {stage_ms}")
    graph_ms = None
    if not args.no_graph:  # the next K sweeps as replays of the captured sweep
        gtot, _ = ctx.time_sweeps(K_steps, use_graph=True)
        graph_ms = gtot / K_steps
        log(f"{K_steps} more sweeps as graph replays: {graph_ms:.3f} ms/sweep")

    # algorithmic bytes of one sampler launch: the two z writes per token (by
    # token and by doc slot); per run (a word's tokens in one doc) its entry
    # and its slot's record of n_dk1 and the row length (16 + 4); phi staging
    # once per word; one read of every scanned DT entry (one scan per run
    # with a full draw; 2-byte entries when K <= 8192, else 4)
    VD = int(info.num_head_words)
    R = int(info.num_runs)
    eb = 2 if K <= 8192 else 4
    s_scanned = n_scanned / K_steps
    b_sample = 4 * N + 20 * R + 4 * K * VD + 4 * tail_nnz + 4 * K + eb * s_scanned
    t_sample = stage_ms["sample"] / K_steps / 1e3 if not chunked else None
    achieved = b_sample / t_sample / 1e9 if not chunked else None
    # whole-sweep algorithmic bytes (same accounting for every stage)
    b_sweep = (b_sample                                                      # sample
               + 2 * N + 2 * N + 4 * K * VD + 4 * tail_nnz                  # n_k + WT rebuild
               + 4 * K * VD + 4 * tail_nnz + 8 * info.num_words             # top-3
               + 2 * N + 4 * N + 4 * N + eb * dt_nnz + 8 * info.num_docs)   # DT + records (z_doc, word, record)
    ll = ctx.loglik()  # first call rebuilds the counts of the final topics
    device_sync()
    t = time.perf_counter()
    ll = ctx.loglik()  # the FP64 log-likelihood pass alone (DT rebuild + reduction), timed
    loglik_ms = (time.perf_counter() - t) * 1e3
    # the committed rocprof measurement of k_sample, used only when it was
    # taken on this configuration over this sweep window (the topics, hence
    # the work, are then the same, sweep by sweep)
    run_cfg = {"workload": args.workload, "K": K, "num_head_words": VD, "run_cap": int(info.run_cap),
               "grid_dim": int(info.grid_dim), "block_dim": int(info.block_dim), "doc_tiles": int(info.doc_tiles),
               "sweeps_before_timing": sweeps_before, "steps": K_steps}
    tr = None if chunked else measured_traffic(args.workload, run_cfg)
    out = {
        "metric": "sampled tokens/s per sweep at K=1024 (+% HBM peak)" if K == 1024 else
                  f"sampled tokens/s per sweep at K={K}",
        "value": value,
        "unit": "tokens/s",
        "n_gpus": 1,
        "steps": K_steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (LDA generative process, Philox-seeded)",
        "config": {"workload": f"{args.workload}: {wl['desc']}", "num_tokens": N, "K": K,
                   "num_head_words": VD, "shortcut": True, "sweeps_before_timing": sweeps_before,
                   "doc_tiles": int(info.doc_tiles), "run_cap": int(info.run_cap), "num_runs": int(info.num_runs),
                   "grid_dim": int(info.grid_dim), "block_dim": int(info.block_dim),
                   "token_segment": int(info.token_segment), "num_chunks": int(info.num_chunks)},
        "roofline": None if chunked else
                    {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": tr["bytes_per_launch"] if tr else None,
                     "kernel": "k_sample", "bytes_per_launch": b_sample, "avg_launch_ms": t_sample * 1e3,
                     "traffic_kind": tr["kind"] if tr else None,
                     "traffic_source": tr["source"] if tr else
                     "none: no committed measurement of this configuration and sweep window",
                     "traffic_window_algorithmic_bytes": tr["algorithmic_bytes_per_launch"] if tr else None},
        "stage_ms_per_sweep": {k: v / K_steps for k, v in stage_ms.items()},
        "sweep_algorithmic_GBs": b_sweep / (ms_per_step / 1e3) / 1e9,
        "full_draw_fraction": n_full / (N * K_steps),
        "dt_entries_per_full_draw": (n_scanned / n_full) if n_full else 0.0,
        "graph_ms_per_step": graph_ms,
        "perplexity_after": math.exp(-ll / N),
        "loglik_ms": loglik_ms,
        "wall_ms_timed_region": wall * 1e3,
    }
    if tuning is not None:
        tuning["graph_vs_plain_launch"] = (tuning["plain_launch_ms_per_sweep"] / graph_ms) if graph_ms else None
        out["autotune"] = tuning
    if not args.no_serve:
        out["serving"] = serving_bench(ctx, ezlda, wl, args)
    ctx.close()
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(corpus, K, wl["seed"], args.cpu_docs)
        log(f"cpu baseline in {time.perf_counter() - t:.1f}s: {out['cpu_baseline']['value']:.3e} tokens/s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
