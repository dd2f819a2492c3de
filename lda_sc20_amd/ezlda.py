"""ctypes binding of include/ezlda.h (libezlda.so, built in-tree).

The library is the product: every call here goes straight to the HIP path.
If libezlda.so is missing the import-time loader raises; there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libezlda.so")

OK, EINVAL, EHIP, EIO, ESTATE, ENOMEM = 0, 1, 2, 3, 4, 5
FOLDIN_TOKEN_BYTES = 16  # EZLDA_FOLDIN_TOKEN_BYTES: server device bytes per resident token
LOGLIK_NOALLOC_MAX_ITEMS = 1 << 20  # EZLDA_LOGLIK_NOALLOC_MAX_ITEMS
STAGES = ("wt", "topk", "dt", "sample")

# Every symbol include/ezlda.h declares (checked by the CPU test suite).
EXPORTS = (
    "ezlda_default_args", "ezlda_last_error", "ezlda_create", "ezlda_destroy", "ezlda_load_docs",
    "ezlda_load_chunked", "ezlda_get_info", "ezlda_build_wt", "ezlda_topk", "ezlda_build_dt",
    "ezlda_sample", "ezlda_shortcut", "ezlda_sample_head", "ezlda_sample_tail", "ezlda_commit_head",
    "ezlda_wt_rowsum", "ezlda_loglik", "ezlda_loglik_noalloc", "ezlda_sweep", "ezlda_sweep_graph", "ezlda_time_sweeps",
    "ezlda_get_topics", "ezlda_set_topics", "ezlda_set_sweep", "ezlda_get_nk", "ezlda_get_wt_dense",
    "ezlda_get_dt_raw", "ezlda_get_top", "ezlda_get_aux", "ezlda_get_inv", "ezlda_save_checkpoint",
    "ezlda_load_checkpoint", "ezlda_save_tl", "ezlda_load_tl", "ezlda_gen_corpus", "ezlda_corpus_num_tokens", "ezlda_corpus_num_docs",
    "ezlda_corpus_num_words", "ezlda_corpus_doc_off", "ezlda_corpus_words", "ezlda_corpus_topics",
    "ezlda_corpus_free", "ezlda_corpus_from_arrays", "ezlda_read_docword", "ezlda_write_chunked",
    "ezlda_read_chunked", "ezlda_selftest_philox", "ezlda_selftest_copy", "ezlda_selftest_gather", "ezlda_get_stats",
    "ezlda_reset_stats", "ezlda_server_create", "ezlda_server_destroy", "ezlda_server_last_error",
    "ezlda_server_submit", "ezlda_server_submit_many", "ezlda_server_wait", "ezlda_server_drain", "ezlda_server_get_stats",
    "ezlda_server_reset_stats", "ezlda_foldin_plain", "ezlda_set_layout", "ezlda_autotune",
    "ezlda_export_model",
)


class EzldaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ezlda error {code}: {msg}")
        self.code = code


class Args(C.Structure):
    _fields_ = [("K", C.c_int32), ("alpha", C.c_float), ("beta", C.c_float), ("block_dim", C.c_int32),
                ("grid_dim", C.c_int32), ("num_chunks", C.c_int32), ("num_streams", C.c_int32),
                ("token_segment", C.c_int32), ("seed", C.c_uint64), ("dense_words", C.c_int32),
                ("dense_threshold", C.c_int32), ("shortcut", C.c_int32), ("topic_bits", C.c_int32),
                ("doc_tiles", C.c_int32), ("run_cap", C.c_int32)]


class Info(C.Structure):
    _fields_ = [("num_tokens", C.c_int64), ("num_words", C.c_int32), ("num_docs", C.c_int32),
                ("num_head_words", C.c_int32), ("K", C.c_int32), ("dt_capacity", C.c_int64),
                ("wt_tail_capacity", C.c_int64), ("num_sample_items", C.c_int32), ("sweep", C.c_uint32),
                ("device_bytes", C.c_int64), ("doc_tiles", C.c_int32), ("num_xcd", C.c_int32),
                ("run_cap", C.c_int32), ("grid_dim", C.c_int32), ("block_dim", C.c_int32),
                ("token_segment", C.c_int32), ("num_runs", C.c_int64), ("num_chunks", C.c_int32)]


class TuneTrial(C.Structure):
    _fields_ = [("args", Args), ("ms_per_sweep", C.c_double)]


class ServerArgs(C.Structure):
    _fields_ = [("max_slots", C.c_int32), ("sweeps", C.c_int32), ("hbm_budget_bytes", C.c_int64),
                ("seed", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64), ("steps", C.c_int64),
                ("active", C.c_int32), ("pending", C.c_int32), ("max_active", C.c_int32),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("docs_per_s", C.c_double)]


class GenParams(C.Structure):
    _fields_ = [("V", C.c_int32), ("D", C.c_int32), ("K_true", C.c_int32), ("len_dist", C.c_int32),
                ("len_mean", C.c_double), ("len_sigma", C.c_double), ("pareto_shape", C.c_double),
                ("pareto_xm", C.c_double), ("len_min", C.c_int32), ("len_max", C.c_int32),
                ("zipf_s", C.c_double), ("phi_conc", C.c_double), ("theta_conc", C.c_double),
                ("seed", C.c_uint64), ("threads", C.c_int32), ("first_doc", C.c_int32)]


_lib = None


def lib() -> C.CDLL:
    """Loads libezlda.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m lda_sc20_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "ezlda_default_args": (None, [C.POINTER(Args)]),
        "ezlda_last_error": (C.c_char_p, [P]),
        "ezlda_create": (P, [C.POINTER(Args)]),
        "ezlda_destroy": (None, [P]),
        "ezlda_load_docs": (C.c_int, [P, C.c_int32, C.c_int32, P, P, P]),
        "ezlda_load_chunked": (C.c_int, [P, C.c_char_p, C.c_int32]),
        "ezlda_get_info": (C.c_int, [P, C.POINTER(Info)]),
        "ezlda_build_wt": (C.c_int, [P]),
        "ezlda_topk": (C.c_int, [P]),
        "ezlda_build_dt": (C.c_int, [P]),
        "ezlda_sample": (C.c_int, [P]),
        "ezlda_shortcut": (C.c_int, [P]),
        "ezlda_sample_head": (C.c_int, [P]),
        "ezlda_sample_tail": (C.c_int, [P]),
        "ezlda_commit_head": (C.c_int, [P]),
        "ezlda_wt_rowsum": (C.c_int, [P]),
        "ezlda_loglik": (C.c_int, [P, C.POINTER(C.c_double)]),
        "ezlda_loglik_noalloc": (C.c_int, [P, C.POINTER(C.c_double)]),
        "ezlda_sweep": (C.c_int, [P, C.c_int32]),
        "ezlda_sweep_graph": (C.c_int, [P, C.c_int32]),
        "ezlda_time_sweeps": (C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(C.c_double), P]),
        "ezlda_get_topics": (C.c_int, [P, P]),
        "ezlda_set_topics": (C.c_int, [P, P]),
        "ezlda_set_sweep": (C.c_int, [P, C.c_uint32]),
        "ezlda_get_nk": (C.c_int, [P, P]),
        "ezlda_get_wt_dense": (C.c_int, [P, P]),
        "ezlda_get_dt_raw": (C.c_int, [P, P, P]),
        "ezlda_get_top": (C.c_int, [P, P]),
        "ezlda_get_aux": (C.c_int, [P, P]),
        "ezlda_get_inv": (C.c_int, [P, P]),
        "ezlda_save_checkpoint": (C.c_int, [P, C.c_char_p]),
        "ezlda_load_checkpoint": (C.c_int, [P, C.c_char_p]),
        "ezlda_save_tl": (C.c_int, [P, C.c_char_p]),
        "ezlda_load_tl": (C.c_int, [P, C.c_char_p]),
        "ezlda_gen_corpus": (P, [C.POINTER(GenParams)]),
        "ezlda_corpus_num_tokens": (C.c_int64, [P]),
        "ezlda_corpus_num_docs": (C.c_int32, [P]),
        "ezlda_corpus_num_words": (C.c_int32, [P]),
        "ezlda_corpus_doc_off": (P, [P]),
        "ezlda_corpus_words": (P, [P]),
        "ezlda_corpus_topics": (P, [P]),
        "ezlda_corpus_free": (None, [P]),
        "ezlda_corpus_from_arrays": (P, [C.c_int32, C.c_int32, P, P, P]),
        "ezlda_read_docword": (P, [C.c_char_p]),
        "ezlda_write_chunked": (C.c_int, [C.c_char_p, P, P, C.c_int32, C.c_int32, C.c_uint64]),
        "ezlda_read_chunked": (P, [C.c_char_p, C.c_int32]),
        "ezlda_selftest_philox": (C.c_int, [P, P, P, C.c_int32]),
        "ezlda_selftest_copy": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
        "ezlda_selftest_gather": (C.c_int, [C.c_int64, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
        "ezlda_get_stats": (C.c_int, [P, P]),
        "ezlda_reset_stats": (C.c_int, [P]),
        "ezlda_set_layout": (C.c_int, [P, C.POINTER(Args)]),
        "ezlda_export_model": (C.c_int, [P, C.c_char_p]),
        "ezlda_autotune": (C.c_int, [P, C.c_int32, C.POINTER(Args), C.POINTER(C.c_double), P, C.c_int32,
                                     C.POINTER(C.c_int32)]),
        "ezlda_server_create": (P, [P, C.POINTER(ServerArgs)]),
        "ezlda_server_destroy": (None, [P]),
        "ezlda_server_last_error": (C.c_char_p, [P]),
        "ezlda_server_submit": (C.c_int, [P, C.c_uint32, P, C.c_int32]),
        "ezlda_server_submit_many": (C.c_int, [P, C.c_int32, P, P, P]),
        "ezlda_server_wait": (C.c_int, [P, C.c_uint32, P, P, C.POINTER(C.c_double), C.c_int32]),
        "ezlda_server_drain": (C.c_int, [P]),
        "ezlda_server_get_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
        "ezlda_server_reset_stats": (C.c_int, [P]),
        "ezlda_foldin_plain": (C.c_int, [P, C.c_uint32, P, C.c_int32, P, P]),
    }
    for name, (res, argt) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = argt
    _lib = L
    return L


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def last_error(ctx=None) -> str:
    m = lib().ezlda_last_error(ctx)
    return m.decode() if m else ""


def default_args(**kw) -> Args:
    a = Args()
    lib().ezlda_default_args(C.byref(a))
    for k, v in kw.items():
        if not hasattr(a, k):
            raise TypeError(f"unknown ezlda argument {k}")
        setattr(a, k, v)
    return a


# ------------------------------------------------------------------ corpora
class Corpus:
    """A doc-major corpus: doc d holds words[doc_off[d]:doc_off[d+1]]."""

    def __init__(self, V: int, doc_off: np.ndarray, words: np.ndarray, topics: Optional[np.ndarray] = None):
        self.V = int(V)
        self.doc_off = np.ascontiguousarray(doc_off, dtype=np.int64)
        self.words = np.ascontiguousarray(words, dtype=np.int32)
        self.topics = None if topics is None else np.ascontiguousarray(topics, dtype=np.uint16)

    @property
    def D(self) -> int:
        return len(self.doc_off) - 1

    @property
    def N(self) -> int:
        return int(self.doc_off[-1])

    @staticmethod
    def _from_handle(h) -> "Corpus":
        L = lib()
        if not h:
            raise EzldaError(EIO, last_error())
        try:
            D = L.ezlda_corpus_num_docs(h)
            N = L.ezlda_corpus_num_tokens(h)
            V = L.ezlda_corpus_num_words(h)
            off = np.ctypeslib.as_array(C.cast(L.ezlda_corpus_doc_off(h), C.POINTER(C.c_int64)), (D + 1,)).copy()
            words = (np.ctypeslib.as_array(C.cast(L.ezlda_corpus_words(h), C.POINTER(C.c_int32)), (N,)).copy()
                     if N else np.zeros(0, np.int32))
            tp = L.ezlda_corpus_topics(h)
            topics = (np.ctypeslib.as_array(C.cast(tp, C.POINTER(C.c_uint16)), (N,)).copy() if tp and N else None)
        finally:
            L.ezlda_corpus_free(h)
        return Corpus(V, off, words, topics)

    def write_chunked(self, prefix: str, K: int, num_chunks: int, z0: Optional[np.ndarray] = None, seed: int = 0):
        L = lib()
        h = L.ezlda_corpus_from_arrays(self.V, self.D, _ptr(self.doc_off), _ptr(self.words), None)
        if not h:
            raise EzldaError(EINVAL, last_error())
        try:
            z = None if z0 is None else np.ascontiguousarray(z0, dtype=np.uint16)
            r = L.ezlda_write_chunked(prefix.encode(), h, _ptr(z), K, num_chunks, seed)
        finally:
            L.ezlda_corpus_free(h)
        if r:
            raise EzldaError(r, last_error())


def gen_corpus(V: int, D: int, K_true: int, len_dist: int = 0, len_mean: float = 100.0, len_sigma: float = 0.8,
               pareto_shape: float = 1.2, pareto_xm: float = 1.0, len_min: int = 1, len_max: int = 65535,
               zipf_s: float = 0.0, phi_conc: float = 0.01, theta_conc: float = 0.1, seed: int = 1001,
               threads: int = 0, first_doc: int = 0) -> Corpus:
    """Synthetic corpus from the LDA generative process (C++ generator)."""
    p = GenParams(V, D, K_true, len_dist, len_mean, len_sigma, pareto_shape, pareto_xm, len_min, len_max, zipf_s,
                  phi_conc, theta_conc, seed, threads, first_doc)
    return Corpus._from_handle(lib().ezlda_gen_corpus(C.byref(p)))


def read_docword(path: str) -> Corpus:
    return Corpus._from_handle(lib().ezlda_read_docword(path.encode()))


def read_chunked(prefix: str, num_chunks: int = 0) -> Corpus:
    return Corpus._from_handle(lib().ezlda_read_chunked(prefix.encode(), num_chunks))


# ------------------------------------------------------------------ context
class Context:
    """One ezlda_ctx: a corpus on the current HIP device (resident, or chunked
    through pinned host memory when num_chunks > 1)."""

    def __init__(self, args: Optional[Args] = None, **kw):
        self._L = lib()
        self.args = args if args is not None else default_args(**kw)
        self._h = self._L.ezlda_create(C.byref(self.args))
        if not self._h:
            raise EzldaError(EHIP, last_error())
        self.corpus: Optional[Corpus] = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.ezlda_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, r: int):
        if r:
            raise EzldaError(r, last_error(self._h))

    # loading
    def load(self, corpus: Corpus, z0: Optional[np.ndarray] = None):
        z = z0 if z0 is not None else corpus.topics
        z = None if z is None else np.ascontiguousarray(z, dtype=np.uint16)
        self._chk(self._L.ezlda_load_docs(self._h, corpus.V, corpus.D, _ptr(corpus.doc_off), _ptr(corpus.words),
                                          _ptr(z)))
        self.corpus = corpus

    def load_chunked(self, prefix: str, num_chunks: int = 0):
        self._chk(self._L.ezlda_load_chunked(self._h, prefix.encode(), num_chunks))

    def info(self) -> Info:
        i = Info()
        self._chk(self._L.ezlda_get_info(self._h, C.byref(i)))
        return i

    # stages
    def build_wt(self):
        self._chk(self._L.ezlda_build_wt(self._h))

    def topk(self):
        self._chk(self._L.ezlda_topk(self._h))

    def build_dt(self):
        self._chk(self._L.ezlda_build_dt(self._h))

    def sample(self):
        self._chk(self._L.ezlda_sample(self._h))

    # the reference's launcher-sized stages (any order of the first three,
    # then commit_head, equals sample())
    def shortcut(self):
        self._chk(self._L.ezlda_shortcut(self._h))

    def sample_head(self):
        self._chk(self._L.ezlda_sample_head(self._h))

    def sample_tail(self):
        self._chk(self._L.ezlda_sample_tail(self._h))

    def commit_head(self):
        self._chk(self._L.ezlda_commit_head(self._h))

    def wt_rowsum(self):
        self._chk(self._L.ezlda_wt_rowsum(self._h))

    def sweep(self, n: int = 1):
        self._chk(self._L.ezlda_sweep(self._h, n))

    def sweep_graph(self, n: int = 1):
        self._chk(self._L.ezlda_sweep_graph(self._h, n))

    def loglik(self) -> float:
        x = C.c_double()
        self._chk(self._L.ezlda_loglik(self._h, C.byref(x)))
        return x.value

    def loglik_noalloc(self) -> float:
        """ezlda_loglik_noalloc: the same value from a call that allocates no
        device memory (partials in the library's static device array)."""
        x = C.c_double()
        self._chk(self._L.ezlda_loglik_noalloc(self._h, C.byref(x)))
        return x.value

    def time_sweeps(self, n: int, use_graph: bool = False):
        tot = C.c_double()
        st = np.zeros(len(STAGES), np.float64)
        self._chk(self._L.ezlda_time_sweeps(self._h, n, int(use_graph), C.byref(tot), _ptr(st)))
        return tot.value, dict(zip(STAGES, st.tolist()))

    def stats(self):
        """(full three-bucket draws, DT entries they read, WT tail nnz, DT nnz)."""
        out = np.zeros(4, np.uint64)
        self._chk(self._L.ezlda_get_stats(self._h, _ptr(out)))
        return tuple(int(x) for x in out)

    def reset_stats(self):
        self._chk(self._L.ezlda_reset_stats(self._h))

    def export_model(self, directory: str):
        """Write model.txt, phi.f32 (V x K float32) and theta.txt into directory."""
        self._chk(self._L.ezlda_export_model(self._h, directory.encode()))

    # layout tuning (topics never depend on it)
    def set_layout(self, **kw):
        """Rebuild the layout with the given block_dim, grid_dim, token_segment,
        dense_words, dense_threshold, doc_tiles or run_cap."""
        a = Args.from_buffer_copy(self.args)
        for k, v in kw.items():
            if not hasattr(a, k):
                raise TypeError(f"unknown ezlda argument {k}")
            setattr(a, k, v)
        self._chk(self._L.ezlda_set_layout(self._h, C.byref(a)))
        self.args = a

    def autotune(self, sweeps_per_trial: int = 3, max_trials: int = 64):
        """Coordinate-descent layout tuning: (chosen Args, ms per sweep, trials)."""
        best = Args()
        ms = C.c_double()
        tr = (TuneTrial * max_trials)()
        nt = C.c_int32()
        self._chk(self._L.ezlda_autotune(self._h, sweeps_per_trial, C.byref(best), C.byref(ms), tr, max_trials,
                                         C.byref(nt)))
        self.args = best
        keys = ("dense_words", "run_cap", "doc_tiles", "block_dim", "token_segment", "grid_dim")
        trials = [dict({k: getattr(tr[i].args, k) for k in keys}, ms_per_sweep=tr[i].ms_per_sweep)
                  for i in range(min(nt.value, max_trials))]
        return best, ms.value, trials

    # state
    def _n(self) -> int:
        return int(self.info().num_tokens)

    def get_topics(self) -> np.ndarray:
        z = np.zeros(self._n(), np.uint16)
        self._chk(self._L.ezlda_get_topics(self._h, _ptr(z)))
        return z

    def set_topics(self, z: np.ndarray):
        z = np.ascontiguousarray(z, dtype=np.uint16)
        if z.ndim != 1 or len(z) != self._n():
            raise EzldaError(EINVAL, f"set_topics needs {self._n()} topics, got {z.size}")
        self._chk(self._L.ezlda_set_topics(self._h, _ptr(z)))

    def set_sweep(self, s: int):
        self._chk(self._L.ezlda_set_sweep(self._h, s))

    def get_nk(self) -> np.ndarray:
        out = np.zeros(self.args.K, np.int32)
        self._chk(self._L.ezlda_get_nk(self._h, _ptr(out)))
        return out

    def get_inv(self) -> np.ndarray:
        out = np.zeros(self.args.K, np.float32)
        self._chk(self._L.ezlda_get_inv(self._h, _ptr(out)))
        return out

    def get_wt_dense(self) -> np.ndarray:
        i = self.info()
        out = np.zeros((i.num_words, i.K), np.int32)
        self._chk(self._L.ezlda_get_wt_dense(self._h, _ptr(out)))
        return out

    def get_dt_raw(self):
        i = self.info()
        ent = np.zeros(max(int(i.dt_capacity), 1), np.uint32)
        nnz = np.zeros(max(i.num_docs, 1), np.int32)
        self._chk(self._L.ezlda_get_dt_raw(self._h, _ptr(ent), _ptr(nnz)))
        return ent[: int(i.dt_capacity)], nnz[: i.num_docs]

    def get_top(self) -> np.ndarray:
        i = self.info()
        out = np.zeros((i.num_words, 3), np.uint16)
        self._chk(self._L.ezlda_get_top(self._h, _ptr(out)))
        return out

    def get_aux(self) -> np.ndarray:
        out = np.zeros(self._n(), np.uint32)
        self._chk(self._L.ezlda_get_aux(self._h, _ptr(out)))
        return out

    def save_checkpoint(self, path: str):
        self._chk(self._L.ezlda_save_checkpoint(self._h, path.encode()))

    def load_checkpoint(self, path: str):
        self._chk(self._L.ezlda_load_checkpoint(self._h, path.encode()))

    def save_tl(self, prefix: str):
        """Rewrite the chunked folder's TL.txt files with the current topics
        and write prefix/ezlda_state.txt (seed, sweep, K)."""
        self._chk(self._L.ezlda_save_tl(self._h, prefix.encode()))

    def load_tl(self, prefix: str):
        """Load the chunked folder and resume at its saved sweep."""
        self._chk(self._L.ezlda_load_tl(self._h, prefix.encode()))


class Server:
    """Fold-in server over a trained Context (its model is frozen at creation)."""

    def __init__(self, ctx: "Context", max_slots: int = 1024, sweeps: int = 20, hbm_budget_bytes: int = 1 << 28,
                 seed: int = 7):
        self._L = lib()
        self.K = ctx.args.K
        self.sweeps = sweeps
        a = ServerArgs(max_slots, sweeps, hbm_budget_bytes, seed)
        self._h = self._L.ezlda_server_create(ctx._h, C.byref(a))
        if not self._h:
            raise EzldaError(EHIP, last_error())
        self._lens = {}  # request id -> tokens, to size each result buffer

    def _chk(self, r):
        if r:
            m = self._L.ezlda_server_last_error(self._h)
            raise EzldaError(r, m.decode() if m else "")

    def close(self):
        if getattr(self, "_h", None):
            self._L.ezlda_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, rid: int, words: np.ndarray):
        w = np.ascontiguousarray(words, dtype=np.int32)
        self._chk(self._L.ezlda_server_submit(self._h, rid, _ptr(w), len(w)))
        self._lens[int(rid)] = len(w)

    def submit_many(self, rids, doc_off, words):
        """Queues len(rids) documents (words[doc_off[d]:doc_off[d+1]]) in one
        call; all or none."""
        ids = np.ascontiguousarray(rids, dtype=np.uint32)
        off = np.ascontiguousarray(doc_off, dtype=np.int64)
        w = np.ascontiguousarray(words, dtype=np.int32)
        if len(off) != len(ids) + 1:
            raise EzldaError(EINVAL, "doc_off needs len(rids) + 1 entries")
        self._chk(self._L.ezlda_server_submit_many(self._h, len(ids), _ptr(ids), _ptr(off), _ptr(w)))
        for d, r in enumerate(ids.tolist()):
            self._lens[r] = int(off[d + 1] - off[d])

    def wait(self, rid: int, n: Optional[int] = None, timeout_ms: int = -1):
        """(z, topic counts, latency ms) of a finished request.  The topic
        buffer is sized from the length the request was submitted with through
        this object (the library copies that many topics); a request submitted
        elsewhere needs its length as n."""
        known = self._lens.get(int(rid))
        if known is None and n is None:
            raise EzldaError(EINVAL, f"length of request {rid} unknown: pass n")
        m = known if known is not None else int(n)
        z = np.zeros(max(m, 1), np.uint16)
        c = np.zeros(self.K, np.int32)
        lat = C.c_double()
        self._chk(self._L.ezlda_server_wait(self._h, rid, _ptr(z), _ptr(c), C.byref(lat), timeout_ms))
        self._lens.pop(int(rid), None)
        return z[:m], c, lat.value

    def drain(self):
        self._chk(self._L.ezlda_server_drain(self._h))

    def stats(self) -> ServerStats:
        st = ServerStats()
        self._chk(self._L.ezlda_server_get_stats(self._h, C.byref(st)))
        return st

    def reset_stats(self):
        self._chk(self._L.ezlda_server_reset_stats(self._h))

    def plain(self, rid: int, words: np.ndarray):
        """One document per kernel launch (all of its sweeps): (z, topic counts)."""
        w = np.ascontiguousarray(words, dtype=np.int32)
        z = np.zeros(max(len(w), 1), np.uint16)
        c = np.zeros(self.K, np.int32)
        self._chk(self._L.ezlda_foldin_plain(self._h, rid, _ptr(w), len(w), _ptr(z), _ptr(c)))
        return z[: len(w)], c


def selftest_philox(ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
    """Philox4x32-10 of each (counter[4], key[2]) row, computed on the GPU."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    k = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros_like(c)
    r = lib().ezlda_selftest_philox(_ptr(c), _ptr(k), _ptr(out), c.shape[0])
    if r:
        raise EzldaError(r, last_error())
    return out


def selftest_copy(nbytes: int, reps: int = 10) -> float:
    """Mean ms of a plain device-to-device copy kernel over nbytes."""
    ms = C.c_double()
    r = lib().ezlda_selftest_copy(nbytes, reps, C.byref(ms))
    if r:
        raise EzldaError(r, last_error())
    return ms.value


def selftest_gather(nbytes: int, n_reads: int, reps: int = 10) -> float:
    """Mean ms of n_reads scattered 64-byte block reads over an nbytes buffer
    (the sampler's row-block access pattern; 64 * n_reads bytes per launch)."""
    ms = C.c_double()
    r = lib().ezlda_selftest_gather(nbytes, n_reads, reps, C.byref(ms))
    if r:
        raise EzldaError(r, last_error())
    return ms.value


COMPACT_K = 8192


def dt_piece(K: int) -> int:
    """Largest count of one compact DT entry: 2^(16 - tb) - 1 with tb =
    max(10, ceil(log2 K)) topic bits (63 for K <= 1024, 31 for K <= 2048...)."""
    tb = 10
    while (1 << tb) < K:
        tb += 1
    return (1 << (16 - tb)) - 1


def dt_capacity(n: np.ndarray, K: int) -> np.ndarray:
    """Entries of a doc's DT row at most: min(n, K), plus one per full piece
    when K <= 8192 (rows then store counts in pieces of at most dt_piece(K))."""
    n = np.asarray(n, dtype=np.int64)
    m = np.minimum(n, K)
    return m if K > COMPACT_K else np.minimum(m + n // dt_piece(K), n)


def dt_offsets(doc_off: Sequence[int], K: int) -> np.ndarray:
    """Row offsets of the DT getters: sum over earlier docs of the capacity."""
    n = np.diff(np.asarray(doc_off, dtype=np.int64))
    return np.concatenate([[0], np.cumsum(dt_capacity(n, K))])
