"""ctypes binding of the CPU oracle (oracle/_build/liboracle*.so).

Test infrastructure only: the checker for parity tests, smoke() and the
bench.py CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_libs = {}


def build() -> None:
    """Compiles both oracle libraries with the committed Makefile."""
    r = subprocess.run(["make", "-s", "-B", "-C", HERE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)


def lib(omp: bool = False) -> C.CDLL:
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    P = C.c_void_p
    sig = {
        "orc_philox": (None, [P, P, P]),
        "orc_create": (P, [C.c_int, C.c_float, C.c_float, C.c_uint64, C.c_int, C.c_int, P, P, P, C.c_int]),
        "orc_destroy": (None, [P]),
        "orc_num_tokens": (C.c_int64, [P]),
        "orc_get_sweep": (C.c_uint32, [P]),
        "orc_full_draws": (C.c_int64, [P]),
        "orc_set_sweep": (None, [P, C.c_uint32]),
        "orc_build_wt": (None, [P]),
        "orc_topk": (None, [P]),
        "orc_build_dt": (None, [P]),
        "orc_sample": (None, [P]),
        "orc_sweep": (None, [P, C.c_int]),
        "orc_loglik": (C.c_double, [P]),
        "orc_get_topics": (None, [P, P]),
        "orc_get_topics_wordmajor": (None, [P, P]),
        "orc_set_topics": (None, [P, P]),
        "orc_get_nk": (None, [P, P]),
        "orc_get_wt_dense": (None, [P, P]),
        "orc_get_dt_dense": (None, [P, P]),
        "orc_get_dt_raw": (None, [P, P, P]),
        "orc_get_top": (None, [P, P]),
        "orc_get_aux": (None, [P, P]),
        "orc_get_inv": (None, [P, P]),
        "orc_get_layout": (None, [P, P, P, P, P]),
        "orc_uniform": (C.c_float, [C.c_uint64, C.c_uint64, C.c_uint32]),
        "orc_num_threads": (C.c_int, []),
        "orc_foldin": (None, [P, C.c_int, P, P, P, C.c_int, C.c_uint64, P, P]),
    }
    for n, (res, argt) in sig.items():
        f = getattr(L, n)
        f.restype = res
        f.argtypes = argt
    _libs[name] = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key, omp: bool = False) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib(omp).orc_philox(_p(c), _p(k), _p(o))
    return o


class Oracle:
    """One oracle state; mirrors the stages of the product context."""

    def __init__(self, K: int, alpha: float, beta: float, seed: int, V: int, doc_off: np.ndarray, words: np.ndarray,
                 z0: Optional[np.ndarray] = None, shortcut: int = 1, omp: bool = False):
        """shortcut: 0 off, 1 (or True) the reference's bound, 2 the GPU path's looser
        bound (n_d - n_dk1) phi_k2 -- every mode gives the same topics."""
        self.L = lib(omp)
        self.K, self.V = K, V
        self.doc_off = np.ascontiguousarray(doc_off, dtype=np.int64)
        self.words = np.ascontiguousarray(words, dtype=np.int32)
        self.D = len(self.doc_off) - 1
        z = None if z0 is None else np.ascontiguousarray(z0, dtype=np.uint16)
        self.h = self.L.orc_create(K, alpha, beta, seed, V, self.D, _p(self.doc_off), _p(self.words), _p(z),
                                   int(shortcut))
        self.N = int(self.L.orc_num_tokens(self.h))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def build_wt(self):
        self.L.orc_build_wt(self.h)

    def topk(self):
        self.L.orc_topk(self.h)

    def build_dt(self):
        self.L.orc_build_dt(self.h)

    def sample(self):
        self.L.orc_sample(self.h)

    def sweep(self, n: int = 1):
        self.L.orc_sweep(self.h, n)

    def loglik(self) -> float:
        return float(self.L.orc_loglik(self.h))

    def full_draws(self) -> int:
        """Tokens that took the full draw (not the shortcut), over all sweeps."""
        return int(self.L.orc_full_draws(self.h))

    @property
    def sweep_index(self) -> int:
        return int(self.L.orc_get_sweep(self.h))

    def set_sweep(self, s: int):
        self.L.orc_set_sweep(self.h, s)

    def get_topics(self) -> np.ndarray:
        z = np.zeros(self.N, np.uint16)
        self.L.orc_get_topics(self.h, _p(z))
        return z

    def set_topics(self, z):
        z = np.ascontiguousarray(z, dtype=np.uint16)
        self.L.orc_set_topics(self.h, _p(z))

    def get_topics_wordmajor(self) -> np.ndarray:
        z = np.zeros(self.N, np.uint16)
        self.L.orc_get_topics_wordmajor(self.h, _p(z))
        return z

    def get_nk(self) -> np.ndarray:
        o = np.zeros(self.K, np.int32)
        self.L.orc_get_nk(self.h, _p(o))
        return o

    def get_inv(self) -> np.ndarray:
        o = np.zeros(self.K, np.float32)
        self.L.orc_get_inv(self.h, _p(o))
        return o

    def get_wt_dense(self) -> np.ndarray:
        o = np.zeros((self.V, self.K), np.int32)
        self.L.orc_get_wt_dense(self.h, _p(o))
        return o

    def get_dt_dense(self) -> np.ndarray:
        o = np.zeros((self.D, self.K), np.int32)
        self.L.orc_get_dt_dense(self.h, _p(o))
        return o

    def get_dt_raw(self):
        n = np.diff(self.doc_off).astype(np.int64)
        m = np.minimum(n, self.K)
        # dt_cap() of the oracle: compact rows (K <= 8192) add one entry per full piece
        tb = max(10, int(np.ceil(np.log2(self.K)))) if self.K > 1 else 10
        piece = (1 << (16 - tb)) - 1
        cap = int((m if self.K > 8192 else np.minimum(m + n // piece, n)).sum())
        ent = np.zeros(max(cap, 1), np.uint32)
        nnz = np.zeros(max(self.D, 1), np.int32)
        self.L.orc_get_dt_raw(self.h, _p(ent), _p(nnz))
        return ent[:cap], nnz[: self.D]

    def get_top(self) -> np.ndarray:
        o = np.zeros((self.V, 3), np.uint16)
        self.L.orc_get_top(self.h, _p(o))
        return o

    def get_aux(self) -> np.ndarray:
        o = np.zeros(self.N, np.uint32)
        self.L.orc_get_aux(self.h, _p(o))
        return o

    def foldin(self, doc_off, words, rids, sweeps: int, seed: int):
        """Fold-in of new docs against the frozen current model: (z, counts[D, K])."""
        doc_off = np.ascontiguousarray(doc_off, dtype=np.int64)
        words = np.ascontiguousarray(words, dtype=np.int32)
        rids = np.ascontiguousarray(rids, dtype=np.uint32)
        nd = len(doc_off) - 1
        z = np.zeros(max(int(doc_off[-1]), 1), np.uint16)
        cnt = np.zeros((max(nd, 1), self.K), np.int32)
        self.L.orc_foldin(self.h, nd, _p(doc_off), _p(words), _p(rids), sweeps, seed, _p(z), _p(cnt))
        return z[: int(doc_off[-1])], cnt[:nd]

    def get_layout(self):
        wo = np.zeros(self.V + 1, np.int64)
        td = np.zeros(self.N, np.int32)
        dtk = np.zeros(self.N, np.int64)
        ro = np.zeros(self.V, np.int32)
        self.L.orc_get_layout(self.h, _p(wo), _p(td), _p(dtk), _p(ro))
        return wo, td, dtk, ro


def uniform(seed: int, tok: int, sweep: int) -> float:
    return float(lib().orc_uniform(seed, tok, sweep))
