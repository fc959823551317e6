"""ctypes front-end of the CPU oracle (oracle/oracle.cpp) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker or the timed CPU baseline.  The
product (gpumap_amd) never imports it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile (g++, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class OrParams(ctypes.Structure):
    _fields_ = [
        ("n_neighbors", ctypes.c_uint32),
        ("n_components", ctypes.c_uint32),
        ("negative_sample_rate", ctypes.c_uint32),
        ("max_epoch", ctypes.c_uint32),
        ("min_dist", ctypes.c_float),
        ("a", ctypes.c_float),
        ("b", ctypes.c_float),
        ("alpha", ctypes.c_float),
        ("metric", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("n_instances", ctypes.c_uint64),
        ("n_features", ctypes.c_uint64),
        ("n_threads", ctypes.c_int32),
        ("max_iter", ctypes.c_int32),
        ("knn_mode", ctypes.c_int32),
    ]


class OrResult(ctypes.Structure):
    _fields_ = [
        ("n_insertion_ops", ctypes.c_uint32),
        ("n_update_ops", ctypes.c_uint32),
        ("n_embedding_ops", ctypes.c_uint32),
        ("insertion_time", ctypes.c_float),
        ("update_time", ctypes.c_float),
        ("embedding_time", ctypes.c_float),
        ("insertion_completed", ctypes.c_int32),
        ("update_completed", ctypes.c_int32),
        ("embedding_completed", ctypes.c_int32),
        ("n_inserted_instances", ctypes.c_uint64),
        ("n_update_changed_edges", ctypes.c_uint64),
        ("embedding_pending_edges", ctypes.c_uint64),
        ("insertion_queue_size", ctypes.c_uint64),
        ("update_queue_size", ctypes.c_uint64),
        ("embedding_queue_size", ctypes.c_uint64),
    ]


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, u32, u64, i32, f32, i64 = (ctypes.c_void_p, ctypes.c_uint32,
                                       ctypes.c_uint64, ctypes.c_int32,
                                       ctypes.c_float, ctypes.c_int64)
        P = ctypes.POINTER
        L.or_create.restype = vp
        L.or_create.argtypes = [P(OrParams)]
        L.or_destroy.argtypes = [vp]
        L.or_append.argtypes = [vp, P(ctypes.c_float), u64, u64]
        L.or_append.restype = i32
        L.or_run.argtypes = [vp, u32, u32, u32, P(OrResult)]
        L.or_get_knn.argtypes = [vp, P(ctypes.c_uint32), P(ctypes.c_float)]
        L.or_get_embedding.argtypes = [vp, P(ctypes.c_float)]
        L.or_set_embedding.argtypes = [vp, P(ctypes.c_float)]
        L.or_get_pool.argtypes = [vp, P(ctypes.c_uint64)]
        L.or_philox4x32.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32),
                                    P(ctypes.c_uint32)]
        L.or_canon_dist.argtypes = [P(ctypes.c_float), P(ctypes.c_float), i32]
        L.or_canon_dist.restype = f32
        L.or_fuzzy.argtypes = [P(ctypes.c_uint32), P(ctypes.c_float), u64, i32,
                               i32, P(ctypes.c_uint32), u64, i32, P(i32), P(i32),
                               P(ctypes.c_float), P(ctypes.c_double),
                               P(ctypes.c_double)]
        L.or_fuzzy.restype = i64
        L.or_sgd.argtypes = [P(ctypes.c_float), u64, i32, P(i32), P(i32),
                             P(ctypes.c_float), i64, i32, i32, f32, f32, f32,
                             f32, u64, i32]
        L.or_sgd_trace.argtypes = [P(ctypes.c_float), u64, i32, P(i32),
                                   P(i32), P(ctypes.c_float), i64, i32, i32,
                                   f32, f32, f32, f32, u64, vp, i64,
                                   P(ctypes.c_float)]
        L.or_sgd_trace.restype = i64
        L.or_rounding_factor.argtypes = [P(i32), i64, u64, f32]
        L.or_rounding_factor.restype = f32
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def philox4x32(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32(_ptr(c, ctypes.c_uint32), _ptr(k, ctypes.c_uint32),
                        _ptr(out, ctypes.c_uint32))
    return out


def canon_dist(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return float(lib().or_canon_dist(_ptr(a, ctypes.c_float),
                                     _ptr(b, ctypes.c_float), a.shape[0]))


def fuzzy(idx, dist, k, euclidean, batch, n_epochs=200):
    """Fuzzy union for the batch rows on a given public graph (n x k)."""
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    dist = np.ascontiguousarray(dist, dtype=np.float32)
    batch = np.ascontiguousarray(batch, dtype=np.uint32)
    n = idx.shape[0]
    cap = 2 * len(batch) * k + 1
    head = np.zeros(cap, np.int32)
    tail = np.zeros(cap, np.int32)
    w = np.zeros(cap, np.float32)
    sig = np.zeros(n, np.float64)
    rho = np.zeros(n, np.float64)
    nnz = lib().or_fuzzy(_ptr(idx, ctypes.c_uint32), _ptr(dist, ctypes.c_float),
                         n, k, int(euclidean), _ptr(batch, ctypes.c_uint32),
                         len(batch), n_epochs, _ptr(head, ctypes.c_int32),
                         _ptr(tail, ctypes.c_int32), _ptr(w, ctypes.c_float),
                         _ptr(sig, ctypes.c_double), _ptr(rho, ctypes.c_double))
    return head[:nnz], tail[:nnz], w[:nnz], sig, rho


def sgd(emb, head, tail, w, n_epochs, nsr, a, b, alpha0=1.0, gamma=1.0,
        seed=42, n_threads=1):
    emb = np.ascontiguousarray(emb, dtype=np.float32).copy()
    head = np.ascontiguousarray(head, dtype=np.int32)
    tail = np.ascontiguousarray(tail, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    lib().or_sgd(_ptr(emb, ctypes.c_float), emb.shape[0], emb.shape[1],
                 _ptr(head, ctypes.c_int32), _ptr(tail, ctypes.c_int32),
                 _ptr(w, ctypes.c_float), len(w), n_epochs, nsr, a, b, alpha0,
                 gamma, seed, n_threads)
    return emb


# or_sgd_event (oracle.h): one applied gradient step
SGD_EVENT = np.dtype([("epoch", np.int32), ("edge", np.int32),
                      ("kind", np.int32), ("other", np.int32),
                      ("alpha", np.float32), ("d2", np.float32),
                      ("coef", np.float32), ("y", np.float32, 4),
                      ("o", np.float32, 4), ("g", np.float32, 4)])
SGD_POSITIVE, SGD_NEGATIVE, SGD_NEGATIVE_SKIPPED = 0, 1, 2


def sgd_trace(emb, head, tail, w, n_epochs, nsr, a, b, alpha0=1.0, gamma=1.0,
              seed=42, cap=1 << 20):
    """Serial SGD with every gradient step recorded: (embedding, events,
    rounding grid)."""
    emb = np.ascontiguousarray(emb, dtype=np.float32).copy()
    head = np.ascontiguousarray(head, dtype=np.int32)
    tail = np.ascontiguousarray(tail, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    ev = np.zeros(cap, SGD_EVENT)
    r = ctypes.c_float(0.0)
    m = lib().or_sgd_trace(_ptr(emb, ctypes.c_float), emb.shape[0],
                           emb.shape[1], _ptr(head, ctypes.c_int32),
                           _ptr(tail, ctypes.c_int32), _ptr(w, ctypes.c_float),
                           len(w), n_epochs, nsr, a, b, alpha0, gamma, seed,
                           ev.ctypes.data, cap, ctypes.byref(r))
    if m > cap:
        raise ValueError(f"{m} SGD steps exceed the trace capacity {cap}")
    return emb, ev[:m], float(r.value)


def rounding_factor(head, n, alpha):
    head = np.ascontiguousarray(head, dtype=np.int32)
    return float(lib().or_rounding_factor(_ptr(head, ctypes.c_int32),
                                          len(head), n, alpha))


class Oracle:
    """The oracle engine with the reference's run(i, u, e) contract."""

    def __init__(self, n_instances, n_features, n_neighbors=15, n_components=2,
                 metric=0, a=1.577, b=0.895, alpha=1.0, max_epoch=200,
                 negative_sample_rate=5, n_threads=1, max_iter=10,
                 knn_mode=0):
        p = OrParams(n_neighbors=n_neighbors, n_components=n_components,
                     negative_sample_rate=negative_sample_rate,
                     max_epoch=max_epoch, min_dist=0.1, a=a, b=b, alpha=alpha,
                     metric=metric, verbose=0, n_instances=n_instances,
                     n_features=n_features, n_threads=n_threads,
                     max_iter=max_iter, knn_mode=knn_mode)
        self._h = lib().or_create(ctypes.byref(p))
        if not self._h:
            raise ValueError("invalid oracle parameters")
        self.k = n_neighbors
        self.c = n_components
        self.n = 0

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_destroy(self._h)
            self._h = None

    def append(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        rc = lib().or_append(self._h, _ptr(x, ctypes.c_float), x.shape[0],
                             x.shape[1])
        if rc != 0:
            raise RuntimeError(f"or_append failed ({rc})")

    def run(self, i, u, e) -> OrResult:
        r = OrResult()
        lib().or_run(self._h, i, u, e, ctypes.byref(r))
        self.n = r.n_inserted_instances
        return r

    def knn(self):
        idx = np.zeros((self.n, self.k), np.uint32)
        dist = np.zeros((self.n, self.k), np.float32)
        lib().or_get_knn(self._h, _ptr(idx, ctypes.c_uint32),
                         _ptr(dist, ctypes.c_float))
        return idx, dist

    def embedding(self):
        e = np.zeros((self.n, self.c), np.float32)
        lib().or_get_embedding(self._h, _ptr(e, ctypes.c_float))
        return e

    def set_embedding(self, e):
        e = np.ascontiguousarray(e, dtype=np.float32)
        assert e.shape == (self.n, self.c)
        lib().or_set_embedding(self._h, _ptr(e, ctypes.c_float))

    def pool(self):
        p = np.zeros((self.n, 32), np.uint64)
        lib().or_get_pool(self._h, _ptr(p, ctypes.c_uint64))
        return p
