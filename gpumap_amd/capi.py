"""ctypes view of the C ABI (include/gpumap_c.h).

The binding a non-pybind11 caller would write (see the reference-side stub in
the integration notes); the GPU parity tests and the bench drive the engine
through it.  ``Engine`` owns one session handle.
"""

from __future__ import annotations

import atexit
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "gpumap_amd", "libgpumap_amd.so")
if not os.path.exists(LIB_PATH):
    raise ImportError("libgpumap_amd.so is not built; run "
                      "`python -m gpumap_amd.build`")
HEADER = os.path.join(ROOT, "include", "gpumap_c.h")

u32, u64, i32, i64, f32 = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32,
                           ctypes.c_int64, ctypes.c_float)
P = ctypes.POINTER


class GmParams(ctypes.Structure):
    _fields_ = [("n_neighbors", u64), ("n_components", u64), ("min_dist", f32),
                ("negative_sample_rate", u64), ("max_epoch", u64), ("a", f32),
                ("b", f32), ("alpha", f32), ("verbose", i32), ("metric", i32),
                ("n_instances", u64), ("n_features", u64), ("knn_mode", i32)]


class GmResult(ctypes.Structure):
    _fields_ = [("n_insertion_ops", u32), ("n_update_ops", u32),
                ("n_embedding_ops", u32), ("insertion_time", f32),
                ("update_time", f32), ("embedding_time", f32),
                ("insertion_completed", i32), ("update_completed", i32),
                ("embedding_completed", i32), ("n_inserted_instances", u64),
                ("n_update_changed_edges", u64), ("embedding_pending_edges", u64),
                ("insertion_queue_size", u64), ("update_queue_size", u64),
                ("embedding_queue_size", u64)]


class GmView(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("rows", u64), ("cols", u64),
                ("row_stride", u64), ("dtype_code", i32), ("dtype_bits", i32),
                ("device_id", i32)]


class GmKernelStats(ctypes.Structure):
    _fields_ = [("launches", u64), ("device_ms", ctypes.c_double),
                ("work", u64 * 4)]


KSTAT_JOIN, KSTAT_SGD, KSTAT_INSERT, KSTAT_SGD_STREAM, KSTAT_GRAPH = (
    0, 1, 2, 3, 4)
KNN_NNDESCENT, KNN_EXACT_FP32, KNN_EXACT_BF16 = 0, 1, 2


def header_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gm_[a-z_]+)\s*\(", text)))


_lib = None


def lib():
    global _lib
    if _lib is None:
        # one HIP runtime per process: torch's, loaded first (see
        # gpumap_amd._native)
        try:
            import torch  # noqa: F401
        except ImportError:  # pragma: no cover - torch is part of the image
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.gm_default_params.argtypes = [P(GmParams)]
        L.gm_create.argtypes = [P(GmParams), P(vp)]
        L.gm_initialize.argtypes = [vp, u64, u64]
        L.gm_append_data.argtypes = [vp, P(f32), u64, u64]
        L.gm_run.argtypes = [vp, u32, u32, u32, P(GmResult)]
        L.gm_run_batch.argtypes = [P(vp), i32, P(u32), P(GmResult)]
        L.gm_close.argtypes = [vp]
        L.gm_destroy.argtypes = [vp]
        L.gm_destroy.restype = None
        L.gm_get_view.argtypes = [vp, i32, P(GmView)]
        L.gm_export_dlpack.argtypes = [vp, i32, P(vp)]
        L.gm_copy_knn.argtypes = [vp, P(u32), P(f32)]
        L.gm_copy_embedding.argtypes = [vp, P(f32)]
        L.gm_set_embedding.argtypes = [vp, P(f32)]
        L.gm_last_error.restype = ctypes.c_char_p
        L.gm_debug_pool.argtypes = [vp, P(u64)]
        L.gm_debug_fuzzy.argtypes = [vp, P(u32), u64, P(i32), P(i32), P(f32),
                                     P(ctypes.c_double), P(ctypes.c_double),
                                     P(i64)]
        L.gm_debug_sgd.argtypes = [P(f32), u64, i32, P(i32), P(i32), P(f32),
                                   i64, i32, i32, f32, f32, f32, f32, u64, i32]
        L.gm_debug_philox.argtypes = [P(u32), P(u32), P(u32)]
        L.gm_append_data_device.argtypes = [vp, vp, u64, u64]
        L.gm_set_profiling.argtypes = [vp, i32]
        L.gm_get_kernel_stats.argtypes = [vp, i32, P(GmKernelStats)]
        L.gm_reset_kernel_stats.argtypes = [vp]
        L.gm_sgd_noalloc.argtypes = [vp, u64, i32, vp, vp, vp, i64, i32, i32,
                                     f32, f32, f32, f32, u64, i32]
        L.gm_release.argtypes = [vp]
        L.gm_is_released.argtypes = [vp, P(i32)]
        L.gm_shutdown.argtypes = []
        L.gm_shutdown.restype = None
        L.gm_debug_e4m3.argtypes = [P(f32), P(ctypes.c_uint8), u64, i32]
        L.gm_screen_shapes.argtypes = [ctypes.c_char_p, u64]
        L.gm_screen_shapes.restype = i64
        # sessions still alive at exit are freed while the HIP runtime is up
        # (registered after torch's hooks, so it runs before them)
        atexit.register(L.gm_shutdown)
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(P(ct))


def check(rc):
    if rc != 0:
        raise RuntimeError(f"gm status {rc}: {lib().gm_last_error().decode()}")


def screen_shapes():
    """The BF16 screen's launch shapes tuned so far in this process (see
    gm_screen_shapes), as a list of dicts."""
    import json

    need = lib().gm_screen_shapes(None, 0)
    if need < 0:
        check(int(need))
    buf = ctypes.create_string_buffer(int(need) + 1)
    got = lib().gm_screen_shapes(buf, len(buf))
    if got < 0:
        check(int(got))
    return json.loads(buf.value.decode())


def e4m3(x, on_device=False):
    """The FP8 screen's float -> E4M3 codes (gm_debug_e4m3)."""
    x = np.ascontiguousarray(x, np.float32).ravel()
    out = np.zeros(x.size, np.uint8)
    check(lib().gm_debug_e4m3(_p(x, f32), _p(out, ctypes.c_uint8), x.size,
                              1 if on_device else 0))
    return out


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    check(lib().gm_debug_philox(_p(c, u32), _p(k, u32), _p(out, u32)))
    return out


def sgd(emb, head, tail, w, n_epochs, nsr, a, b, alpha0=1.0, gamma=1.0, seed=42,
        per_epoch=False, stream=False, pull=False):
    """gm_debug_sgd: all epochs in one launch of the kernel that fits, one
    launch per epoch (per_epoch), the streaming kernel (stream) or the
    row-owner kernel (pull)."""
    emb = np.ascontiguousarray(emb, np.float32).copy()
    head = np.ascontiguousarray(head, np.int32)
    tail = np.ascontiguousarray(tail, np.int32)
    w = np.ascontiguousarray(w, np.float32)
    check(lib().gm_debug_sgd(_p(emb, f32), emb.shape[0], emb.shape[1],
                             _p(head, i32), _p(tail, i32), _p(w, f32), len(w),
                             n_epochs, nsr, a, b, alpha0, gamma, seed,
                             3 if pull else 2 if stream else
                             1 if per_epoch else 0))
    return emb


def sgd_noalloc_device(emb, head, tail, w, n_epochs, nsr, a, b, alpha0=1.0,
                       gamma=1.0, seed=42, per_epoch=False, stream=False):
    """gm_sgd_noalloc on torch device tensors (emb float32 (n, c), updated in
    place; head, tail int32 and w float32 of one length)."""
    import torch

    for t in (emb, head, tail, w):
        if not (t.is_cuda and t.is_contiguous()):
            raise ValueError("device tensors, contiguous, expected")
    assert emb.dtype == torch.float32 and w.dtype == torch.float32
    assert head.dtype == torch.int32 and tail.dtype == torch.int32
    torch.cuda.current_stream().synchronize()
    check(lib().gm_sgd_noalloc(emb.data_ptr(), emb.shape[0], emb.shape[1],
                               head.data_ptr(), tail.data_ptr(), w.data_ptr(),
                               w.numel(), n_epochs, nsr, a, b, alpha0, gamma,
                               seed, 2 if stream else 1 if per_epoch else 0))


def sgd_noalloc(emb, head, tail, w, n_epochs, nsr, a, b, alpha0=1.0, gamma=1.0,
                seed=42, per_epoch=False, stream=False):
    """sgd_noalloc_device on host arrays (copied to and from the device)."""
    import torch

    dev = torch.device("cuda")
    e = torch.from_numpy(np.ascontiguousarray(emb, np.float32).copy()).to(dev)
    h = torch.from_numpy(np.ascontiguousarray(head, np.int32)).to(dev)
    t = torch.from_numpy(np.ascontiguousarray(tail, np.int32)).to(dev)
    ww = torch.from_numpy(np.ascontiguousarray(w, np.float32)).to(dev)
    sgd_noalloc_device(e, h, t, ww, n_epochs, nsr, a, b, alpha0, gamma, seed,
                       per_epoch, stream)
    return e.cpu().numpy()


class Engine:
    def __init__(self, n_instances, n_features, n_neighbors=15, n_components=2,
                 metric=0, a=1.577, b=0.895, alpha=1.0, max_epoch=200,
                 negative_sample_rate=5, knn_mode=None):
        p = GmParams()
        lib().gm_default_params(ctypes.byref(p))
        p.n_instances, p.n_features = n_instances, n_features
        p.n_neighbors, p.n_components, p.metric = n_neighbors, n_components, metric
        p.a, p.b, p.alpha, p.max_epoch = a, b, alpha, max_epoch
        p.negative_sample_rate = negative_sample_rate
        if knn_mode is not None:
            p.knn_mode = knn_mode
        h = ctypes.c_void_p()
        check(lib().gm_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.k, self.c, self.n = n_neighbors, n_components, 0

    def close(self):
        if self.h:
            lib().gm_destroy(self.h)
            self.h = None

    def release(self):
        """gm_release: free the device memory now (no exports are made
        through this class)."""
        check(lib().gm_release(self.h))

    @property
    def released(self) -> bool:
        r = i32(0)
        check(lib().gm_is_released(self.h, ctypes.byref(r)))
        return bool(r.value)

    def __del__(self):
        self.close()

    def append(self, x):
        x = np.ascontiguousarray(x, np.float32)
        check(lib().gm_append_data(self.h, _p(x, f32), x.shape[0], x.shape[1]))

    def append_device(self, ptr: int, rows: int, cols: int):
        """Append rows already in device memory (a raw device pointer)."""
        check(lib().gm_append_data_device(self.h, ctypes.c_void_p(ptr), rows,
                                          cols))

    def set_profiling(self, on: bool = True):
        check(lib().gm_set_profiling(self.h, 1 if on else 0))

    def kernel_stats(self, which: int) -> GmKernelStats:
        st = GmKernelStats()
        check(lib().gm_get_kernel_stats(self.h, which, ctypes.byref(st)))
        return st

    def reset_kernel_stats(self):
        check(lib().gm_reset_kernel_stats(self.h))

    def run(self, i, u, e):
        r = GmResult()
        check(lib().gm_run(self.h, i, u, e, ctypes.byref(r)))
        self.n = r.n_inserted_instances
        return r

    @staticmethod
    def run_batch(engines, ops):
        """One cycle of several sessions together (gm_run_batch); `ops` holds
        one (insertion, update, embedding) triple per engine."""
        cnt = len(engines)
        hs = (ctypes.c_void_p * cnt)(
            *[e.h.value if isinstance(e.h, ctypes.c_void_p) else e.h
              for e in engines])
        flat = np.ascontiguousarray(np.asarray(ops, np.uint32).reshape(cnt, 3))
        rs = (GmResult * cnt)()
        check(lib().gm_run_batch(hs, cnt, _p(flat, u32), rs))
        for e, r in zip(engines, rs):
            e.n = r.n_inserted_instances
        return list(rs)

    def knn(self):
        idx = np.zeros((self.n, self.k), np.uint32)
        dist = np.zeros((self.n, self.k), np.float32)
        check(lib().gm_copy_knn(self.h, _p(idx, u32), _p(dist, f32)))
        return idx, dist

    def embedding(self):
        e = np.zeros((self.n, self.c), np.float32)
        check(lib().gm_copy_embedding(self.h, _p(e, f32)))
        return e

    def set_embedding(self, e):
        e = np.ascontiguousarray(e, np.float32)
        check(lib().gm_set_embedding(self.h, _p(e, f32)))

    def pool(self):
        p = np.zeros((self.n, 32), np.uint64)
        check(lib().gm_debug_pool(self.h, _p(p, u64)))
        return p

    def fuzzy(self, batch):
        batch = np.ascontiguousarray(batch, np.uint32)
        cap = 2 * len(batch) * self.k + 1
        head = np.zeros(cap, np.int32)
        tail = np.zeros(cap, np.int32)
        w = np.zeros(cap, np.float32)
        sig = np.zeros(self.n, np.float64)
        rho = np.zeros(self.n, np.float64)
        nnz = i64(0)
        check(lib().gm_debug_fuzzy(self.h, _p(batch, u32), len(batch),
                                   _p(head, i32), _p(tail, i32), _p(w, f32),
                                   _p(sig, ctypes.c_double),
                                   _p(rho, ctypes.c_double), ctypes.byref(nnz)))
        m = nnz.value
        return head[:m], tail[:m], w[:m], sig, rho
