"""Serving form of the progressive engine: many sessions on one device.

No counterpart in the reference, where one ``GPUMAP`` object runs at a time
(python/gpumap/gpumap.py:33-162).  A :class:`Server` admits sessions while
their device footprint fits its budget and runs their requests by continuous
batching: a loop thread takes every pending request (at most one per session)
and runs them as one batched cycle (``_core.run_batch``, the C ABI's
``gm_run_batch``).  In a batched cycle the sessions' insertion work overlaps
on the device and their per-cycle SGDs share one cooperative launch; a
request that arrives while a cycle runs joins the next one, and a session
that closes frees its footprint for the sessions waiting to be admitted.

Each session's results are those of ``GPUMAP.run`` on its own (the same
graph bit for bit, the same op counts and queue states); only the phase
times are the batch's, and the Hogwild SGD makes embeddings agree only in
distribution, as two runs alone do.
"""

from __future__ import annotations

import threading
import time
from concurrent.futures import Future

from .gpumap import GPUMAP
from ._native import _core

_POOL_W = 32            # pool slots per row
_CAND = 256             # NN-Descent candidates per inserted row
_SURVIVOR_CAP = 2048    # BF16 screening survivors per batch row (x8 above
                        # 2^18 rows)
_MAX_SPLIT = 128        # column splits of the exact insertion's tiles
_LB_STORE_MAX = 8 << 30
_SGD_SLOTS = 4 + 512    # cooperative SGD barrier words and wave slots
_GRANULE = 2 << 20      # device allocations are reserved in 2 MiB granules


def _allocations(n: int, d: int, k: int, c: int, knn_mode: int,
                 b: int) -> list[int]:
    """Bytes of each device allocation of a session of n rows (Engine::
    allocate) and of its insertion scratch for batches of up to b rows."""
    dp = (d + 3) // 4 * 4
    nk = n * k
    sizes = [n * dp * 4,                      # rows
             n * _POOL_W * 8,                 # pools
             nk * 4, nk * 4,                  # public ids, distances
             n * c * 4, n * c * 4,            # embedding, its snapshot
             n * (4 if c == 3 else c) * 8,    # the SGD's row records
             n, n, n * 4, n * 4, n * 4,       # flags, marks, list, pmax, deg
             (max(2 * nk, n) // 1024 + 2) * 4,
             n * _POOL_W * 4, n * _POOL_W * 4,        # sampled lists
             n * _POOL_W * 8, n * _POOL_W * 8]        # reverse lists
    sizes += [n * 4] * 6                      # masks, owners, rows, counts
    sizes += [n * 8, n * 8]                   # sigma, rho
    sizes += [nk * 4, nk]                     # weights, weight flags
    sizes += [2 * nk * 4] * 6                 # COO and SGD schedule
    sizes += [n * 4]                          # embedding batch
    sizes += [256 * 8, 256 * 8, 8, 64, 8, 128, 4, 4, _SGD_SLOTS * 4]
    # the exact tiles' per-(row, split) top lists (FP32 and BF16 modes) or
    # the NN-Descent candidates
    if knn_mode in (0, 4):
        sizes += [b * _POOL_W * 4, b * 4, b * _CAND * 4, b * _CAND * 4, b * 4]
    else:
        sizes += [b * _MAX_SPLIT * _POOL_W * 8]
    if knn_mode in (2, 3):
        # BF16: 4 bytes per padded feature (hi + lo); FP8: 2, and the
        # rows' inverse scales
        db = (d + 15) // 16 * 16 if knn_mode == 2 else (d + 63) // 64 * 64
        rows = (n + 127) // 128 * 128
        cap = (_SURVIVOR_CAP if n <= 1 << 18 and knn_mode == 2
               else 8 * _SURVIVOR_CAP)
        per = 4 if knn_mode == 2 else 2
        sizes += [rows * db * per, rows * 4, 16]          # split rows, norms
        if knn_mode == 3:
            sizes += [rows * 4]                           # inverse scales
        sizes += [(b + 127) // 128 * 128 * db * 4, b * 4,  # batch operands
                  b * cap * 4, 2 * b * 4, b * cap * 8]
        if n * b * 4 <= _LB_STORE_MAX:
            sizes += [n * b * 4]                          # lower-bound store
    if knn_mode == 4:
        # the local join's hi/lo operand store (32-feature chunks), norms
        sizes += [n * ((d + 31) // 32 * 32) * 4, n * 4]
    return sizes


def session_footprint(n_instances: int, n_features: int, n_neighbors: int = 15,
                      n_components: int = 2, knn_mode: int = 2,
                      max_batch_rows: int = 1000) -> int:
    """Device bytes a session of this shape can hold: each of the engine's
    allocations (its session arrays and the scratch of insertions of up to
    `max_batch_rows` rows) rounded up to the allocator's 2 MiB granule."""
    g = _GRANULE
    return int(sum((s + g - 1) // g * g
                   for s in _allocations(n_instances, n_features, n_neighbors,
                                         n_components, knn_mode,
                                         max_batch_rows)))


def _device_free_bytes() -> int | None:
    try:
        import torch

        if torch.cuda.is_available():
            return int(torch.cuda.mem_get_info()[0])
    except Exception:  # pragma: no cover - torch is part of the image
        pass
    return None


class Admission:
    """Footprint bookkeeping of a server: `acquire` waits until a footprint
    fits the budget (and a session slot is free), `release` frees it at once
    and `release_when` once a probe reports that the session's device
    memory has really been freed."""

    def __init__(self, budget_bytes: int, max_sessions: int,
                 poll_interval: float = 0.02):
        self.budget = int(budget_bytes)
        self.max_sessions = int(max_sessions)
        self.used = 0
        self.sessions = 0
        self.poll_interval = float(poll_interval)
        self._freeing: list[tuple[int, object]] = []
        self._cv = threading.Condition()

    def _sweep(self) -> None:
        # caller holds the lock: footprints whose memory is gone are returned
        keep = []
        for nbytes, probe in self._freeing:
            if probe():
                self.used -= nbytes
                self.sessions -= 1
            else:
                keep.append((nbytes, probe))
        self._freeing = keep

    def acquire(self, nbytes: int, timeout: float | None = None) -> None:
        if nbytes > self.budget:
            raise MemoryError(f"session needs {nbytes} bytes, more than the "
                              f"server's budget of {self.budget}")
        deadline = None if timeout is None else time.monotonic() + timeout
        with self._cv:
            while True:
                self._sweep()
                if (self.used + nbytes <= self.budget
                        and self.sessions < self.max_sessions):
                    break
                left = None if deadline is None else deadline - time.monotonic()
                if left is not None and left <= 0:
                    raise TimeoutError("no room for the session before the "
                                       "admission timeout")
                if self._freeing:
                    # memory held by a closed session's exported tensors is
                    # freed without a notification: poll for it
                    left = (self.poll_interval if left is None
                            else min(left, self.poll_interval))
                self._cv.wait(left)
            self.used += nbytes
            self.sessions += 1

    def release(self, nbytes: int) -> None:
        with self._cv:
            self.used -= nbytes
            self.sessions -= 1
            self._cv.notify_all()

    def release_when(self, nbytes: int, probe) -> None:
        """Return `nbytes` once `probe()` is true (checked now and by every
        waiting `acquire`)."""
        with self._cv:
            self._freeing.append((nbytes, probe))
            self._sweep()
            self._cv.notify_all()

    def pending_release(self) -> int:
        """Bytes of closed sessions whose device memory is still held."""
        with self._cv:
            self._sweep()
            return sum(nb for nb, _ in self._freeing)


class ServedSession:
    """A session admitted by a :class:`Server`: the ``GPUMAP`` surface, with
    run() and run_with_n_ops() served in batched cycles.  The getters return
    views of the session's device arrays; read them between the session's
    own requests (a cycle in flight writes them)."""

    def __init__(self, server: "Server", model: GPUMAP, footprint: int):
        self._server = server
        self._model = model
        self.footprint = footprint
        self._closed = False
        # (insertion, update, embedding) ops of each cycle the session ran,
        # in order: its run() calls replayed by run_with_n_ops give the same
        # results
        self.ops_history: list[tuple[int, int, int]] = []

    @property
    def params(self):
        return self._model.params

    @property
    def strategy(self):
        return self._model.strategy

    @property
    def n_inserted_instances(self) -> int:
        return self._model.n_inserted_instances

    def submit(self, target_latency: float | None = None,
               n_ops: tuple[int, int, int] | None = None) -> Future:
        """Queue one cycle: by the strategy for `target_latency`, or the
        given (insertion, update, embedding) op counts."""
        if self._closed:
            raise RuntimeError("session is closed")
        return self._server._submit(self, target_latency, n_ops)

    def run(self, target_latency: float = 10.0):
        return self.submit(target_latency=target_latency).result()

    def run_with_n_ops(self, n_insertion_ops: int, n_update_ops: int,
                       n_embedding_ops: int):
        return self.submit(n_ops=(n_insertion_ops, n_update_ops,
                                  n_embedding_ops)).result()

    def get_embedding(self):
        return self._model.get_embedding()

    def get_neighbor_indices(self):
        return self._model.get_neighbor_indices()

    def get_neighbor_dists(self):
        return self._model.get_neighbor_dists()

    def close(self):
        if not self._closed:
            self._closed = True
            self._server._close_session(self)


class Server:
    """Continuous batching of many sessions on the current device.

    `hbm_budget_bytes` defaults to 90% of the device memory free when the
    server starts; `max_sessions` bounds the sessions admitted at once and
    `max_batch_rows` the rows a session inserts per cycle when its strategy
    allocates the ops (run(target_latency)), which bounds both the footprint
    estimate and the time one session's insertion adds to a shared cycle;
    explicit op counts (run_with_n_ops) are run as given.  A
    cycle starts once every session of the previous cycle that is still open
    has its next request pending, or `batch_window` seconds after the first
    pending request, whichever comes first (0 starts it at once with
    whatever is pending): clients that answer each result with their next
    request share every cycle, and idle sessions delay nothing."""

    def __init__(self, hbm_budget_bytes: int | None = None,
                 max_sessions: int = 64, max_batch_rows: int = 1000,
                 batch_window: float = 0.001):
        if hbm_budget_bytes is None:
            free = _device_free_bytes()
            if free is None:
                raise RuntimeError("no device to serve on")
            hbm_budget_bytes = int(0.9 * free)
        self.admission = Admission(hbm_budget_bytes, max_sessions)
        self.max_batch_rows = max_batch_rows
        self.batch_window = float(batch_window)
        self._last: list[ServedSession] = []
        self._cv = threading.Condition()
        self._pending: dict[ServedSession, tuple] = {}
        self._queued: list[tuple] = []
        self._running: set = set()  # sessions of the cycle in flight
        self._stop = False
        self.cycles = 0          # batched cycles run
        self.requests_run = 0    # requests over those cycles
        self.max_batch = 0       # most sessions in one cycle
        self._thread = threading.Thread(target=self._loop, daemon=True,
                                        name="gpumap-serve")
        self._thread.start()

    # -- sessions --------------------------------------------------------
    def open(self, data_loader, params, strategy=None,
             insertion_estimator=None, update_estimator=None,
             embedding_estimator=None, default_n_ops: int = 100,
             admission_timeout: float | None = None) -> ServedSession:
        """Admit a session (waiting until its footprint fits) and create it
        with the ``GPUMAP`` constructor's arguments."""
        n = params.n_instances or data_loader.n_instances
        d = params.n_features or data_loader.n_features
        mode = getattr(params, "knn_mode", "exact_bf16")
        mode = {"nndescent": 0, "exact_fp32": 1, "exact_bf16": 2,
                "exact_fp8": 3, "nndescent_bf16": 4}.get(mode, mode)
        fp = session_footprint(n, d, params.n_neighbors, params.n_components,
                               int(mode), self.max_batch_rows)
        self.admission.acquire(fp, admission_timeout)
        try:
            model = GPUMAP(data_loader, params, strategy, insertion_estimator,
                           update_estimator, embedding_estimator,
                           default_n_ops)
        except BaseException:
            self.admission.release(fp)
            raise
        return ServedSession(self, model, fp)

    def _close_session(self, s: ServedSession) -> None:
        with self._cv:
            # every request of a closing session still waiting is cancelled:
            # the pending one and those queued behind it
            req = self._pending.pop(s, None)
            dropped = [q for q in self._queued if q[0] is s]
            self._queued = [q for q in self._queued if q[0] is not s]
            self._last = [q for q in self._last if q is not s]
            self._cv.notify_all()
            # a cycle in flight still uses the session's device arrays (the
            # loop thread itself never waits: it closes sessions only from
            # result callbacks, after the cycle's device work)
            if threading.current_thread() is not self._thread:
                self._cv.wait_for(lambda: s not in self._running)
        if req is not None:
            req[2].cancel()
        for q in dropped:
            q[2].cancel()
        # the device memory is freed once no tensor from a getter holds it;
        # the footprint returns to the budget only then
        model = s._model
        model.release()
        self.admission.release_when(s.footprint, lambda: model.released)

    # -- requests --------------------------------------------------------
    def _submit(self, s, target_latency, n_ops) -> Future:
        fut: Future = Future()
        with self._cv:
            if self._stop:
                raise RuntimeError("server is closed")
            if s in self._pending:
                # one request per session per cycle; later ones queue
                self._queued.append((s, (target_latency, n_ops), fut))
            else:
                self._pending[s] = ((target_latency, n_ops), None, fut)
                # wake the loop for the cycle's first request (it then waits
                # for the batch window) and when the batch is complete, not
                # for every request: each wake-up contends for the GIL with
                # the clients still answering
                if len(self._pending) == 1 or self._complete():
                    self._cv.notify_all()
        return fut

    def _complete(self) -> bool:
        # every session of the previous cycle has its next request pending
        return all(q in self._pending or q._closed for q in self._last)

    def _loop(self) -> None:
        while True:
            with self._cv:
                self._cv.wait_for(lambda: self._pending or self._stop)
                if self._stop and not self._pending:
                    return
                if self.batch_window > 0:
                    # let the other sessions' next requests join this cycle
                    self._cv.wait_for(lambda: self._stop or self._complete(),
                                      self.batch_window)
                batch = list(self._pending.items())
                self._pending.clear()
                self._last = [s for s, _ in batch]
                self._running = set(self._last)
                # the next queued request of each session joins the
                # following cycle
                rest = []
                for q in self._queued:
                    if q[0] not in self._pending:
                        self._pending[q[0]] = (q[1], None, q[2])
                    else:
                        rest.append(q)
                self._queued = rest
            try:
                self._run_cycle(batch)
            finally:
                self._cycle_done()

    def _cycle_done(self) -> None:
        # the cycle's device work has finished: its sessions may be released
        with self._cv:
            self._running = set()
            self._cv.notify_all()

    def _run_cycle(self, batch) -> None:
        live, ops = [], []
        for s, ((target_latency, n_ops), _, fut) in batch:
            if not fut.set_running_or_notify_cancel():
                continue
            m = s._model
            try:
                if n_ops is None:
                    a = m.strategy.allocate(target_latency)
                    ni = min(a[0], self.max_batch_rows,
                             m.data_loader.n_instances
                             - m.n_inserted_instances)
                    o = (ni, a[1], a[2])
                else:
                    o = tuple(int(v) for v in n_ops)
                m._feed(o[0])
            except BaseException as e:
                fut.set_exception(e)
                continue
            live.append((s, fut))
            ops.append(o)
        if not live:
            return
        try:
            results = _core.run_batch([s._model._impl for s, _ in live], ops)
        except BaseException as e:
            self._cycle_done()
            for _, fut in live:
                fut.set_exception(e)
            return
        self._cycle_done()
        self.cycles += 1
        self.requests_run += len(live)
        self.max_batch = max(self.max_batch, len(live))
        for (s, fut), o, r in zip(live, ops, results):
            s.ops_history.append(tuple(int(v) for v in o))
            try:
                fut.set_result(s._model._finish(r))
            except BaseException as e:  # pragma: no cover - strategy errors
                fut.set_exception(e)

    def close(self) -> None:
        """Finish the queued requests and stop the loop."""
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        self._thread.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
