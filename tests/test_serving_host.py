"""Host logic of the serving form: footprint estimate and admission."""

import threading
import time

import pytest

from gpumap_amd.serving import Admission, session_footprint


def test_footprint_grows_with_shape_and_counts_bf16_scratch():
    base = session_footprint(70_000, 784)
    assert base > 70_000 * 784 * 8          # rows and their BF16 split
    assert session_footprint(140_000, 784) > 1.7 * base
    assert session_footprint(70_000, 784, knn_mode=1) < base
    # the lower-bound store (N x batch floats) is counted up to its cap
    assert (session_footprint(70_000, 784, max_batch_rows=2000)
            - session_footprint(70_000, 784, max_batch_rows=1000)
            >= 70_000 * 1000 * 4)


def test_admission_waits_for_room_and_rejects_oversize():
    a = Admission(budget_bytes=100, max_sessions=8)
    a.acquire(60)
    with pytest.raises(MemoryError):
        a.acquire(101)
    with pytest.raises(TimeoutError):
        a.acquire(50, timeout=0.05)
    got = []

    def waiter():
        a.acquire(50)
        got.append(time.monotonic())

    t = threading.Thread(target=waiter)
    t.start()
    time.sleep(0.05)
    assert not got                      # still waiting for room
    released = time.monotonic()
    a.release(60)
    t.join(timeout=5)
    assert got and got[0] >= released
    assert a.used == 50 and a.sessions == 1


def test_admission_bounds_the_session_count():
    a = Admission(budget_bytes=10**9, max_sessions=2)
    a.acquire(1)
    a.acquire(1)
    with pytest.raises(TimeoutError):
        a.acquire(1, timeout=0.05)
    a.release(1)
    a.acquire(1, timeout=1)


def test_admission_returns_a_footprint_only_once_its_memory_is_freed():
    a = Admission(budget_bytes=100, max_sessions=8, poll_interval=0.01)
    a.acquire(80)
    freed = threading.Event()
    a.release_when(80, freed.is_set)
    assert a.pending_release() == 80
    with pytest.raises(TimeoutError):
        a.acquire(50, timeout=0.05)     # the closed session still holds it
    got = []

    def waiter():
        a.acquire(50, timeout=5)
        got.append(time.monotonic())

    t = threading.Thread(target=waiter)
    t.start()
    time.sleep(0.05)
    assert not got
    t0 = time.monotonic()
    freed.set()                          # e.g. the last exported tensor died
    t.join(timeout=5)
    assert got and got[0] >= t0
    assert a.used == 50 and a.sessions == 1 and a.pending_release() == 0


class _FakeModel:
    """Stands in for a GPUMAP session on the host: its first cycle blocks in
    _feed until the test lets it go."""

    def __init__(self, gate):
        self.gate = gate
        self.released = False
        self._impl = object()   # run_batch rejects it: the cycle fails

    def _feed(self, n):
        self.gate.wait(5)

    def release(self):
        self.released = True


def test_closing_a_session_cancels_its_queued_requests_and_waits_for_its_cycle():
    from concurrent.futures import CancelledError

    from gpumap_amd.serving import Server, ServedSession

    srv = Server(hbm_budget_bytes=10**9, batch_window=0)
    gate = threading.Event()
    model = _FakeModel(gate)
    srv.admission.acquire(10)
    s = ServedSession(srv, model, 10)
    f1 = s.submit(n_ops=(1, 0, 0))       # taken by the loop, blocks in _feed
    deadline = time.monotonic() + 5
    while s not in srv._running and time.monotonic() < deadline:
        time.sleep(0.005)
    assert s in srv._running
    f2 = s.submit(n_ops=(1, 0, 0))       # pending for the next cycle
    f3 = s.submit(n_ops=(1, 0, 0))       # queued behind it
    closer = threading.Thread(target=s.close)
    closer.start()
    time.sleep(0.05)
    assert closer.is_alive()             # waits for the cycle in flight
    assert not model.released
    gate.set()
    closer.join(timeout=5)
    assert not closer.is_alive()
    assert model.released
    with pytest.raises(Exception):
        f1.result(timeout=5)             # the fake session's cycle failed
    for f in (f2, f3):
        with pytest.raises(CancelledError):
            f.result(timeout=5)
    assert srv.admission.pending_release() == 0
    assert srv.admission.used == 0
    srv.close()
