"""Host logic of the serving form: footprint estimate and admission."""

import threading
import time

import pytest

from gpumap_amd.serving import Admission, session_footprint


def test_footprint_grows_with_shape_and_counts_bf16_scratch():
    base = session_footprint(70_000, 784)
    assert base > 70_000 * 784 * 8          # rows and their BF16 split
    assert session_footprint(140_000, 784) > 1.9 * base
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
