"""Query-server scheduling without a GPU (null executor: batches complete after a fixed delay and
compute nothing): memory admission, continuous batching, cancellation, timeouts, error paths."""
from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import load_rtg

rtg = load_rtg()


def server(**kw):
    kw.setdefault("null_executor", 1)
    return rtg.Server(None, None, None, 0, **kw)


def test_admission_by_memory_budget():
    s = server(slots=1, null_latency_us=200_000, mem_budget_bytes=1000)
    keys = np.arange(100, dtype=np.uint32)  # footprint 16 + 8*100 = 816 bytes
    r1, _ = s.submit_lookup(keys)
    with pytest.raises(rtg.RTGError) as e:
        s.submit_lookup(keys)
    assert e.value.code == 6  # RTG_EAGAIN: rejected, nothing queued
    assert s.wait(r1)[0] == rtg.REQ_DONE
    r3, _ = s.submit_lookup(keys)  # budget freed by completion
    assert s.wait(r3)[0] == rtg.REQ_DONE
    st = s.stats()
    assert (st["admitted"], st["rejected"], st["completed"]) == (2, 1, 2)
    s.close()


def test_continuous_batching_groups_waiting_requests():
    s = server(slots=2, null_latency_us=3000)
    ids = [s.submit_pair(i, i + 1) for i in range(300)]
    for r in ids:
        assert s.wait(r)[0] == rtg.REQ_DONE
    st = s.stats()
    assert st["completed"] == 300
    assert st["batches"] < 300 and st["max_batch_requests"] > 1
    assert st["max_busy_slots"] == 2
    assert st["p99_us"] >= st["p50_us"] > 0
    s.close()


def test_batch_capacity_limits():
    s = server(slots=1, null_latency_us=1000, max_batch_requests=7, max_batch_keys=64)
    ids = [s.submit_pair(0, 1) for _ in range(20)]
    ids += [s.submit_lookup(np.arange(40, dtype=np.uint32))[0] for _ in range(5)]
    for r in ids:
        assert s.wait(r)[0] == rtg.REQ_DONE
    st = s.stats()
    assert st["max_batch_requests"] <= 7
    with pytest.raises(rtg.RTGError):
        s.submit_lookup(np.arange(65, dtype=np.uint32))  # larger than a batch
    s.close()


def test_cancel_pending_only():
    s = server(slots=1, null_latency_us=100_000)
    a = s.submit_pair(1, 2)
    time.sleep(0.02)  # a is running
    b = s.submit_pair(3, 4)
    c = s.submit_pair(5, 6)
    s.cancel(c)
    with pytest.raises(rtg.RTGError):
        s.cancel(a)  # already running
    assert s.wait(c)[0] == rtg.REQ_CANCELLED
    assert s.wait(a)[0] == rtg.REQ_DONE and s.wait(b)[0] == rtg.REQ_DONE
    assert s.stats()["cancelled"] == 1
    s.close()


def test_queue_timeout():
    s = server(slots=1, null_latency_us=50_000, timeout_us=5_000)
    first = s.submit_pair(0, 0)
    time.sleep(0.005)
    late = [s.submit_pair(0, 0) for _ in range(3)]
    assert s.wait(first)[0] == rtg.REQ_DONE
    assert all(s.wait(r)[0] == rtg.REQ_TIMEOUT for r in late)
    s.close()


def test_error_paths_and_destroy_with_queue():
    s = server(slots=1, null_latency_us=100_000)
    with pytest.raises(rtg.RTGError):
        s.poll(123456)
    with pytest.raises(rtg.RTGError):
        s.submit_lookup(np.zeros(0, np.uint32))
    for _ in range(5):
        s.submit_pair(0, 1)
    s.close()  # waits for the queue to drain, then stops


def test_replay_null_executor():
    s = server(slots=4, null_latency_us=100)
    st, _, _ = s.replay(rate_qps=50_000, nqueries=2000, warmup=100, pair_frac=0.7, max_keys=64, seed=3)
    assert st["completed"] + st["rejected"] + st["cancelled"] == 1900
    assert st["completed"] > 0 and st["p99_us"] > 0
    s.close()


def test_request_table_across_chunks():
    """More than 2^16 requests (the id table's chunk) in flight and collected out of order:
    every id stays valid until collected, then is unknown."""
    s = server(slots=4, null_latency_us=100, max_batch_requests=4096)
    ids = [s.submit_pair(i, i) for i in range(70_000)]
    order = ids[65_530:65_545] + ids[::-1]
    seen = set()
    for r in order:
        if r in seen:
            continue
        seen.add(r)
        assert s.wait(r)[0] == rtg.REQ_DONE
    with pytest.raises(rtg.RTGError):
        s.poll(ids[0])
    assert s.stats()["completed"] == 70_000
    s.close()


def test_cancel_races_dispatch():
    """Cancels issued while the dispatcher is taking batches: every request ends exactly once,
    as CANCELLED when its cancel succeeded and DONE when the cancel found it already claimed,
    and the counters add up (a request is never both cancelled and launched)."""
    import threading

    s = server(slots=4, null_latency_us=30, max_batch_requests=16, mem_budget_bytes=16 * 4096)
    n = 6000
    ids = [0] * n
    cancelled_ok = [False] * n
    submitted = threading.Semaphore(0)

    def canceller():
        for i in range(n):
            submitted.acquire()
            try:
                s.cancel(ids[i])
                cancelled_ok[i] = True
            except rtg.RTGError as e:
                assert e.code == 6  # RTG_EAGAIN: already claimed by a batch or finished

    t = threading.Thread(target=canceller)
    t.start()
    for i in range(n):
        while True:
            try:
                ids[i] = s.submit_pair(i % 97, i % 89)
                break
            except rtg.RTGError as e:  # budget full: let the batches drain
                assert e.code == 6
                time.sleep(50e-6)
        submitted.release()
    t.join()
    finals = [s.wait(r)[0] for r in ids]
    for ok, st in zip(cancelled_ok, finals):
        assert st == (rtg.REQ_CANCELLED if ok else rtg.REQ_DONE)
    st = s.stats()
    assert st["cancelled"] == sum(cancelled_ok)
    assert st["completed"] == n - sum(cancelled_ok)
    assert st["completed"] + st["cancelled"] == st["admitted"] == n
    # the admitted bytes all returned to the budget: a full budget's worth is admitted again
    s2_ids = [s.submit_pair(0, 1) for _ in range(4096)]
    for r in s2_ids:
        assert s.wait(r)[0] == rtg.REQ_DONE
    s.close()
