"""Host logic of the server's admission (no GPU): budget accounting,
oversize requests, cancellation and the split of a counted batch."""
import collections
import threading
from concurrent.futures import Future

import pytest


def fake_server(pkg, budget, oversize="alone", max_batch=8, out_guess=1.0):
    S = pkg.Server
    s = object.__new__(S)
    s.lock = threading.Condition()
    s.pending = collections.deque()
    s.in_use = 0
    s.budget = budget
    s.oversize = oversize
    s.max_batch = max_batch
    s.out_guess = out_guess
    s.running = False  # an empty queue returns [] instead of waiting
    s.paused = False
    s.stats = dict(peak_in_use=0, cancelled=0, rejected=0, splits=0)
    return s


def req(pkg, nbytes, units=0):
    from importlib import import_module
    serve = import_module(pkg.__name__ + ".serve")
    return serve._Req(None, None, None, Future(), 0.0, nbytes, units)


def admit(s):
    with s.lock:
        return s._admit()


def test_admission_respects_budget(pkg):
    s = fake_server(pkg, budget=10_000_000)
    rs = [req(pkg, 2_000_000) for _ in range(5)]
    s.pending.extend(rs)
    b = admit(s)
    # each estimate is 2 x 2 MB + 64 KiB: two fit, the third does not
    assert b == rs[:2]
    assert s.in_use == sum(r.reserved for r in b) <= s.budget
    assert all(r.future.running() for r in b)
    assert list(s.pending) == rs[2:]


def test_cancelled_requests_are_dropped(pkg):
    s = fake_server(pkg, budget=1 << 30)
    rs = [req(pkg, 1000) for _ in range(4)]
    assert rs[1].future.cancel()
    s.pending.extend(rs)
    b = admit(s)
    assert b == [rs[0], rs[2], rs[3]]
    assert s.stats["cancelled"] == 1


def test_lone_oversize_reserves_counting_memory_or_is_rejected(pkg):
    s = fake_server(pkg, budget=3_000_000)
    big = req(pkg, 2_000_000)  # estimate 4 MB + 64 KiB: over the budget
    s.pending.append(big)
    assert admit(s) == [big]
    assert big.reserved == 2_000_000 + (1 << 16) <= s.budget
    s = fake_server(pkg, budget=1_000_000, oversize="reject")
    huge, small = req(pkg, 2_000_000), req(pkg, 1000)
    s.pending.extend([huge, small])
    assert admit(s) == [small]
    with pytest.raises(MemoryError, match="budget"):
        huge.future.result(timeout=0)
    assert s.stats["rejected"] == 1
    s = fake_server(pkg, budget=1_000_000, oversize="alone")
    huge = req(pkg, 2_000_000)
    s.pending.append(huge)
    assert admit(s) == [huge]  # runs alone, over the budget by policy


def test_fit_splits_counted_batch(pkg):
    s = fake_server(pkg, budget=10_000)
    batch = [req(pkg, 0) for _ in range(4)]
    for r in batch:
        r.reserved = 100
    s.in_use = 400
    with s.lock:
        assert s._fit(batch, [3000, 3000, 3000, 3000]) == 3
        assert s._fit(batch, [20_000, 1, 1, 1]) == 1  # alone, policy "alone"
        s.oversize = "reject"
        assert s._fit(batch, [20_000, 1, 1, 1]) == 0
        assert s._fit(batch, [2000, 2000, 2000, 2000]) == 4
