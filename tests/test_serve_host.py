"""Host logic of the continuous-batching request loop (rumdeed_amd/serve.py),
run against a stub pool with the same contract as rumdeed_amd.diode.Pool:
try_open returns a slot or None when no contiguous range of the store is
free, step returns one record per slot.  No device is involved."""
import threading

import numpy as np
import pytest

import rumdeed_amd.serve as S

NM = 1e-9


class _Rec:
    def __init__(self, step):
        self.step = step
        self.n = 10
        self.n_round = 2
        self.emitted = 1
        self.absorbed = 1
        self.dropped = 0
        self.ramo_current = 1e-3

    def as_dict(self):
        return {"step": self.step, "n": self.n}


class StubPool:
    """First-fit contiguous ranges over a store of `total` particles."""

    def __init__(self, total, max_slots, device=0):
        self.total = total
        self.max_slots = max_slots
        self.ranges = {}  # slot -> (base, cap)
        self.steps = {}

    @staticmethod
    def footprint(capacity):
        return 200 * int(capacity)

    def set_graphs(self, on):
        pass

    def close(self):
        pass

    def try_open(self, cap):
        free_slots = [s for s in range(self.max_slots) if s not in self.ranges]
        if not free_slots:
            return None
        base = 0
        for b, c in sorted(self.ranges.values()):
            if b - base >= cap:
                break
            base = b + c
        if base + cap > self.total:
            return None
        s = free_slots[0]
        self.ranges[s] = (base, cap)
        self.steps[s] = 0
        return s

    def close_slot(self, s):
        del self.ranges[s]
        del self.steps[s]

    def select(self, s):
        pass

    def set_planar(self, *a, **k):
        pass

    def set_time_step(self, *a):
        pass

    def set_seed(self, *a):
        pass

    def set_emitter(self, *a, **k):
        pass

    def step(self, precision):
        recs = [None] * self.max_slots
        for s in self.ranges:
            self.steps[s] += 1
            recs[s] = _Rec(self.steps[s])
        return recs


def _serve(server, reqs, **kw):
    out = {}

    def go():
        out["r"] = server.run(reqs, realtime=False, keep_records=True, **kw)

    t = threading.Thread(target=go, daemon=True)
    t.start()
    t.join(20.0)
    assert not t.is_alive(), "Server.run did not return"
    return out["r"]


@pytest.fixture
def stub(monkeypatch):
    monkeypatch.setattr(S, "Pool", StubPool)


def test_oversize_request_is_refused_not_waited_for(stub):
    """An 8 kV / 1000 nm / 1000 nm request needs more particles than a pool
    of 1 << 15 holds: it comes back refused, and the requests queued
    behind it are still served."""
    reqs = [S.Request(rid=0, V=8000.0, gap=1000 * NM, side=1000 * NM, phi=4.7, steps=3),
            S.Request(rid=1, V=1600.0, gap=200 * NM, side=100 * NM, phi=4.7, steps=4)]
    assert S.capacity_estimate(reqs[0]) > 1 << 15
    srv = S.Server(total_particles=1 << 15, max_slots=2)
    res, stats = _serve(srv, reqs)
    assert [r.rid for r in res] == [0, 1]
    assert res[0].error and "exceeds" in res[0].error and not res[0].valid
    assert res[1].error == "" and res[1].valid and len(res[1].records) == 4
    assert stats["refused"] == 1 and stats["requests"] == 1


def test_request_over_the_hbm_budget_is_refused(stub):
    reqs = [S.Request(rid=0, V=1600.0, gap=200 * NM, side=100 * NM, phi=4.7, steps=2,
                      capacity=4096)]
    srv = S.Server(total_particles=1 << 15, max_slots=2, hbm_budget=200 * 1024)
    res, stats = _serve(srv, reqs)
    assert "budget" in res[0].error and stats["refused"] == 1


def test_request_that_waits_for_room_is_admitted_later(stub):
    """Two requests that each fill most of the store run one after the
    other: the second waits (it fits an empty store) rather than being
    refused."""
    reqs = [S.Request(rid=i, V=1600.0, gap=200 * NM, side=100 * NM, phi=4.7, steps=3 + i,
                      capacity=3 << 13) for i in range(2)]
    srv = S.Server(total_particles=1 << 15, max_slots=4)
    res, stats = _serve(srv, reqs)
    assert all(r.valid for r in res)
    assert stats["max_active"] == 1 and stats["pool_steps"] == 3 + 4


def test_capacity_estimate_is_not_clamped():
    """The slot capacity follows the space-charge estimate without a cap,
    so a large request is sized (or refused), never silently truncated."""
    r = S.Request(rid=0, V=8000.0, gap=1000 * NM, side=1000 * NM, phi=4.7)
    F = 8000.0 / (1000 * NM)
    n = 3.0 * S.EPS0 * F * (1000 * NM) ** 2 / S.Q0
    cap = S.capacity_estimate(r)
    assert cap >= n and cap % 256 == 0 and cap > 1 << 16


def test_dropped_electrons_mark_a_result_invalid():
    r = S.Result(rid=0, arrival=0.0, admitted=0.0, finished=1.0, steps=1, dropped=3)
    assert not r.valid
    assert S.Result(rid=0, arrival=0.0, admitted=0.0, finished=1.0, steps=1).valid


def test_stream_requests_shape():
    reqs = S.stream_requests(n=64, rate=4.0, seed=2024, steps=10)
    gaps = np.array([r.gap for r in reqs]) / NM
    assert gaps.min() >= 200 and gaps.max() <= 2000
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0)
    assert [r.seed for r in reqs] == list(range(64))
