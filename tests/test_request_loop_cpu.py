"""CPU tests of the request loop's admission rules (no GPU: requests that
can never be admitted are rejected before any device work starts)."""
import pytest

import licom_ensemble as le
import licom_synth as ls


class FakeClock:
    """Deterministic clock: sleep() advances it, nothing else does."""

    def __init__(self):
        self.t = 0.0
        self.sleeps = 0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        assert dt > 0, "the loop must never spin without sleeping"
        self.sleeps += 1
        self.t += dt


def small_case(seed=1):
    return ls.make_case("C1", nx=16, ny=8, km=6, seed=seed)


def test_over_cap_request_is_rejected_not_blocking():
    c = small_case()
    reqs = [le.Request(0, 0.0, c, 24), le.Request(1, 0.01, c, 24)]
    cap = 0.5 * reqs[0].bytes               # neither fits
    loop = le.RequestLoop(hbm_cap_bytes=cap, free_bytes=None)
    clk = FakeClock()
    stats = loop.run(reqs, clock=clk, sleep=clk.sleep)
    assert stats["requests"] == 0 and stats["rejected"] == 2
    assert [r.rid for r in loop.rejected] == [0, 1]
    assert all("cap" in r.error for r in reqs)
    assert all(r.model is None and r.finished is None for r in reqs)


def test_request_larger_than_free_device_memory_is_rejected():
    c = small_case(2)
    r = le.Request(0, 0.002, c, 24)
    loop = le.RequestLoop(hbm_cap_bytes=200e9, free_bytes=lambda: 0.5 * r.bytes,
                          free_margin_bytes=0.0)
    clk = FakeClock()
    stats = loop.run([r], clock=clk, sleep=clk.sleep)
    assert stats["rejected"] == 1 and "free" in r.error
    # it waited for its arrival by sleeping, not by spinning
    assert clk.sleeps >= 1 and clk.t >= r.arrival


def test_srpt_pick_orders_by_remaining_work_and_fills():
    """The srpt policy runs the requests with the least remaining work
    (steps x cells per step) first, as many as fill the GPU's cells per step,
    and always at least one."""
    loop = le.RequestLoop(policy="srpt", fill_cells=1.0e6, free_bytes=None)

    class C:
        def __init__(self, nx, ny, km):
            self.nx, self.ny, self.km = nx, ny, km

    class R:
        def __init__(self, rid, shape, steps, done=0, arrival=0.0):
            self.rid, self.case, self.n_steps, self.done_steps, self.arrival = rid, C(*shape), steps, done, arrival

    small, mid, big = (96, 64, 30), (128, 96, 30), (360, 218, 55)
    a = R(0, small, 200, done=190)          # 10 steps x 184k cells left
    b = R(1, mid, 24)                       # 24 x 369k
    c = R(2, small, 240)                    # 240 x 184k
    d = R(3, big, 30)                       # 30 x 4.3M: larger than the fill alone
    assert [r.rid for r in loop._srpt_pick([d, c, b, a])] == [0, 1, 2]
    assert [r.rid for r in loop._srpt_pick([d])] == [3]
    with pytest.raises(ValueError):
        le.RequestLoop(policy="fifo")


def test_solo_pick_least_remaining_work_bound_or_unstarted():
    """The solo policy's lane takes, among the requests bound to it and those
    not started, the one with the least remaining work (steps not yet launched
    x cells per step); requests bound to another lane are never taken."""
    class C:
        def __init__(self, nx, ny, km):
            self.nx, self.ny, self.km = nx, ny, km

    class R:
        def __init__(self, rid, shape, steps, done=0, queued=0, lane=None, arrival=0.0):
            self.rid, self.case, self.n_steps = rid, C(*shape), steps
            self.done_steps, self.queued, self.lane, self.arrival = done, queued, lane, arrival

    small, big = (96, 64, 30), (360, 218, 55)
    a = R(0, big, 240, done=200, lane=1)         # 40 x 4.3M, on lane 1
    b = R(1, small, 240, done=100, lane=0)       # 140 x 184k, on lane 0
    c = R(2, small, 24, arrival=0.5)             # 24 x 184k, not started
    d = R(3, small, 24, arrival=0.4)             # same work, arrived first
    e = R(4, small, 30, done=20, queued=10, lane=0)   # everything launched
    pick = le.RequestLoop._solo_pick
    assert pick([a, b, c, d, e], 0).rid == 3
    assert pick([a, b, c, e], 0).rid == 2
    assert pick([a, b, e], 0).rid == 1
    assert pick([a, b, e], 1).rid == 0
    assert pick([a, e], 0) is None
    assert pick([a, b], 2) is None
    with pytest.raises(ValueError):
        le.RequestLoop(policy="solo", lanes=0)
