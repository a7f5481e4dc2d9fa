"""CPU tests of the request loop's admission rules (no GPU: requests that
can never be admitted are rejected before any device work starts)."""
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
