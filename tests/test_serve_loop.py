"""Scheduling of the serving request loop (ftk2_amd/serve.py), on the CPU: a fake
server with the ftk2h_server contract (budgeted admission, one slab per stream per
step) driven by a deterministic clock."""
import numpy as np
import pytest

from ftk2_amd import serve


class FakeServer:
    def __init__(self, budget):
        self.budget = budget
        self.used = 0
        self.next = 1
        self.streams = {}
        self.steps = 0
        self.clock = 0.0
        self.max_active = 0
        self.open_order = []

    def open(self, spatial, kind, m, dtype, ncomp, max_t, thr=(0.0, 0.0)):
        need = int(np.prod(spatial)) * max_t
        if need > self.budget:
            raise RuntimeError("never fits")
        if self.used + need > self.budget:
            return None
        sid = self.next
        self.next += 1
        self.used += need
        self.streams[sid] = {"need": need, "q": 0, "closed": False, "slabs": 0, "done": False, "spatial": spatial}
        self.open_order.append(spatial)
        return sid

    def append(self, sid, comps):
        self.streams[sid]["q"] += 1

    def close_stream(self, sid):
        self.streams[sid]["closed"] = True

    def step(self):
        self.steps += 1
        self.clock += 1.0
        live = [s for s in self.streams.values() if not s["done"]]
        self.max_active = max(self.max_active, len(live))
        done = []
        for sid, s in self.streams.items():
            if s["done"]:
                continue
            need_slabs = max(0, s["q"] - 1)
            if s["slabs"] < need_slabs:
                s["slabs"] += 1
            if s["closed"] and s["slabs"] == need_slabs:
                s["done"] = True
                done.append(sid)
        return done

    def take(self, sid):
        s = self.streams.pop(sid)
        self.used -= s["need"]
        return ("result", s["spatial"], s["slabs"])


def make_requests(n, rng, rate):
    arr = serve.poisson_arrivals(n, rate, rng)
    reqs = []
    for i in range(n):
        T = int(rng.integers(1, 9))
        side = int(rng.integers(2, 6))
        reqs.append(serve.Request(i, float(arr[i]), (side, side), 0, 2, 1, [[None, None]] * T,
                                  simplices=(side - 1) ** 2 * max(0, T - 1) * 12))
    return reqs


def run_fake(server, reqs):
    clk = lambda: server.clock
    def slp(dt):
        server.clock += dt
    return serve.run(server, reqs, keep_results=True, clock=clk, sleep=slp)


def test_every_request_answered_with_its_own_result():
    rng = np.random.default_rng(0)
    reqs = make_requests(40, rng, rate=2.0)
    srv = FakeServer(budget=10 ** 9)
    outs, wall, rej = run_fake(srv, reqs)
    assert rej == []
    assert [o.rid for o in outs] == list(range(40))
    for o in outs:
        rq = reqs[o.rid]
        assert o.result == ("result", rq.spatial, max(0, len(rq.timesteps) - 1))
        assert o.done >= o.admitted >= o.arrival
        assert o.latency >= 1.0  # at least one step
    assert srv.max_active > 1  # streams overlap (continuous batching)


def test_budget_admits_first_come_first_served_and_rejects_oversize():
    rng = np.random.default_rng(1)
    reqs = make_requests(30, rng, rate=50.0)  # everything arrives at once
    big = serve.Request(99, 0.0, (100, 100), 0, 2, 1, [[None, None]] * 4)
    srv = FakeServer(budget=250)  # the largest request needs 5*5*8 = 200
    outs, wall, rej = run_fake(srv, reqs + [big])
    assert rej == [99]
    assert len(outs) == 30
    # admission order follows arrival order
    order = sorted(reqs, key=lambda r: r.arrival)
    assert srv.open_order == [r.spatial for r in order]
    assert srv.used == 0


def test_latency_grows_with_load():
    lo = make_requests(60, np.random.default_rng(3), rate=0.05)
    hi = make_requests(60, np.random.default_rng(3), rate=50.0)
    o_lo, w_lo, _ = run_fake(FakeServer(400), lo)
    o_hi, w_hi, _ = run_fake(FakeServer(400), hi)
    p_lo = serve.summarize(o_lo, lo, w_lo)
    p_hi = serve.summarize(o_hi, hi, w_hi)
    assert p_hi["p99_ms"] > p_lo["p99_ms"]
    assert p_hi["p50_ms"] >= p_lo["p50_ms"]
