"""Host-only tests of the continuous-batching request loop (ttsh3d.serve)
with a fake device: an HBM budget allocator and a step counter, no GPU."""
import math

import pytest

from ttsh3d import serve


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeBackend:
    """Admits by byte budget and slot count; each step advances the clock."""

    def __init__(self, clock, budget, max_slots, step_cost=1e-3, bytes_per_cell=88):
        self.clock, self.budget, self.max_slots = clock, budget, max_slots
        self.step_cost, self.bpc = step_cost, bytes_per_cell
        self.used = 0
        self.peak = 0
        self.slots = {}
        self.admit_log = []

    def prepare(self, r):
        pass

    def admit(self, r):
        need = self.bpc * r.N ** 3
        if self.used + need > self.budget or len(self.slots) >= self.max_slots:
            return None
        slot = min(set(range(self.max_slots)) - set(self.slots))
        self.slots[slot] = [r, 0, need]
        self.used += need
        self.peak = max(self.peak, self.used)
        self.admit_log.append((r.rid, self.clock.t, len(self.slots)))
        return slot

    def step(self, n):
        for s in self.slots.values():
            s[1] = min(s[1] + n, s[0].steps)
        self.clock.t += n * self.step_cost

    def sync(self):
        pass

    def wait(self, keep):
        pass

    def ready(self, slot):
        r, steps, _ = self.slots[slot]
        return steps >= r.steps

    def steps_done(self, slot):
        return self.slots[slot][1]

    def finish(self, slot):
        r, steps, need = self.slots.pop(slot)
        self.used -= need
        return (1.0, 0.5, 0.1), True


def test_all_complete_within_budget():
    clock = FakeClock()
    reqs = serve.make_requests(60, rate=200.0, sizes=(32, 48, 64), steps=(16, 160), seed=3)
    budget = 88 * (2 * 64 ** 3 + 48 ** 3)
    be = FakeBackend(clock, budget, max_slots=8)
    done, wall = serve.Server(be, macro=16, clock=clock, sleep=clock.sleep).run(reqs)
    assert len(done) == 60 and all(r.ok for r in done)
    assert be.peak <= budget
    assert be.used == 0
    for r in done:
        assert r.admitted >= r.arrival - 1e-12
        assert r.completed >= r.admitted
    s = serve.summarize(done, wall)
    assert s["completed"] == 60 and s["latency_p99_s"] >= s["latency_p50_s"] > 0


def test_continuous_admission_overlaps_running_work():
    """A request that arrives while another runs is admitted before that one ends."""
    clock = FakeClock()
    a = serve.Request(0, 0.0, 64, 1600, -1.0, 1.0, 1)
    b = serve.Request(1, 0.1, 32, 160, -1.0, 1.0, 2)
    be = FakeBackend(clock, budget=10 ** 12, max_slots=4, step_cost=1e-3)
    done, _ = serve.Server(be, macro=16, clock=clock, sleep=clock.sleep).run([a, b])
    ra = next(r for r in done if r.rid == 0)
    rb = next(r for r in done if r.rid == 1)
    assert rb.admitted < ra.completed
    assert rb.completed < ra.completed
    assert math.isclose(rb.admitted, 0.112, abs_tol=0.02)


def test_fifo_when_budget_blocks():
    clock = FakeClock()
    reqs = [serve.Request(i, 0.0, 64, 64, -1.0, 1.0, i) for i in range(5)]
    be = FakeBackend(clock, budget=88 * 64 ** 3 * 2, max_slots=8)
    done, _ = serve.Server(be, macro=16, clock=clock, sleep=clock.sleep).run(reqs)
    order = [rid for rid, _, _ in be.admit_log]
    assert order == sorted(order)
    assert max(n for _, _, n in be.admit_log) == 2


def test_oversized_request_rejected():
    clock = FakeClock()
    reqs = [serve.Request(0, 0.0, 256, 32, -1.0, 1.0, 0), serve.Request(1, 0.0, 32, 32, -1.0, 1.0, 1)]
    be = FakeBackend(clock, budget=88 * 64 ** 3, max_slots=4)
    done, _ = serve.Server(be, macro=16, clock=clock, sleep=clock.sleep).run(reqs)
    by = {r.rid: r for r in done}
    assert not by[0].ok and by[1].ok


def test_request_generator_is_seeded():
    a = serve.make_requests(20, 1.0, seed=9)
    b = serve.make_requests(20, 1.0, seed=9)
    assert [(r.arrival, r.N, r.steps, r.cA, r.La) for r in a] == [(r.arrival, r.N, r.steps, r.cA, r.La) for r in b]
    assert all(64 <= r.N <= 192 and 1000 <= r.steps <= 20000 for r in a)
    assert all(-10.0 <= r.cA <= -0.1 and 0.5 <= r.La ** 2 <= 2.0 for r in a)


def test_bounded_bypass_of_a_blocked_head():
    """A head request that does not fit lets later arrivals that fit pass it,
    at most `bypass` times; then admission waits for the head."""
    clock = FakeClock()
    big = lambda i: serve.Request(i, 0.0, 64, 640, -1.0, 1.0, i)
    small = lambda i: serve.Request(i, 0.0, 32, 64, -1.0, 1.0, i)
    reqs = [big(0), big(1), big(2), small(3), small(4), small(5)]
    budget = 88 * (2 * 64 ** 3 + 2 * 32 ** 3)
    for bypass, passed in ((0, []), (1, [3]), (8, [3, 4, 5])):
        clock.t = 0.0
        be = FakeBackend(clock, budget, max_slots=8)
        serve.Server(be, macro=16, clock=clock, sleep=clock.sleep, bypass=bypass).run([
            serve.Request(r.rid, r.arrival, r.N, r.steps, r.cA, r.La, r.seed) for r in reqs])
        order = [rid for rid, _, _ in be.admit_log]
        head_at = order.index(2)
        assert order[:2] == [0, 1]
        assert sorted(order[2:head_at]) == passed, (bypass, order)


def test_summary_leaves_out_rejected_requests():
    """A rejected request (never stepped) counts as rejected only: it adds no
    cells, no completion and no latency."""
    clock = FakeClock()
    reqs = [serve.Request(0, 0.0, 256, 32, -1.0, 1.0, 0), serve.Request(1, 0.0, 32, 32, -1.0, 1.0, 1)]
    be = FakeBackend(clock, budget=88 * 64 ** 3, max_slots=4)
    done, wall = serve.Server(be, macro=16, clock=clock, sleep=clock.sleep).run(reqs)
    s = serve.summarize(done, wall)
    assert s["completed"] == 1 and s["rejected"] == 1 and s["failed"] == 0
    ok = next(r for r in done if r.rid == 1)
    assert s["latency_p50_s"] == pytest.approx(ok.completed - ok.arrival, abs=1e-4)
    assert s["mcell_steps_per_s"] == pytest.approx(32 ** 3 * 32 / wall / 1e6, rel=1e-2)
