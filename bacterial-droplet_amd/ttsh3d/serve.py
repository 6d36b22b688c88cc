"""Continuous-batching request loop for droplet simulations on one GPU.

This replaces the reference's serial sweep (forEach.pl:33-133 runs one
droplet per process, one after another, through run_cuda.pl) with a server:
requests (a droplet's size, parameters, seed and step count) arrive over
time; the loop admits every arrived request whose device state fits the HBM
budget and a free slot, steps all resident droplets together in one batch
(one launch per kernel for every droplet, replayed as a HIP graph of
`macro` steps), retires droplets that reached their step count, and admits
more.  Each droplet's state stays device-resident from admission to
retirement.  Droplets advance exactly their own step count (the batch stops
each at its limit), so a droplet's result is the same alone or in a mix.

Reported: completed sims per hour, p50/p99 completion latency (completion
minus arrival), queueing delay, device occupancy.

    python -m ttsh3d.serve --requests 500 --rate-frac 0.8 ...
"""
from __future__ import annotations

import argparse
import collections
import concurrent.futures as cf
import dataclasses
import json
import math
import sys
import time

import numpy as np


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    N: int
    steps: int
    cA: float
    La: float
    seed: int
    admitted: float = math.nan
    completed: float = math.nan
    ok: bool = True          # finished with finite reductions
    rejected: bool = False   # never admitted: does not fit an empty device
    reductions: tuple = ()


def make_requests(n, rate, sizes=(64, 96, 128, 192), steps=(1000, 20000), seed=1005):
    """Seeded Poisson arrivals at `rate` per second; L uniform over sizes,
    steps uniform in [steps[0], steps[1]], cA = default x logU[0.1, 10],
    La^2 = default x U[0.5, 2] (defaults of run_cuda.pl:30-44: cA -1, La 1)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        N = int(rng.choice(sizes))
        st = int(rng.integers(steps[0], steps[1] + 1))
        cA = -1.0 * math.exp(rng.uniform(math.log(0.1), math.log(10.0)))
        La = math.sqrt(1.0 * rng.uniform(0.5, 2.0))
        out.append(Request(i, t, N, st, cA, La, seed * 1000 + i))
    return out


def _done_future(value):
    f = cf.Future()
    f.set_result(value)
    return f


class GpuBackend:
    """The libttsh3d batch (ttsh_batch_*)."""

    def __init__(self, max_slots, hbm_budget, dt0=0.01, keep_states=False):
        from . import Batch, default_params, droplet_bytes, init_droplet, F32
        self._default_params = default_params
        self._init = init_droplet
        self._F32 = F32
        self.batch = Batch(max_slots=max_slots, hbm_budget=hbm_budget, use_graph=True)
        self.bytes_for = droplet_bytes
        self.dt0 = dt0
        self.pool = cf.ThreadPoolExecutor(max_workers=8)
        self.states = {}
        self.keep_states = keep_states
        self.final = {}  # request id -> (phi, P, v) when keep_states
        self._rid = {}

    def prepare(self, r: Request):
        """Start building the request's initial state on a host thread."""
        if r.rid not in self.states:
            self.states[r.rid] = self.pool.submit(self._init, r.N, 0.3 * r.N, 2.0, 0, r.seed)

    def admit(self, r: Request):
        p = self._default_params(N=r.N, dt0=self.dt0, cA=r.cA, La=r.La)
        self.prepare(r)
        phi, P, v = self.states.pop(r.rid).result()
        slot = self.batch.try_admit(p, phi.astype(np.float32), P.astype(np.float32), v.astype(np.float32),
                                    dtype=self._F32, limit=r.steps)
        if slot is None:
            self.states[r.rid] = _done_future((phi, P, v))
        else:
            self._rid[slot] = r.rid
        return slot

    def step(self, n):
        self.batch.step(n)

    def steps_done(self, slot):
        return self.batch.steps_done(slot)

    def wait(self, keep):
        self.batch.wait(keep)

    def ready(self, slot):
        return self.batch.ready(slot)

    def finish(self, slot):
        red, ok = self.batch.reductions(slot)
        if self.keep_states:
            self.final[self._rid[slot]] = self.batch.get_state(slot)
        self.batch.retire(slot)
        self._rid.pop(slot, None)
        return tuple(float(x) for x in red), ok

    def sync(self):
        self.batch.synchronize()

    def close(self):
        self.pool.shutdown(wait=False)
        self.batch.close()


class Server:
    """bypass: how many later arrivals may be admitted ahead of a head
    request that does not fit yet (0 = strict FIFO); the head keeps its
    place, and once it has been passed that often, admission waits for it."""

    def __init__(self, backend, macro=16, clock=time.perf_counter, sleep=time.sleep, bypass=8):
        self.b = backend
        self.macro = macro
        self.clock = clock
        self.sleep = sleep
        self.bypass = bypass

    def run(self, requests, log=None):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        active = {}  # slot -> request
        done = []
        passes = {}  # head request id -> later arrivals admitted ahead of it
        t0 = self.clock()
        busy_steps = 0
        for r in list(pending)[:16]:
            self.b.prepare(r)
        while pending or active:
            now = self.clock() - t0
            # admit arrived requests in arrival order while memory and slots
            # allow; later arrivals that fit may pass a head that does not,
            # at most `bypass` times per head, after which admission waits
            # for the head
            i = 0
            while i < len(pending) and pending[i].arrival <= now:
                head = pending[0]
                if i > 0 and passes.get(head.rid, 0) >= self.bypass:
                    break
                r = pending[i]
                slot = self.b.admit(r)
                if slot is None:
                    if i == 0 and not active:
                        # cannot fit even on an empty device: reject it
                        pending.popleft()
                        r.ok = False
                        r.rejected = True
                        r.admitted = r.completed = self.clock() - t0
                        done.append(r)
                        continue
                    i += 1
                    continue
                del pending[i]
                if i > 0:
                    passes[head.rid] = passes.get(head.rid, 0) + 1
                r.admitted = self.clock() - t0
                active[slot] = r
                for q in list(pending)[:8]:
                    self.b.prepare(q)
            if not active:
                # idle until the next arrival
                if pending:
                    wait = pending[0].arrival - (self.clock() - t0)
                    if wait > 0:
                        self.sleep(min(wait, 0.05))
                continue
            # queue the next macro step, then wait for the one before it: the
            # device works on the new one while the host retires and admits
            self.b.step(self.macro)
            busy_steps += self.macro
            if all(self.b.steps_done(slot) >= r.steps for slot, r in active.items()):
                self.b.wait(0)  # nothing left running: wait for the last steps
            else:
                self.b.wait(1)
            for slot, r in list(active.items()):
                if self.b.ready(slot):
                    r.reductions, r.ok = self.b.finish(slot)
                    r.completed = self.clock() - t0
                    done.append(r)
                    del active[slot]
                    if log:
                        log(f"done rid={r.rid} N={r.N} steps={r.steps} latency={r.completed - r.arrival:.3f}s")
        wall = self.clock() - t0
        return done, wall


def summarize(done, wall):
    """Throughput and latency over the requests that ran to completion with
    finite results; rejected requests (never admitted) and diverged ones
    (non-finite reductions) are counted apart and enter neither."""
    ok = [r for r in done if r.ok]
    lat = np.array([r.completed - r.arrival for r in ok])
    queue = np.array([r.admitted - r.arrival for r in ok])
    cells = sum(r.N ** 3 * r.steps for r in ok)
    return {
        "completed": len(ok),
        "failed": sum(1 for r in done if not r.ok and not r.rejected),
        "rejected": sum(1 for r in done if r.rejected),
        "wall_s": round(wall, 3),
        "sims_per_hour": round(len(ok) / wall * 3600.0, 2),
        "latency_p50_s": round(float(np.percentile(lat, 50)), 4) if len(lat) else None,
        "latency_p99_s": round(float(np.percentile(lat, 99)), 4) if len(lat) else None,
        "queue_p50_s": round(float(np.percentile(queue, 50)), 4) if len(queue) else None,
        "mcell_steps_per_s": round(cells / wall / 1e6, 1),
    }


def measure_capacity(sizes, steps, budget, max_slots, seed=77):
    """Cell-steps/s of a full batch of the request mix (used to set the rate)."""
    rng = np.random.default_rng(seed)
    be = GpuBackend(max_slots, budget)
    reqs = [Request(i, 0.0, int(rng.choice(sizes)), 10 ** 9, -1.0, 1.0, seed + i) for i in range(max_slots)]
    cells = 0
    for r in reqs:
        if be.admit(r) is None:
            break
        cells += r.N ** 3
    be.step(32)
    be.sync()
    t = time.perf_counter()
    be.step(64)
    be.sync()
    dt = time.perf_counter() - t
    be.close()
    return cells * 64 / dt


def main(argv=None):
    ap = argparse.ArgumentParser(description="continuous-batching droplet server (Poisson arrivals)")
    ap.add_argument("--requests", type=int, default=500)
    ap.add_argument("--sizes", type=int, nargs="+", default=[64, 96, 128, 192])
    ap.add_argument("--steps", type=int, nargs=2, default=[1000, 20000])
    ap.add_argument("--rate-frac", type=float, default=0.8, help="arrival rate as a fraction of capacity")
    ap.add_argument("--rate", type=float, default=None,
                    help="arrival rate per second (default: rate-frac of the capacity measured first)")
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--max-slots", type=int, default=64)
    ap.add_argument("--macro", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1005)
    ap.add_argument("--part", type=int, nargs=2, default=None, metavar=("FIRST", "COUNT"),
                    help="serve requests FIRST .. FIRST+COUNT-1 of the stream, arrivals shifted to start at "
                         "the first one's (a stream longer than one run can hold is served in parts)")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    budget = int(a.budget_gib * (1 << 30))
    cap = measure_capacity(a.sizes, a.steps, budget, a.max_slots)
    mean_work = np.mean([N ** 3 for N in a.sizes]) * (a.steps[0] + a.steps[1]) / 2.0
    rate = a.rate if a.rate else a.rate_frac * cap / mean_work
    print(f"capacity {cap / 1e6:.0f} Mcell-steps/s; arrival rate {rate:.3f}/s", flush=True)
    reqs = make_requests(a.requests, rate, tuple(a.sizes), tuple(a.steps), a.seed)
    if a.part:
        first, count = a.part
        reqs = reqs[first:first + count]
        t0 = reqs[0].arrival
        for r in reqs:
            r.arrival -= t0
    be = GpuBackend(a.max_slots, budget)
    srv = Server(be, macro=a.macro)
    done, wall = srv.run(reqs, log=lambda m: print(m, flush=True))
    be.close()
    out = summarize(done, wall)
    out.update({"requests": a.requests, "rate_per_s": rate, "capacity_mcell_steps_per_s": round(cap / 1e6, 1),
                "sizes": a.sizes, "steps": a.steps, "rate_frac": a.rate_frac, "part": a.part,
                "latency_s": {str(r.rid): round(r.completed - r.arrival, 4) for r in done if r.ok}})
    print(json.dumps(out), flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
