"""Continuous-batching request loop over an Engine.

A request is one independent MPM run: a Problem advanced for a requested
number of steps (or to its endtime).  The server keeps every admitted run
resident in one device pool and advances all of them together a few steps
at a time (one launch per phase and shape group, optionally graph-replayed).
After every chunk it retires the runs that finished (their final state is
downloaded and handed back) and admits waiting requests, in arrival order,
as soon as pool space (particle and node ranges, instance slots) is free:
new runs join while others are on the GPU.  Admission is bounded by the
engine's HBM budget.

Time to solution of a request = completion time - arrival time, with the
completion time taken when the host sees the run finished after a chunk.
"""
import collections
import math
import time

import numpy as np

from ._lib import MpmError


class Request:
    __slots__ = ("rid", "problem", "steps", "arrival", "admitted", "finished", "iid", "result", "n", "kind",
                 "final_step")

    def __init__(self, rid, problem, steps, arrival=0.0, kind=""):
        self.rid = rid
        self.problem = problem
        self.steps = int(steps)
        self.arrival = arrival
        self.admitted = None
        self.finished = None
        self.iid = None
        self.result = None
        self.n = problem.n
        self.kind = kind
        self.final_step = 0


class Server:
    """engine: an Engine; chunk: steps per launch round; graph: replay the
    chunk from a captured HIP graph ("lazy": capture only once the resident
    set has stayed the same for a round, since a capture costs several plain
    rounds and admissions change the set often); keep_state: download the
    final particle state of every run (else only its scalars); pipeline:
    retire from a status snapshot taken after the previous round, so the host
    waits for round k-1 while round k runs (finished runs are seen one round
    later; the device is never idle between rounds; needs keep_state off)."""

    def __init__(self, engine, chunk=8, graph="lazy", keep_state=False, clock=time.perf_counter, pipeline=False):
        self.eng = engine
        self.chunk = int(chunk)
        self.graph = graph
        self.keep_state = keep_state
        self.clock = clock
        self.pipeline = bool(pipeline) and not keep_state
        self._snap_ids = None      # running ids covered by the pending status snapshot
        self.queue = collections.deque()
        self.running = {}          # iid -> Request
        self.done = []
        self.failed = []           # runs evicted because their state went non-finite
        self.max_running = 0
        self.steps_issued = 0

    def submit(self, req):
        self.queue.append(req)

    def admit(self, now):
        """First fit in arrival order: a request that does not fit now stays
        queued without blocking smaller ones behind it."""
        admitted = 0
        if not self.queue:
            return 0
        free_p, free_n, free_i = self.eng.capacity()
        for req in list(self.queue):
            if free_i <= 0:
                break
            # largest free ranges: skip what cannot fit without a failing call
            if req.n > free_p or req.problem.nb_node > free_n:
                continue
            try:
                iid = self.eng.add(req.problem)
            except MpmError as e:
                if e.code == -6:   # MPM_ERR_NOMEM: pool full for now
                    continue
                raise
            self.eng.set_step_limit(iid, req.steps)
            req.iid = iid
            req.admitted = now
            self.running[iid] = req
            self.queue.remove(req)
            admitted += 1
            free_p, free_n, free_i = self.eng.capacity()
        self.max_running = max(self.max_running, len(self.running))
        return admitted

    def advance(self):
        if not self.running:
            return 0
        self.eng.step(self.chunk, graph=self.graph, graph_steps=self.chunk)
        self.steps_issued += self.chunk
        return self.chunk

    def retire(self, now_fn):
        if self.pipeline:
            return self._retire_pipelined(now_fn)
        if not self.running:
            return 0
        ids = list(self.running)
        st = self.eng.status(ids)
        out = 0
        # a run whose stress update went non-finite is evicted and reported;
        # its NaNs stay in its own particle and node ranges, so the others
        # go on unaffected
        for iid, (negvol, nonfinite) in self.eng.numeric_flags(ids).items():
            if nonfinite > 0:
                req = self.running.pop(iid)
                req.final_step = st[iid][1]
                req.result = {"error": "non-finite state", "nonfinite": nonfinite, "negvol": negvol,
                              "scalars": self.eng.scalars(iid)}
                self.eng.remove(iid)
                req.finished = now_fn()
                self.failed.append(req)
        ids = list(self.running)
        for iid in ids:
            done, istep = st[iid]
            if done != 1:
                continue
            req = self.running.pop(iid)
            req.final_step = istep
            req.result = {"scalars": self.eng.scalars(iid)}
            if self.keep_state:
                req.result["particles"] = self.eng.particles(iid)
            self.eng.remove(iid)
            req.finished = now_fn()
            self.done.append(req)
            out += 1
        return out

    def _retire_pipelined(self, now_fn):
        out = 0
        if self._snap_ids:
            snap = self.eng.snapshot_read(self._snap_ids)   # waits for the previous round only
            for iid, (done, sc) in snap.items():
                if done != 1 or iid not in self.running:
                    continue
                req = self.running.pop(iid)
                req.final_step = int(sc[0])
                req.result = {"scalars": sc}
                self.eng.remove(iid)
                req.finished = now_fn()
                self.done.append(req)
                out += 1
        self._snap_ids = None
        if self.running:
            self.eng.snapshot()                               # after the round just queued
            self._snap_ids = list(self.running)
        return out

    def idle(self):
        return not self.running and not self.queue


def poisson_arrivals(rate, duration, rng):
    """Arrival times of a Poisson process of the given rate on [0, duration)."""
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append(t)


def run_poisson(server, pool, rate, duration, seed=0, drain=True, max_wall=None):
    """Drive the server with Poisson arrivals of requests drawn from pool
    (a list of (problem, steps, kind)) for `duration` seconds of wall time;
    with drain, keep stepping until every arrived request has finished.
    Returns the statistics of the completed requests."""
    rng = np.random.default_rng(seed)
    arr = poisson_arrivals(rate, duration, rng)
    picks = rng.integers(0, len(pool), size=len(arr))
    clock = server.clock
    t0 = clock()
    now = lambda: clock() - t0
    nxt = 0
    limit = max_wall if max_wall is not None else 10 * duration + 60
    while True:
        t = now()
        while nxt < len(arr) and arr[nxt] <= t:
            prob, steps, kind = pool[picks[nxt]]
            server.submit(Request(nxt, prob, steps, arrival=arr[nxt], kind=kind))
            nxt += 1
        server.admit(t)
        if server.running:
            server.advance()
            server.retire(now)
        elif nxt < len(arr):
            time.sleep(max(0.0, min(arr[nxt] - now(), 0.01)))
        elif server.idle() or not drain:
            break
        if now() > limit:
            break
    wall = now()
    return summarize(server.done, wall, rate, duration, len(arr), server)


def summarize(done, wall, rate, duration, arrived, server):
    tts = np.array([r.finished - r.arrival for r in done]) if done else np.zeros(0)
    wait = np.array([r.admitted - r.arrival for r in done]) if done else np.zeros(0)
    psteps = float(sum(r.n * r.final_step for r in done))
    first = min((r.arrival for r in done), default=0.0)
    last = max((r.finished for r in done), default=0.0)
    span = max(last - first, 1e-9)
    pct = lambda a, q: float(np.percentile(a, q)) if len(a) else None
    return {
        "rate_req_per_s": rate,
        "window_s": duration,
        "arrived": arrived,
        "completed": len(done),
        "p50_tts_s": pct(tts, 50),
        "p99_tts_s": pct(tts, 99),
        "mean_tts_s": float(tts.mean()) if len(tts) else None,
        "p50_queue_s": pct(wait, 50),
        "p99_queue_s": pct(wait, 99),
        "particle_steps": psteps,
        "particle_steps_per_s": psteps / span,
        "max_concurrent": server.max_running,
        "failed": len(getattr(server, "failed", ())),
        "wall_s": wall,
        "kinds": dict(collections.Counter(r.kind for r in done)),
    }


def request_pool(mpm, decks, count=64, seed=0, slope_frac=0.7, steps=(100, 300), slope_n=(2e4, 4e5),
                 bar_radius=(2.0, 5.0), bar_length=32.4):
    """Mixed request pool: slopes as in the batched ensemble (random size in
    slope_n particles and random strength) and Taylor bars (impact speed
    U(100, 250) m/s, radius U(bar_radius) mm, length bar_length mm); every
    request asks for a step count drawn from U(steps)."""
    rng = np.random.default_rng(seed)
    n_slope = int(round(slope_frac * count))
    pool = []
    ens = decks.slope_ensemble(n_slope, seed=seed + 1, n_min=slope_n[0], n_max=slope_n[1])
    for k, (text, n, meta) in enumerate(ens):
        p = decks.load(text, mpm, "req_slope%d" % k)
        pool.append((p, int(rng.integers(steps[0], steps[1] + 1)), "slope"))
    for k in range(count - n_slope):
        text, n = decks.taylor_bar(radius=float(rng.uniform(*bar_radius)), speed=float(rng.uniform(100.0, 250.0)),
                                   length=bar_length)
        p = decks.load(text, mpm, "req_bar%d" % k)
        pool.append((p, int(rng.integers(steps[0], steps[1] + 1)), "bar"))
    order = rng.permutation(len(pool))
    return [pool[i] for i in order]
