"""Host logic of the continuous-batching server (no GPU): admission by free
pool space in arrival order with first fit, retirement of finished runs,
per-request step limits, Poisson arrivals and the reported statistics.  The
engine is replaced by a stand-in with the same interface that models pool
capacity and the device-side step limit."""
import numpy as np

from conftest import load_package


class FakeProblem:
    def __init__(self, n, nb_node=10):
        self.n = n
        self.nb_node = nb_node


class FakeEngine:
    def __init__(self, cap_p, cap_i):
        self.mpm = load_package()
        self.cap_p, self.cap_i = cap_p, cap_i
        self.live = {}
        self.next = 0
        self.log = []

    def capacity(self):
        return self.cap_p - sum(v[0] for v in self.live.values()), 10 ** 9, self.cap_i - len(self.live)

    def add(self, prob):
        if prob.n > self.capacity()[0] or len(self.live) >= self.cap_i:
            raise self.mpm.MpmError(-6, "mpm_engine_add", "pool full")
        iid = self.next
        self.next += 1
        self.live[iid] = [prob.n, 0, 0]
        self.log.append(("add", prob.n))
        return iid

    def set_step_limit(self, iid, s):
        self.live[iid][2] = s

    def step(self, n, graph=False, graph_steps=1):
        for v in self.live.values():
            v[1] = min(v[1] + n, v[2]) if v[2] else v[1] + n

    def status(self, ids):
        return {i: (int(self.live[i][2] and self.live[i][1] >= self.live[i][2]), self.live[i][1]) for i in ids}

    def scalars(self, iid):
        return np.array([self.live[iid][1], 0, 0, 0, 0, 0, 0, 0], dtype=float)

    def particles(self, iid):
        return np.zeros((1, self.live[iid][0]))

    def remove(self, iid):
        self.log.append(("remove", self.live.pop(iid)[0]))

    bad = ()   # instance ids whose state the stand-in reports non-finite

    def numeric_flags(self, ids):
        return {i: (0, 3 if i in self.bad else 0) for i in ids}

    # status snapshot: the state after the work issued so far (the stand-in's
    # steps complete at once)
    def snapshot(self):
        self.snap = {i: (int(v[2] and v[1] >= v[2]), v[1]) for i, v in self.live.items()}

    def snapshot_read(self, ids):
        return {i: (self.snap[i][0] if i in self.live else -1,
                    np.array([self.snap[i][1], 0, 0, 0, 0, 0, 0, 0], dtype=float)) for i in ids}


def test_first_fit_admission_and_retirement():
    serve = load_package().serve
    eng = FakeEngine(cap_p=1000, cap_i=3)
    srv = serve.Server(eng, chunk=4, graph=False, clock=lambda: 0.0)
    sizes, steps = [600, 500, 300, 200], [16, 8, 8, 8]
    for k, (n, s) in enumerate(zip(sizes, steps)):
        srv.submit(serve.Request(k, FakeProblem(n), s))
    srv.admit(0.0)
    # 600 fits, 500 does not (400 left), 300 fits, 200 does not (100 left)
    assert sorted(r.n for r in srv.running.values()) == [300, 600]
    assert [r.n for r in srv.queue] == [500, 200]
    t = [0.0]
    rounds = 0
    while not srv.idle():
        t[0] += 1.0
        srv.admit(t[0])
        srv.advance()
        srv.retire(lambda: t[0])
        rounds += 1
        assert rounds < 50
    assert sorted(r.rid for r in srv.done) == [0, 1, 2, 3]
    assert all(r.final_step == r.steps for r in srv.done)
    # the 300 run finishes after 2 chunks and frees room for 200, not 500
    order = [e for e in eng.log if e[0] == "add"]
    assert order[:3] == [("add", 600), ("add", 300), ("add", 200)]
    assert srv.max_running == 2


def test_poisson_arrivals_and_summary():
    serve = load_package().serve
    rng = np.random.default_rng(3)
    arr = serve.poisson_arrivals(100.0, 10.0, rng)
    assert abs(len(arr) - 1000) < 4 * np.sqrt(1000)
    assert np.all(np.diff(arr) > 0) and arr[-1] < 10.0
    done = []
    for k in range(100):
        r = serve.Request(k, FakeProblem(10 + k), 5, arrival=0.1 * k)
        r.admitted = r.arrival + 0.01
        r.finished = r.arrival + 0.5 + 0.01 * k
        r.final_step = 5
        done.append(r)

    class S:
        max_running = 7

    st = serve.summarize(done, 20.0, 5.0, 10.0, 100, S())
    tts = np.array([0.5 + 0.01 * k for k in range(100)])
    assert st["completed"] == 100
    assert abs(st["p50_tts_s"] - np.percentile(tts, 50)) < 1e-12
    assert abs(st["p99_tts_s"] - np.percentile(tts, 99)) < 1e-12
    assert st["particle_steps"] == sum((10 + k) * 5 for k in range(100))
    assert st["max_concurrent"] == 7


def test_pipelined_retirement_sees_each_round_one_round_later():
    """Pipelined serving retires from the snapshot taken after the previous
    round: every run still finishes with its own step count, one round later
    than the synchronous loop sees it, and never twice."""
    serve = load_package().serve
    results = {}
    for pipe in (False, True):
        eng = FakeEngine(cap_p=1000, cap_i=3)
        srv = serve.Server(eng, chunk=4, graph=False, clock=lambda: 0.0, pipeline=pipe)
        for k, (n, s) in enumerate(zip([600, 500, 300, 200], [16, 8, 8, 8])):
            srv.submit(serve.Request(k, FakeProblem(n), s))
        t = [0.0]
        rounds = 0
        while not srv.idle():
            t[0] += 1.0
            srv.admit(t[0])
            if srv.running:
                srv.advance()
            srv.retire(lambda: t[0])
            rounds += 1
            assert rounds < 50
        assert sorted(r.rid for r in srv.done) == [0, 1, 2, 3]
        assert all(r.final_step == r.steps for r in srv.done)
        assert len(eng.live) == 0
        results[pipe] = {r.rid: r.finished for r in srv.done}
    # the first run to finish (rid 2: 8 steps from round 1) is seen in round 2
    # synchronously and in round 3 pipelined
    assert results[False][2] == 2.0 and results[True][2] == 3.0


def test_a_non_finite_run_is_evicted_and_reported():
    serve = load_package().serve
    eng = FakeEngine(cap_p=1000, cap_i=4)
    eng.bad = (1,)
    srv = serve.Server(eng, chunk=4, graph=False, clock=lambda: 0.0)
    for k, n in enumerate([100, 200, 300]):
        srv.submit(serve.Request(k, FakeProblem(n), 12))
    srv.admit(0.0)
    while not srv.idle():
        srv.advance()
        srv.retire(lambda: 0.0)
    assert [r.rid for r in srv.failed] == [1]
    assert srv.failed[0].result["error"] == "non-finite state" and srv.failed[0].result["nonfinite"] == 3
    assert sorted(r.rid for r in srv.done) == [0, 2] and all(r.final_step == 12 for r in srv.done)
    assert ("remove", 200) in eng.log
    st = serve.summarize(srv.done, 1.0, 1.0, 1.0, 3, srv)
    assert st["failed"] == 1 and st["completed"] == 2
