"""The continuous-batching scheduler (swalbe.serve.Server) on a simulated
device: admission order, footprint limits, step budgets, overlap of
requests and latency accounting.  No GPU."""
import numpy as np
import pytest

from swalbe import serve
from swalbe._lib import ENOMEM, SwalbeError


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeBackend:
    """Device stand-in: films are counters of steps done; time advances by
    `sec_per_lup` per lattice update of every resident unfinished film."""

    def __init__(self, clock, capacity, max_films=64, sec_per_lup=1e-9, fail_adds=0):
        self.clock, self.capacity, self.max_films = clock, capacity, max_films
        self.sec_per_lup = sec_per_lup
        self.films = {}
        self.used = 0
        self.peak = 0
        self.fail_adds = fail_adds
        self.log = []

    def footprint(self, r):
        return 100 * r.Lx * r.Ly

    def bytes_free(self):
        return self.capacity - self.used

    def add(self, r):
        if self.fail_adds and self.films:  # an empty arena is one free block
            self.fail_adds -= 1
            raise SwalbeError("fragmented", ENOMEM)
        need = self.footprint(r)
        assert self.used + need <= self.capacity, "admitted beyond the arena"
        assert len(self.films) < self.max_films
        slot = next(i for i in range(self.max_films) if i not in self.films)
        self.films[slot] = {"r": r, "done": 0}
        self.used += need
        self.peak = max(self.peak, self.used)
        self.log.append(("add", r.rid, self.clock()))
        return slot

    def step(self, n):
        lups = 0
        for f in self.films.values():
            k = min(n, f["r"].nsteps - f["done"])
            f["done"] += k
            lups += k * f["r"].Lx * f["r"].Ly
        self.clock.t += lups * self.sec_per_lup

    def wait(self, steps_behind):
        pass

    def get(self, slot, name):
        f = self.films[slot]
        return np.full((f["r"].Lx, f["r"].Ly), float(f["done"]))

    def remove(self, slot):
        f = self.films.pop(slot)
        self.used -= self.footprint(f["r"])
        self.log.append(("remove", f["r"].rid, self.clock()))


def _requests(n=40, seed=3, rate=200.0):
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        Lx = int(rng.choice([64, 128, 256, 512]))
        T = int(rng.choice([100, 1000, 5000]))
        out.append(serve.Request(k, Lx, Lx // 2, None, None, T, arrival=t))
    return out


def _run(reqs, capacity=100 * 512 * 256 * 3, tick=500, **kw):
    clock = FakeClock()
    be = FakeBackend(clock, capacity, **kw)
    srv = serve.Server(be, tick=tick, clock=clock, sleep=clock.sleep)
    done = srv.serve(reqs)
    return srv, be, done


def test_every_request_gets_exactly_its_budget():
    reqs = _requests()
    srv, be, done = _run(reqs)
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
    for r in done:
        assert np.all(r.result["height"] == r.nsteps), r.rid
    assert not be.films and be.used == 0


def test_fifo_admission_never_before_arrival():
    reqs = _requests()
    srv, be, done = _run(reqs)
    adds = [(rid, t) for ev, rid, t in be.log if ev == "add"]
    by_arrival = [r.rid for r in sorted(reqs, key=lambda r: r.arrival)]
    assert [rid for rid, _ in adds] == by_arrival
    for r in done:
        assert r.t_admit >= r.arrival - 1e-12
        assert r.t_done >= r.t_admit


def test_footprint_limit_and_overlap():
    reqs = _requests(60)
    cap = 100 * 512 * 256 * 2
    srv, be, done = _run(reqs, capacity=cap)
    assert be.peak <= cap
    assert srv.max_resident > 1  # requests run together, not one at a time
    assert len(done) == len(reqs)


def test_continuous_admission_while_others_run():
    """A request that arrives while another is mid-run is admitted before
    the first one finishes."""
    a = serve.Request(0, 128, 64, None, None, 20_000, arrival=0.0)
    b = serve.Request(1, 128, 64, None, None, 1_000, arrival=1e-4)
    srv, be, done = _run([a, b], tick=100, sec_per_lup=1e-9)
    ta = {rid: t for ev, rid, t in be.log if ev == "add"}
    tr = {rid: t for ev, rid, t in be.log if ev == "remove"}
    assert ta[1] < tr[0]
    assert tr[1] < tr[0]  # the short request finishes first


def test_oversized_request_is_rejected():
    r = serve.Request(0, 4096, 4096, None, None, 10, arrival=0.0)
    with pytest.raises(RuntimeError, match="does not fit"):
        _run([r], capacity=1000)


def test_zero_step_request_and_fragmented_arena():
    reqs = [serve.Request(0, 64, 32, None, None, 300, arrival=0.0),
            serve.Request(1, 64, 32, None, None, 0, arrival=0.0)]
    srv, be, done = _run(reqs, fail_adds=1)
    assert sorted(r.rid for r in done) == [0, 1]
    res = {r.rid: r.result["height"] for r in done}
    assert np.all(res[0] == 300) and np.all(res[1] == 0)


def test_latency_stats():
    reqs = _requests(30)
    srv, be, done = _run(reqs)
    st = serve.latency_stats(done)
    assert st["requests"] == 30
    assert 0 < st["p50_s"] <= st["p99_s"]
    assert st["solves_per_s"] == pytest.approx(30 / st["makespan_s"])


def test_request_generators_follow_the_workload_spec():
    sweep = serve.coalescence_sweep(64, seed=2, nsteps=10)
    assert all(r.Lx % 64 == 0 and 128 <= r.Lx <= 1024 and r.Ly == r.Lx // 2 for r in sweep)
    assert len({r.init.seed for r in sweep}) == 64
    for r in sweep:
        assert 0.005 <= r.param.γ <= 0.02 and 1 / 18 <= r.param.θ <= 1 / 6
        R = r.init.r1
        assert 0.15 * r.Ly <= R <= 0.3 * r.Ly and r.init.r2 == R
        # tangent caps: centres 2 R sin(πθ) apart on the x mid-line
        assert r.init.c2x - r.init.c1x == pytest.approx(2 * R * np.sin(np.pi * r.init.theta1))
        assert r.init.c1y == r.init.c2y
    reqs = serve.poisson_requests(400, rate=200.0, seed=4)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert gaps.mean() == pytest.approx(1 / 200, rel=0.15)
    assert all(r.Ly in (r.Lx, r.Lx // 2) and r.nsteps in (1000, 10000, 50000) for r in reqs)
    assert serve.domain_seed(1, 0) != serve.domain_seed(1, 1) != serve.domain_seed(2, 1)
