"""Serving loop: scheduler logic on the CPU (simulated clock and backend),
and on the GPU every request's result in the mix equals the request served
alone (bitwise)."""
import numpy as np
import pytest

from linqt_amd import serve


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeBackend:
    """1D jobs advance `chunk` steps per step_1d (1 ms each); 2D jobs take M us."""

    def __init__(self, clock, chunk=32):
        self.clock = clock
        self.chunk = chunk
        self.left = {}
        self.used = 0
        self.peak = 0
        self.starts = []

    def footprint(self, r):
        return r.M * r.R * (100 if r.kind == "dos" else 400)

    def start_1d(self, r):
        self.left[r.rid] = r.M - 1
        self.used += self.footprint(r)
        self.peak = max(self.peak, self.used)
        self.starts.append(r.rid)

    def step_1d(self):
        self.clock.t += 1e-3
        done = []
        for rid in list(self.left):
            self.left[rid] -= self.chunk
            if self.left[rid] <= 0:
                done.append(rid)
                del self.left[rid]
        return done

    def collect_1d(self, r):
        self.used -= self.footprint(r)
        return ("dos", r.rid)

    def run_2d(self, r):
        self.starts.append(r.rid)
        self.peak = max(self.peak, self.used + self.footprint(r))
        self.clock.t += r.M * 1e-6
        return ("2d", r.rid)


def test_poisson_arrivals_rate_and_determinism():
    t = serve.poisson_arrivals(20000, 200.0, seed=5)
    assert np.all(np.diff(t) > 0)
    assert abs(len(t) / t[-1] - 200.0) < 5.0
    assert np.array_equal(t, serve.poisson_arrivals(20000, 200.0, seed=5))


def test_scheduler_completes_all_within_budget():
    reqs = serve.request_mix(300, 200.0, seed=3, systems_1d=["a", "b"], systems_2d=["c"])
    clock = FakeClock()
    be = FakeBackend(clock)
    budget = max(be.footprint(r) for r in reqs) * 3
    sched = serve.Scheduler(be, budget, clock=clock, sleep=clock.sleep)
    out = sched.run(reqs)
    assert all(r.finish is not None and r.result is not None for r in out)
    assert all(r.start >= r.arrival - 1e-12 and r.finish >= r.start for r in out)
    assert be.peak <= budget
    # 2D requests start in arrival order
    twod = [rid for rid in be.starts if out[rid].kind != "dos"]
    assert twod == sorted(twod, key=lambda i: (out[i].arrival, i))
    s = serve.summary(out)
    assert s["requests"] == 300 and s["latency_ms"]["p50"] <= s["latency_ms"]["p99"]


def test_scheduler_rejects_oversized_request():
    reqs = serve.request_mix(5, 200.0, seed=1, systems_1d=["a"], systems_2d=[])
    clock = FakeClock()
    be = FakeBackend(clock)
    with pytest.raises(ValueError):
        serve.Scheduler(be, 10, clock=clock, sleep=clock.sleep).run(reqs)


@pytest.mark.gpu
def test_served_results_equal_alone(ctx):
    import linqt_amd as lq
    from linqt_amd import lattices

    systems = {}
    for name, Hh, a in (("sq24", lattices.square_anderson(24, W=1.0, seed=1), 4.6),
                        ("cub10", lattices.cubic_anderson(10, W=2.0, seed=2), 7.2),
                        ("km8", lattices.kane_mele(8, W=1.0, seed=3), 4.6)):
        systems[name] = (lq.SparseMatrix.from_csr(ctx, Hh).Rescale(1 / a, 0.0), None, None, Hh.n, 2 * a, 0.0)
    Hh, Vx, Vy = lattices.haldane_honeycomb(8, seed=4)
    systems["hc8"] = (lq.SparseMatrix.from_csr(ctx, Hh).Rescale(1 / 4.6, 0.0), lq.SparseMatrix.from_csr(ctx, Vx),
                      lq.SparseMatrix.from_csr(ctx, Vy), Hh.n, 9.2, 0.0)
    reqs = serve.request_mix(16, 2000.0, seed=7, systems_1d=["sq24", "cub10", "km8"], systems_2d=["hc8"],
                             M_1d=(16, 64, 200), M_2d=(16, 32), R=(1, 3, 8))
    be = serve.GpuBackend(lq, ctx, systems, chunk=8)
    out = serve.Scheduler(be, 1 << 34).run(reqs)
    assert any(r.kind != "dos" for r in out) and any(r.kind == "dos" for r in out)
    for r in out:
        ref = be.alone(r)
        assert np.array_equal(r.result[0], ref[0]), (r.rid, r.kind)
        for a, b in zip(r.result[1], ref[1]):
            assert np.array_equal(a, b), (r.rid, r.kind)
    be.close()
