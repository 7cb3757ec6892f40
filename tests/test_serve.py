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
    """1D jobs advance `chunk` steps per step_1d (1 ms each); a 2D job takes
    `t2d` seconds per order M on one of `slots_2d` workers, concurrently with
    the 1D steps."""

    def __init__(self, clock, chunk=32, slots_2d=2, t2d=1e-6):
        self.clock = clock
        self.chunk = chunk
        self.slots_2d = slots_2d
        self.t2d = t2d
        self.left = {}
        self.used = 0
        self.peak = 0
        self.starts = []
        self.running_2d = {}  # rid -> (finish time, footprint)
        self.concurrent_2d = 0

    def footprint(self, r):
        return r.M * r.R * (100 if r.kind == "dos" else 400)

    def _use(self, b):
        self.used += b
        self.peak = max(self.peak, self.used)

    def start_1d(self, r):
        self.left[r.rid] = r.M - 1
        self._use(self.footprint(r))
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

    def submit_2d(self, r):
        assert len(self.running_2d) < self.slots_2d
        self.starts.append(r.rid)
        self._use(self.footprint(r))
        self.running_2d[r.rid] = (self.clock.t + r.M * self.t2d, self.footprint(r))
        self.concurrent_2d = max(self.concurrent_2d, len(self.running_2d))

    def poll_2d(self):
        done = [rid for rid, (t, _) in self.running_2d.items() if t <= self.clock.t]
        out = []
        for rid in done:
            self.used -= self.running_2d.pop(rid)[1]
            out.append((rid, ("2d", rid)))
        return out

    def wait_2d(self, timeout):
        target = self.clock.t + timeout
        if self.running_2d:
            target = min(target, min(t for t, _ in self.running_2d.values()))
        self.clock.t = max(self.clock.t, target)


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
    assert sched.reserved == 2 * max(be.footprint(r) for r in reqs if r.kind != "dos")
    # 2D requests start in arrival order (each fits the budget next to the batch here)
    twod = [rid for rid in be.starts if out[rid].kind != "dos"]
    assert twod == sorted(twod, key=lambda i: (out[i].arrival, i))
    assert be.concurrent_2d == 2
    s = serve.summary(out)
    assert s["requests"] == 300 and s["latency_ms"]["p50"] <= s["latency_ms"]["p99"]


def test_long_2d_solves_do_not_block_the_batch():
    """DOS requests arriving while 2D solves of a whole second run are admitted
    and finish during them; a 2D request arriving while another runs starts on
    the free worker at once."""
    reqs = [serve.Request(0, "bastin", "c", 1000, 1, 0.0), serve.Request(1, "greenwood", "c", 1000, 1, 0.1)]
    reqs += [serve.Request(2 + i, "dos", "a", 256, 1, 0.2 + 0.01 * i) for i in range(20)]
    clock = FakeClock()
    be = FakeBackend(clock, t2d=1e-3)  # 1 s per 2D solve
    out = serve.Scheduler(be, 10 ** 12, clock=clock, sleep=clock.sleep).run(reqs)
    assert abs(out[1].start - 0.1) < 2e-3  # admitted on arrival, beside request 0
    assert out[0].finish >= 1.0 and out[1].finish >= 1.1
    for r in out[2:]:
        assert r.start - r.arrival < 2e-3 and r.finish < 0.5, (r.rid, r.start, r.finish)


def test_scheduler_rejects_oversized_request():
    reqs = serve.request_mix(5, 200.0, seed=1, systems_1d=["a"], systems_2d=[])
    clock = FakeClock()
    be = FakeBackend(clock)
    with pytest.raises(ValueError):
        serve.Scheduler(be, 10, clock=clock, sleep=clock.sleep).run(reqs)


def _systems(lq, ctx, spec):
    """name -> (H~, V_x, V_y, n, W, Ec) on context ctx."""
    from linqt_amd import lattices

    out = {}
    for name, kind, L, seed in spec:
        if kind == "sq":
            Hh, a = lattices.square_anderson(L, W=1.0, seed=seed), 4.6
        elif kind == "cub":
            Hh, a = lattices.cubic_anderson(L, W=2.0, seed=seed), 7.2
        elif kind == "km":
            Hh, a = lattices.kane_mele(L, W=1.0, seed=seed), 4.6
        if kind != "hc":
            out[name] = (lq.SparseMatrix.from_csr(ctx, Hh).Rescale(1 / a, 0.0), None, None, Hh.n, 2 * a, 0.0)
        else:
            Hh, Vx, Vy = lattices.haldane_honeycomb(L, seed=seed)
            out[name] = (lq.SparseMatrix.from_csr(ctx, Hh).Rescale(1 / 4.6, 0.0), lq.SparseMatrix.from_csr(ctx, Vx),
                         lq.SparseMatrix.from_csr(ctx, Vy), Hh.n, 9.2, 0.0)
    return out


def _served_equal_alone(lq, ctx, spec, reqs, chunk):
    ctxs = [lq.Context(0), lq.Context(0)]
    be = serve.GpuBackend(lq, ctx, _systems(lq, ctx, spec), chunk=chunk,
                          workers_2d=[(c, _systems(lq, c, spec)) for c in ctxs])
    try:
        out = serve.Scheduler(be, 1 << 36).run(reqs)
        assert any(r.kind != "dos" for r in out) and any(r.kind == "dos" for r in out)
        for r in out:
            ref = be.alone(r)
            assert np.array_equal(r.result[0], ref[0]), (r.rid, r.kind)
            for a, b in zip(r.result[1], ref[1]):
                assert np.array_equal(a, b), (r.rid, r.kind)
        return out
    finally:
        be.close()
        for c in ctxs:
            c.close()


@pytest.mark.gpu
def test_served_results_equal_alone(ctx):
    import linqt_amd as lq

    spec = [("sq24", "sq", 24, 1), ("cub10", "cub", 10, 2), ("km8", "km", 8, 3), ("hc8", "hc", 8, 4)]
    reqs = serve.request_mix(16, 2000.0, seed=7, systems_1d=["sq24", "cub10", "km8"], systems_2d=["hc8"],
                             M_1d=(16, 64, 200), M_2d=(16, 32), R=(1, 3, 8))
    _served_equal_alone(lq, ctx, spec, reqs, chunk=8)


@pytest.mark.gpu
@pytest.mark.slow
def test_served_mix_at_serving_sizes_with_overlapping_2d(ctx):
    """N 16k-33k, M 256-512, Greenwood and Bastin requests, two of them arriving
    1 ms apart so the second starts while the first is on the GPU (two 2D
    workers); every result bitwise equal to the request served alone."""
    import linqt_amd as lq

    spec = [("cub32", "cub", 32, 11), ("km64", "km", 64, 12), ("hc96", "hc", 96, 13)]
    reqs = [serve.Request(0, "bastin", "hc96", 256, 1, 0.0, seed=1), serve.Request(1, "greenwood", "hc96", 256, 2, 0.001, seed=2)]
    more = serve.request_mix(10, 500.0, seed=9, systems_1d=["cub32", "km64"], systems_2d=["hc96"],
                             M_1d=(256, 512), M_2d=(256,), R=(1, 2, 4))
    for i, r in enumerate(more):
        reqs.append(serve.Request(2 + i, r.kind, r.system, r.M, r.R, 0.002 + r.arrival, seed=r.seed))
    out = _served_equal_alone(lq, ctx, spec, reqs, chunk=32)
    a, b = out[0], out[1]
    assert b.start < a.finish, "the second 2D request started while the first was running"
    print("[case] served latencies ms:", [round(r.latency * 1e3, 1) for r in out], flush=True)
