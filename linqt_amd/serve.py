"""Serving loop for moment requests on one GPU (continuous batching).

Requests arrive over time (a seeded Poisson stream in the benchmark):

* ``dos``: SpectralMoments (chebyshev_solver.cpp:176-200) + Jackson + DOS on
  an energy grid -- these join the batched engine (``Batch``) at the next
  chunk boundary, so many of them advance together;
* ``greenwood`` / ``bastin``: Kubo_solver_FFT::compute (Kubo_solver_FFT.cpp)
  + post-processing -- handed to the backend's 2D workers (each its own
  context, i.e. its own HIP stream and buffers), so their recursion and
  contraction kernels run on the device beside the DOS batch's launches and
  beside each other, and the loop keeps admitting arrivals and stepping the
  batch while they run.

Admission is by device footprint: a request starts only when the bytes it
holds while running fit the server's HBM budget next to what is already
running (and, for 2D requests, when a worker is free); otherwise it waits in
arrival order (first fit, so a small request may pass a large one).  Every
request's result is what the same request gives when served alone: the
batched engine's moments are bitwise independent of the mix, and a 2D solve
runs the same kernels on the same inputs on whichever context serves it.

The scheduler (``Scheduler``) only sees a backend with ``footprint``,
``start_1d``, ``step_1d``, ``collect_1d``, ``submit_2d``, ``poll_2d``,
``wait_2d`` and ``slots_2d``; ``GpuBackend`` implements them on the C ABI.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Callable

import numpy as np


@dataclasses.dataclass
class Request:
    rid: int
    kind: str              # "dos", "greenwood", "bastin"
    system: str            # key of the operator set in the backend's registry
    M: int
    R: int
    arrival: float         # seconds from the start of the run
    seed: int = 0
    nump: int = 0          # energies (0: 2 M for 2D, 2048 for DOS)
    start: float | None = None
    finish: float | None = None
    result: object = None

    @property
    def moments(self) -> int:
        return self.M * self.R if self.kind == "dos" else self.M * self.M * self.R

    @property
    def latency(self) -> float | None:
        return None if self.finish is None else self.finish - self.arrival


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times of n requests with exponential gaps of mean 1/rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def request_mix(n: int, rate: float, seed: int, systems_1d: list[str], systems_2d: list[str],
                frac_2d: float = 0.3, M_1d=(256, 512, 1024, 2048, 4096), M_2d=(128, 256), R=(1, 2, 4, 8),
                sizes: dict | None = None, max_2d_work: float | None = None):
    """n seeded requests: a fraction frac_2d Greenwood/Bastin, the rest DOS.
    With sizes (system -> N) and max_2d_work, a 2D request's M is drawn from
    the orders with M^2 N <= max_2d_work (the smallest order if none)."""
    rng = np.random.default_rng(seed + 1)
    t = poisson_arrivals(n, rate, seed)
    out = []
    for i in range(n):
        if systems_2d and rng.uniform() < frac_2d:
            kind = "greenwood" if rng.uniform() < 0.5 else "bastin"
            system = str(rng.choice(systems_2d))
            orders = list(M_2d)
            if sizes is not None and max_2d_work is not None:
                orders = [m for m in M_2d if m * m * sizes[system] <= max_2d_work] or [min(M_2d)]
            out.append(Request(i, kind, system, int(rng.choice(orders)), int(rng.choice(R[:2])), float(t[i]),
                               seed=seed * 1000 + i))
        else:
            out.append(Request(i, "dos", str(rng.choice(systems_1d)), int(rng.choice(M_1d)),
                               int(rng.choice(R)), float(t[i]), seed=seed * 1000 + i))
    return out


class Scheduler:
    """Arrival-ordered, footprint-admitted continuous batching.

    One loop: admit arrived requests (DOS into the batch, 2D to a free
    worker), advance the batch by one chunk, collect finished 2D solves.  The
    loop never waits on a 2D solve; with nothing to step it sleeps until the
    next arrival or until a 2D solve finishes.  Memory: each 2D worker holds
    the largest 2D footprint of the stream (reserved up front), DOS requests
    share the rest of the budget."""

    def __init__(self, backend, hbm_budget: int, clock: Callable[[], float] | None = None,
                 sleep: Callable[[float], None] | None = None):
        self.b = backend
        self.budget = hbm_budget
        self.clock = clock or time.perf_counter
        self.sleep = sleep or time.sleep
        self.used = 0
        self.log: list[tuple[float, str, int]] = []

    def run(self, requests: list[Request]) -> list[Request]:
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        # every 2D worker keeps the buffers of the largest 2D request it may run
        # (no allocation churn between requests), so that much is reserved per
        # worker up front and DOS requests are admitted into the rest
        fp2d = [self.b.footprint(r) for r in pending if r.kind != "dos"]
        self.reserved = max(fp2d, default=0) * (self.b.slots_2d if fp2d else 0)
        for r in pending:
            fp = self.b.footprint(r)
            if (r.kind == "dos" and fp > self.budget - self.reserved) or fp > self.budget:
                raise ValueError(f"request {r.rid} needs {fp} bytes, above the budget "
                                 f"({self.reserved} reserved for 2D workers)")
        waiting: list[Request] = []
        running_1d: dict[int, Request] = {}
        running_2d: dict[int, Request] = {}
        t0 = self.clock()
        now = lambda: self.clock() - t0

        def finish(r):
            r.finish = now()
            if r.kind == "dos":
                self.used -= self.b.footprint(r)
            self.log.append((r.finish, "finish", r.rid))

        while pending or waiting or running_1d or running_2d:
            while pending and pending[0].arrival <= now():
                waiting.append(pending.pop(0))
            # first-fit admission in arrival order: DOS into the batch, 2D to a free worker
            for r in list(waiting):
                if r.kind == "dos":
                    fp = self.b.footprint(r)
                    if self.used + fp > self.budget - self.reserved:
                        continue
                    self.used += fp
                elif len(running_2d) >= self.b.slots_2d:
                    continue
                waiting.remove(r)
                r.start = now()
                self.log.append((r.start, "start", r.rid))
                if r.kind == "dos":
                    self.b.start_1d(r)
                    running_1d[r.rid] = r
                else:
                    running_2d[r.rid] = r
                    self.b.submit_2d(r)
            if running_1d:
                for rid in self.b.step_1d():
                    r = running_1d.pop(rid)
                    r.result = self.b.collect_1d(r)
                    finish(r)
            for rid, res in self.b.poll_2d():
                r = running_2d.pop(rid)
                r.result = res
                finish(r)
            if not running_1d:
                # nothing to step: sleep until the next arrival or a 2D solve ends
                dt = max(0.0, pending[0].arrival - now()) if pending else None
                if running_2d:
                    self.b.wait_2d(dt if dt is not None else 0.05)
                elif dt is not None and not waiting:
                    self.sleep(dt)
        return requests


def summary(requests: list[Request]) -> dict:
    done = [r for r in requests if r.finish is not None]
    lat = np.array([r.latency for r in done]) * 1e3
    span = max(r.finish for r in done) - min(r.arrival for r in done)
    out = {"requests": len(done), "makespan_s": span, "throughput_req_s": len(done) / span,
           "moments_per_s": sum(r.moments for r in done) / span,
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(lat.max())}}
    for kind in ("dos", "greenwood", "bastin"):
        lk = [r.latency * 1e3 for r in done if r.kind == kind]
        if lk:
            out[f"latency_ms_{kind}"] = {"n": len(lk), "p50": float(np.percentile(lk, 50)),
                                        "p99": float(np.percentile(lk, 99))}
    return out


class GpuBackend:
    """Runs requests on the GPU.  ``systems`` maps a name to (H, V_x, V_y, n, W,
    Ec) with H already rescaled on context ``ctx`` (V may be None for DOS-only
    systems); DOS requests run in a ``Batch`` on ``ctx``.  ``workers_2d`` is a
    list of (context, systems) pairs -- the same systems uploaded on that
    context -- and each runs one Greenwood/Bastin solve at a time on its own
    host thread; with none, 2D solves run on ``ctx`` when submitted."""

    def __init__(self, lq, ctx, systems: dict, chunk: int = 32, hbm_budget: int = 0, workers_2d=None):
        import queue
        import threading

        self.lq = lq
        self.ctx = ctx
        self.systems = systems
        self.batch = lq.Batch(ctx, hbm_budget=hbm_budget, chunk=chunk)
        self.chunk = chunk
        self.jobs: dict[int, int] = {}
        self.workers = list(workers_2d or [])
        self.slots_2d = max(1, len(self.workers))
        self._done_2d: list[tuple[int, object]] = []
        self._cv = threading.Condition()
        self._queues = []
        self._threads = []
        self._busy = [0] * len(self.workers)
        for i, (wctx, wsys) in enumerate(self.workers):
            q = queue.Queue()
            t = threading.Thread(target=self._worker, args=(i, wctx, wsys, q), daemon=True)
            t.start()
            self._queues.append(q)
            self._threads.append(t)

    def states(self, r: Request):
        from .lattices import random_phase_states

        return random_phase_states(self.systems[r.system][3], r.R, seed=r.seed)

    def footprint(self, r: Request) -> int:
        n = self.systems[r.system][3]
        if r.kind == "dos":
            return self.lq.Batch.job_bytes(n, r.R, r.M, self.chunk)
        # the resident 2D pipeline of one operator (sigma_xx) for all R states in
        # one group: M left and M right vectors, phi and two work vectors per
        # state, plus split-K partial tables; a worker's buffers stay at its
        # largest request, which the scheduler reserves, so with the budget
        # below the device's memory every request runs the same grouping, and
        # gives the same result, as when alone
        return (2 * r.M + 3) * n * r.R * 16 + (256 << 20)

    def start_1d(self, r: Request):
        self.jobs[self.batch.submit(self.systems[r.system][0], self.states(r), r.M)] = r.rid

    def step_1d(self) -> list[int]:
        self.batch.run(max_chunks=1)
        done = [j for j in self.jobs if self.batch.state(j) == 2]
        self._done = {self.jobs[j]: j for j in done}
        return [self.jobs.pop(j) for j in done]

    def collect_1d(self, r: Request):
        lq = self.lq
        H, _, _, n, W, Ec = self.systems[r.system]
        mu = self.batch.result(self._done[r.rid])
        muj = lq.jackson_1d(self.ctx, mu, W, 1e-3)
        return mu, lq.dos_fftgrid(self.ctx, muj, n, W, Ec, r.nump or 2048)

    def _solve_2d(self, ctx, systems, r: Request):
        lq = self.lq
        H, Vx, Vy, n, W, Ec = systems[r.system]
        formula = lq.LQ_KUBO_GREENWOOD if r.kind == "greenwood" else lq.LQ_KUBO_BASTIN
        nump = r.nump or 2 * r.M
        fd = lq.kubo_fft(H, Vx, Vx, self.states(r), r.M, nump, formula)
        return fd, lq.kubo_post(ctx, formula, fd, n, W, Ec)

    def _worker(self, i, ctx, systems, q):
        while True:
            r = q.get()
            if r is None:
                return
            try:
                res = self._solve_2d(ctx, systems, r)
            except Exception as e:  # reported by poll_2d's caller through the result
                res = e
            with self._cv:
                self._busy[i] -= 1
                self._done_2d.append((r.rid, res))
                self._cv.notify_all()

    def submit_2d(self, r: Request):
        if not self.workers:
            res = self._solve_2d(self.ctx, self.systems, r)
            with self._cv:
                self._done_2d.append((r.rid, res))
            return
        with self._cv:
            i = min(range(len(self.workers)), key=lambda w: self._busy[w])
            self._busy[i] += 1
        self._queues[i].put(r)

    def poll_2d(self) -> list[tuple[int, object]]:
        with self._cv:
            out, self._done_2d = self._done_2d, []
        for rid, res in out:
            if isinstance(res, Exception):
                raise res
        return out

    def wait_2d(self, timeout: float):
        with self._cv:
            if not self._done_2d:
                self._cv.wait(timeout)

    def run_2d(self, r: Request):
        """A 2D request solved synchronously on the first 2D context."""
        if self.workers:
            return self._solve_2d(self.workers[0][0], self.workers[0][1], r)
        return self._solve_2d(self.ctx, self.systems, r)

    def alone(self, r: Request):
        """The same request served by itself (fresh batch for DOS)."""
        if r.kind != "dos":
            return self.run_2d(r)
        lq = self.lq
        b = lq.Batch(self.ctx, chunk=self.chunk)
        j = b.submit(self.systems[r.system][0], self.states(r), r.M)
        b.run()
        mu = b.result(j)
        b.close()
        H, _, _, n, W, Ec = self.systems[r.system]
        muj = lq.jackson_1d(self.ctx, mu, W, 1e-3)
        return mu, lq.dos_fftgrid(self.ctx, muj, n, W, Ec, r.nump or 2048)

    def close(self):
        for q in self._queues:
            q.put(None)
        for t in self._threads:
            t.join()
        self.batch.close()
