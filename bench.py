#!/usr/bin/env python
"""Headline benchmark: Chebyshev moments/s of the Kubo-Bastin 2D-moment path.

Default workload (one GPU): Haldane honeycomb 512 x 512 cells (N = 524288),
t = 1, t2 = 0.1 (phase pi/2), Anderson disorder U[-1, 1] (seeded), complex128
CSR H rescaled by 1/4.6, V_x = i[H, X], V_y = i[H, Y]; M = 512 x 512 moments,
R = 4 random-phase vectors; sigma_xx and sigma_xy by the FFT-grid Kubo-Bastin
formula on 2048 energies.  One step = one full solve of both components
(Kubo_solver_FFT::compute + post-processing for OPR in {V_x, V_y}).

  value = 2 components * M^2 * R / (seconds per step)

`--workload dos` runs the DOS configuration instead (square 256 x 256,
M = 1024, R = 8, Jackson + DOS on 2048 energies; value = M * R / s).

`--workload serve` replays a seeded Poisson stream of DOS, Greenwood and
Bastin requests (200 req/s; DOS on cubic Anderson / Kane-Mele systems with
N up to 2.1M and M up to 4096, 2D on Haldane honeycombs) through the serving
loop (continuous batching, admission by device footprint); one step is the
whole stream, value = moments served / s, with p50/p99 request latency.

`--workload batch` runs 64 independent DOS instances (cubic Anderson
L in {16..96}, Kane-Mele-like honeycomb N = 16k..1M, M in {256..2048},
R in {1..16}, drawn from a fixed seed) through the batched engine; one step
is the whole batch, value = sum of M * R over the jobs / s.  The same jobs
run one after another through the single-job solver once, for comparison.

`--workload ingest` measures operator ingestion: the Hamiltonian of a clean
512 x 512 Haldane honeycomb generated on the GPU from its unit-cell hopping
list (wannier2sparse's wrap + CSR + the Hermitian completion), value = stored
CSR entries made ready on the GPU per second; the same operator's text file
through the native reader and the reference's stream reader are timed beside it.

Prints one JSON line.  The roofline object describes the kernel class that
takes most of the step time, measured with HIP events on the library's own
stream; cpu_baseline times the CPU restatement of the reference algorithm on
a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6
INT8_MFMA_PEAK_TOPS = 5000.0
FP32_MFMA_PEAK_TFS = 157.3
BF16_MFMA_PEAK_TFS = 2500.0
PRECISIONS = {"fp64": 0, "fp32x2": 2, "bf16x3": 3, "int8x4": 4, "int8x5": 5, "int8x6": 6, "int8x7": 7, "int8x8": 8}


def kernel_source_sha16():
    """Hash of the HIP sources the kernels are built from (the key of the
    committed PMC traffic summary)."""
    import glob
    import hashlib

    h = hashlib.sha256()
    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "linqt_amd", "csrc")
    for path in sorted(glob.glob(os.path.join(root, "*.hip")) + glob.glob(os.path.join(root, "*.h"))
                       + glob.glob(os.path.join(root, "modules", "*.hip"))):
        h.update(os.path.basename(path).encode())
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def host_cpu():
    """Model name and logical CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return {"model": model, "logical_cpus_available": n}


def _pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def build_bastin(lq, ctx, L):
    from linqt_amd import lattices

    H, Vx, Vy = lattices.haldane_honeycomb(L, t=1.0, t2=0.1, phi=np.pi / 2, W=2.0, seed=20261203)
    Hg = lq.SparseMatrix.from_csr(ctx, H).Rescale(1 / 4.6, 0.0)
    Vxg = lq.SparseMatrix.from_csr(ctx, Vx)
    Vyg = lq.SparseMatrix.from_csr(ctx, Vy)
    return (H, Vx, Vy), (Hg, Vxg, Vyg)


def cpu_baseline_bastin(host_ops, M, R, nump, phi):
    """Reference algorithm (Kubo_solver_FFT, one run per sigma component):
    per state and component 2M H-steps, M+1 V-applies and N sites x 8 DFTs.
    Timed: a few H steps and V applies at full N, and the per-site DFT stage
    on a sample of sites; the step time is extrapolated from these."""
    import oracle as O

    H, Vx, Vy = host_ops
    n = H.n
    Ho = O.Op.from_csr(H).rescale(1 / 4.6, 0.0)
    Vo = O.Op.from_csr(Vx)
    x = np.ascontiguousarray(phi[:1])
    y = np.zeros_like(x)
    O.omp_cheb_step(Ho, x, y)  # builds the completed CSR outside the timing
    O.omp_mv(Vo, x, y)
    reps, samples = 5, 7

    def median_time(f, k):
        ts = []
        for _ in range(samples):
            t0 = time.perf_counter()
            for _ in range(k):
                f()
            ts.append((time.perf_counter() - t0) / k)
        return float(np.median(ts)), float(min(ts)), float(max(ts))

    tH, tHmin, tHmax = median_time(lambda: O.omp_cheb_step(Ho, x, y), reps)
    tV, _, _ = median_time(lambda: O.omp_mv(Vo, x, y), reps)
    S = 2048
    rng = np.random.default_rng(0)
    Ls = (rng.normal(size=(M, S)) + 1j * rng.normal(size=(M, S))) / np.sqrt(n)
    Rs = (rng.normal(size=(M, S)) + 1j * rng.normal(size=(M, S))) / np.sqrt(n)
    tsite, _, _ = median_time(lambda: O.omp_bastin_ffts(Ls, Rs, nump), 1)
    tsite /= S
    per_state_comp = 2 * M * tH + (M + 1) * tV + n * tsite
    step = 2 * R * per_state_comp
    return {
        "value": 2 * M * M * R / step,
        "unit": "moments/s",
        "cores": O.omp_threads(),
        "kind": "port",
        "host": host_cpu(),
        "sample": (f"CPU restatement of Kubo_solver_FFT (OpenMP, {O.omp_threads()} threads), medians of {samples} "
                   f"samples (H step {tHmin * 1e3:.2f}-{tHmax * 1e3:.2f} ms over the samples): timed {reps} H steps "
                   f"({tH * 1e3:.2f} ms), {reps} V applies ({tV * 1e3:.2f} ms) at N={n} and the per-site Bastin "
                   f"DFT stage on {S} sites ({tsite * 1e6:.1f} us/site, nump={nump}); step extrapolated as "
                   f"2R x (2M t_H + (M+1) t_V + N t_site) = {step:.1f} s"),
    }


def batch_instances(lq, ctx, njobs):
    """The 64-instance mix: even jobs cubic Anderson, odd jobs Kane-Mele-like."""
    from linqt_amd import lattices

    rng = np.random.default_rng(20261205)
    specs, mats = [], []
    for i in range(njobs):
        if i % 2 == 0:
            L = int(rng.choice([16, 24, 32, 48, 64, 96]))
            W = float(rng.uniform(1.0, 4.0))
            Hh = lattices.cubic_anderson(L, W=W, seed=20261205 + i)
            a = 6.0 + W / 2 + 0.01
            kind = "cubic"
        else:
            L = int(rng.choice([64, 91, 128, 181, 256, 362, 512]))
            W = float(rng.uniform(0.5, 2.0))
            Hh = lattices.kane_mele(L, W=W, seed=20261205 + i)
            a = 4.6
            kind = "kanemele"
        M = int(rng.choice([256, 512, 1024, 2048]))
        R = int(rng.choice([1, 2, 4, 8, 16]))
        H = lq.SparseMatrix.from_csr(ctx, Hh).Rescale(1 / a, 0.0)
        phi = lattices.random_phase_states(Hh.n, R, seed=20261205 + 1000 + i)
        specs.append({"kind": kind, "L": L, "W": W, "a": a, "M": M, "R": R})
        mats.append((H, Hh, lq.DeviceStates(ctx, phi)))
    return specs, mats


_SERVE_HOST = {}


def serve_systems(lq, ctx):
    """name -> (H~, V_x, V_y, N, W, Ec) for the serving mix, uploaded on ctx
    (the host matrices are generated once)."""
    from linqt_amd import lattices

    if not _SERVE_HOST:
        for L in (24, 48, 96):
            _SERVE_HOST[f"cubic{L}"] = (lattices.cubic_anderson(L, W=2.0, seed=20261207 + L), None, None, 7.01)
        for L in (64, 256, 724):
            _SERVE_HOST[f"kanemele{L}"] = (lattices.kane_mele(L, W=1.0, seed=20261207 + L), None, None, 4.6)
        for L in (96, 256, 1024):
            _SERVE_HOST[f"haldane{L}"] = (*lattices.haldane_honeycomb(L, seed=20261207 + L), 4.6)
    out = {}
    for name, (H, Vx, Vy, a) in _SERVE_HOST.items():
        out[name] = (lq.SparseMatrix.from_csr(ctx, H).Rescale(1 / a, 0.0),
                     None if Vx is None else lq.SparseMatrix.from_csr(ctx, Vx),
                     None if Vy is None else lq.SparseMatrix.from_csr(ctx, Vy), H.n, 2 * a, 0.0)
    return out


def cpu_baseline_batch(specs, mats):
    """Times the CPU restatement on one state of a mid-size instance for 64
    steps and extrapolates by the row-step work (nnz + N R) of the batch."""
    import oracle as O

    k = min(range(len(specs)), key=lambda i: abs(mats[i][1].n - 100000))
    Hh = mats[k][1]
    Ho = O.Op.from_csr(Hh).rescale(1 / specs[k]["a"], 0.0)
    phi = np.zeros((1, Hh.n), complex)
    phi[0, 0] = 1.0
    O.omp_spectral_moments(Ho, phi, 2)
    steps = 64
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.omp_spectral_moments(Ho, phi, steps)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    unit = t / (steps * (2 * Hh.nnz - Hh.n + Hh.n))
    work = sum(sp["M"] * (sp["R"] * (2 * m[1].nnz - m[1].n) + m[1].n * sp["R"]) for sp, m in zip(specs, mats))
    step = unit * work
    return {
        "value": sum(sp["M"] * sp["R"] for sp in specs) / step,
        "unit": "moments/s",
        "cores": O.omp_threads(),
        "kind": "port",
        "host": host_cpu(),
        "sample": (f"CPU restatement of SpectralMoments (OpenMP, {O.omp_threads()} threads): 1 state x {steps} "
                   f"moments of {specs[k]['kind']}{specs[k]['L']} (N={Hh.n}) took {t:.3f} s (median of 5); extrapolated by "
                   f"row-step work (nnz + N per state-step) to the whole batch = {step:.1f} s"),
    }


def cpu_baseline_dos(host_H, M, R, phi):
    import oracle as O

    Ho = O.Op.from_csr(host_H).rescale(1 / 4.6, 0.0)
    Msample = min(M, 256)
    O.omp_spectral_moments(Ho, phi[:1], 2)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.omp_spectral_moments(Ho, phi[:1], Msample)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    per_moment_state = t / Msample
    return {
        "value": 1.0 / per_moment_state,
        "unit": "moments/s",
        "cores": O.omp_threads(),
        "kind": "port",
        "host": host_cpu(),
        "sample": (f"CPU restatement of SpectralMoments (OpenMP, {O.omp_threads()} threads): 1 state x {Msample} "
                   f"moments at N={host_H.n} took {t:.3f} s (median of 5, {min(ts):.3f}-{max(ts):.3f} s); rate per "
                   f"moment extrapolated to M={M}, R={R}"),
    }


def haldane_hoppings(t=1.0, t2=0.1, mass=0.3):
    """Unit-cell hopping list (cell, edge, value) of a Haldane honeycomb with
    a staggered mass: 3 nearest-neighbour bonds A->B, 3 second-neighbour bonds
    per sublattice with phase pi/2 (i t2 / -i t2), each with its Hermitian
    partner, plus +-mass on site."""
    base = [((0, 0, 0), (0, 1), -t), ((-1, 0, 0), (0, 1), -t), ((0, -1, 0), (0, 1), -t),
            ((1, 0, 0), (0, 0), 1j * t2), ((0, 1, 0), (0, 0), -1j * t2), ((-1, 1, 0), (0, 0), 1j * t2),
            ((1, 0, 0), (1, 1), -1j * t2), ((0, 1, 0), (1, 1), 1j * t2), ((-1, 1, 0), (1, 1), -1j * t2)]
    cell, edge, val = [], [], []
    for c, e, v in base:
        cell += [c, tuple(-x for x in c)]
        edge += [e, e[::-1]]
        val += [v, np.conj(v)]
    for o, m in ((0, mass), (1, -mass)):
        cell.append((0, 0, 0))
        edge.append((o, o))
        val.append(m)
    return np.array(cell, np.int32), np.array(edge, np.int32), np.array(val, np.complex128)


def run_ingest(args, lq, ctx, log):
    """Operator ingestion: the Hamiltonian of a 512 x 512 Haldane honeycomb
    (N = 524288) made ready on the GPU.  One step = lq_csr_from_hoppings
    (supercell wrap, CSR and Hermitian completion generated on the device,
    sliced-ELL upload).  Also timed once: the same operator's reference text
    file (6 digits, as wannier2sparse writes it) through the native parallel
    reader, its binary form, and the full file -> device operator path; the
    CPU baseline is the reference's stream reader restated in C."""
    import tempfile

    import oracle

    L = args.L or 512
    cell, edge, val = haldane_hoppings()
    dims = (L, L, 1)
    A0 = lq.SparseMatrix.from_hoppings(ctx, 2, cell, edge, val, dims)
    m = A0.export()
    n = A0.n
    A0.close()
    tmpdir = tempfile.mkdtemp(prefix="lq_ingest_")
    ptxt = os.path.join(tmpdir, "haldane.HAM.CSR")
    pbin = ptxt + ".bin"
    lq.write_csr_file(ptxt, m, lq.LQ_FMT_TEXT_REF)
    lq.write_csr_file(pbin, m, lq.LQ_FMT_BINARY)
    txt_bytes = os.path.getsize(ptxt)
    log(f"ingest: N={n} stored entries={m.nnz} text file {txt_bytes / 1e6:.1f} MB")

    def step():
        A = lq.SparseMatrix.from_hoppings(ctx, 2, cell, edge, val, dims)
        A.close()

    for w in range(args.warmup):
        step()
    ctx.synchronize()
    lat = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        step()
        ctx.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
        log(f"step {k + 1}/{args.steps}: {lat[-1]:.1f} ms")
    elapsed = time.perf_counter() - t_start
    # one instrumented generation: HIP events around the count and fill kernels
    ctx.set_timing(True)
    ctx.reset_timing()
    A = lq.SparseMatrix.from_hoppings(ctx, 2, cell, edge, val, dims)
    ctx.synchronize()
    tm = ctx.timing()
    ctx.set_timing(False)
    completed = A.info()["nnz"]
    A.close()
    kern_bytes = 21.0 * completed + 8.0 * n  # count: 4 B per row; fill: col + value + kind per entry, rowptr
    kern_gbs = kern_bytes / (tm["post_ms"] * 1e-3) / 1e9

    def timed(fn, reps=3):
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    t_txt = timed(lambda: lq.read_csr_file(ptxt))
    t_bin = timed(lambda: lq.read_csr_file(pbin))

    def load():
        B = lq.SparseMatrix.from_file(ctx, ptxt)
        ctx.synchronize()
        B.close()

    t_load = timed(load)
    entries = m.nnz
    ms = elapsed * 1e3 / args.steps
    out = {"metric": "CSR entries made ready on the GPU per second", "value": entries / (ms * 1e-3),
           "unit": "entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "complex128",
           "data": "synthetic",
           "latency_ms": {"p50": _pct(lat, 50), "p99": _pct(lat, 99)},
           "config": {"workload": f"ingest: Haldane honeycomb {L}x{L} cells N={n} (t=1, t2=0.1, mass 0.3), "
                                  f"{entries} stored entries, {completed} completed",
                      "N": n, "stored_entries": entries, "completed_entries": completed, "M": 0},
           "roofline": {"bound": "hbm", "kernel": "k_sc_count + k_sc_fill", "achieved": kern_gbs,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kern_gbs / HBM_PEAK_GBS, "traffic": None,
                        "per_launch": f"{kern_bytes:.4g} B (21 B per completed entry + 8 B per row)",
                        "avg_launch_ms": tm["post_ms"]},
           "phases": {"generate_ms": ms, "kernels_ms": tm["post_ms"],
                      "text_parse_ms": t_txt * 1e3, "text_MB_per_s": txt_bytes / t_txt / 1e6,
                      "text_entries_per_s": entries / t_txt, "binary_read_ms": t_bin * 1e3,
                      "text_file_to_device_operator_ms": t_load * 1e3, "text_file_MB": txt_bytes / 1e6}}
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        oracle.csr_read_stream(ptxt)
        t_ref = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": entries / t_ref, "unit": "entries/s", "cores": 1, "kind": "port",
                               "sample": f"the reference's stream CSR reader (Sparse::OPERATOR_FromCSRFile, "
                                         f"token-by-token stdio) on the same {txt_bytes / 1e6:.1f} MB text file: "
                                         f"{t_ref:.2f} s"}
    else:
        out["cpu_baseline"] = None
    for p in (ptxt, pbin):
        os.remove(p)
    os.rmdir(tmpdir)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("bastin", "dos", "batch", "serve", "ingest"), default="bastin")
    ap.add_argument("--requests", type=int, default=200, help="requests of the serve workload")
    ap.add_argument("--rate", type=float, default=200.0, help="mean arrival rate of the serve workload (req/s)")
    ap.add_argument("--jobs", type=int, default=64, help="instances of the batch workload")
    ap.add_argument("--L", type=int, default=None, help="lattice size (cells per side)")
    ap.add_argument("--M", type=int, default=None)
    ap.add_argument("--R", type=int, default=None)
    ap.add_argument("--nump", type=int, default=2048)
    ap.add_argument("--algo", choices=("direct", "doubling"), default="direct",
                    help="1D moment algorithm of the dos workload (doubling: OP = identity, half the recursion steps)")
    ap.add_argument("--precision", choices=tuple(PRECISIONS), default="fp64",
                    help="contraction precision of the 2D table (int8xS: S int8 digits on INT8 MFMA)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="only time a device-to-device copy of 1 GiB (the reference kernel for PMC byte counts)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(over several GPUs the path would split the random states between them and sum "
              "the moment tables with one all-reduce; only the one-GPU share is built)", file=sys.stderr)
        sys.exit(2)

    import linqt_amd as lq
    from linqt_amd import lattices

    ctx = lq.Context(0)
    ctx.set_timing(False)  # the timed steps run uninstrumented; one extra step is instrumented below
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)

    if args.calibrate:
        from linqt_amd import _lib

        nbytes = 1 << 30
        src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
        _lib.call("lq_memcpy_d2d", ctx.h, dst, src, nbytes)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.call("lq_memcpy_d2d", ctx.h, dst, src, nbytes)
        dt = (time.perf_counter() - t0) / reps
        out = {"calibrate": "d2d copy", "bytes_read": nbytes, "bytes_written": nbytes, "reps": reps,
               "ms": dt * 1e3, "GB/s (read+write)": 2 * nbytes / dt / 1e9}
        ctx.free(src)
        ctx.free(dst)
        # the library's hand-written streaming kernels (16-byte non-temporal
        # accesses, grid = 8 blocks per CU): 4 GiB read / copied / written
        probe = {}
        for kind, name in ((0, "read"), (1, "copy"), (2, "write")):
            ms, moved = C.c_double(), C.c_double()
            _lib.call("lq_hbm_probe", ctx.h, 4 << 30, kind, 10, C.byref(ms), C.byref(moved))
            probe[name] = {"ms": ms.value, "GB/s": moved.value / ms.value / 1e6}
        out["streaming_kernels"] = probe
        out["ceiling_GB/s"] = max(v["GB/s"] for v in probe.values())
        print(json.dumps(out), flush=True)
        return

    if args.workload == "ingest":
        run_ingest(args, lq, ctx, log)
        return

    if args.workload == "bastin":
        L = args.L or 512
        M = args.M or 512
        R = args.R or 4
        nump = args.nump
        host_ops, (Hg, Vxg, Vyg) = build_bastin(lq, ctx, L)
        n = host_ops[0].n
        phi = lattices.random_phase_states(n, R, seed=20261203)
        dphi = lq.DeviceStates(ctx, phi)
        log(f"bastin N={n} M={M} R={R} nump={nump}")

        prec = PRECISIONS[args.precision]

        def step(precision=prec):
            fd = lq.kubo_fft(Hg, Vxg, [Vxg, Vyg], dphi, M, nump, lq.LQ_KUBO_BASTIN, precision=precision)
            return [lq.kubo_post(ctx, lq.LQ_KUBO_BASTIN, fd[o], n, 9.2, 0.0) for o in range(2)]

        moments_per_step = 2 * M * M * R
        config = {"workload": f"kubo_bastin_2d honeycomb-haldane {L}x{L} cells N={n} t2=0.1 W=2 "
                              f"M={M}x{M} R={R} sigma_xx+sigma_xy nump={nump}",
                  "N": n, "M": M, "R": R, "nump": nump, "components": 2, "precision": args.precision}
    elif args.workload == "serve":
        from linqt_amd import serve

        systems = serve_systems(lq, ctx)
        # two 2D workers, each its own context (stream) with the systems uploaded on it
        wctx = [lq.Context(0) for _ in range(2)]
        workers = [(c, serve_systems(lq, c)) for c in wctx]
        names_1d = [k for k, v in systems.items() if v[1] is None]
        names_2d = [k for k, v in systems.items() if v[1] is not None]
        reqs_proto = serve.request_mix(args.requests, args.rate, 20261207, names_1d, names_2d, frac_2d=0.3,
                                       M_1d=(256, 512, 1024, 2048, 4096), M_2d=(256, 512, 1024, 2048, 4096),
                                       R=(1, 2, 4, 8), sizes={k: v[3] for k, v in systems.items()},
                                       max_2d_work=float(2 ** 40))
        log(f"serve: {len(reqs_proto)} requests at {args.rate} req/s over {len(systems)} systems")
        served = []

        def step():
            reqs = [serve.Request(r.rid, r.kind, r.system, r.M, r.R, r.arrival, r.seed) for r in reqs_proto]
            be = serve.GpuBackend(lq, ctx, systems, chunk=32, workers_2d=workers)
            serve.Scheduler(be, hbm_budget=int(200e9)).run(reqs)
            be.close()
            served.append(serve.summary(reqs))
            return reqs

        moments_per_step = sum(r.moments for r in reqs_proto)
        n = max(v[3] for v in systems.values())
        M = max(r.M for r in reqs_proto)
        R = max(r.R for r in reqs_proto)
        config = {"workload": f"serve: {args.requests} Poisson requests at {args.rate:g} req/s, DOS (cubic Anderson "
                              f"L24-96, Kane-Mele N16k-2.1M, M256-4096) + Greenwood/Bastin (Haldane N18k-2.1M, "
                              f"M256-4096 with M^2 N <= 2^40), two 2D workers beside the DOS batch, seed 20261207",
                  "requests": args.requests, "rate": args.rate, "max_N": n, "M": M, "R": R, "precision": "fp64",
                  "systems": {k: v[3] for k, v in systems.items()}}
    elif args.workload == "batch":
        if args.precision != "fp64":
            sys.exit("bench.py: --precision applies to the 2D (bastin) workload")
        specs, mats = batch_instances(lq, ctx, args.jobs)
        log(f"batch: {len(specs)} jobs, sum N = {sum(m[1].n for m in mats)}")

        def step():
            b = lq.Batch(ctx, chunk=32)
            ids = [b.submit(H, ds, sp["M"]) for sp, (H, _, ds) in zip(specs, mats)]
            b.run()
            out = [b.result(j) for j in ids]
            b.close()
            return out

        moments_per_step = sum(sp["M"] * sp["R"] for sp in specs)
        n = sum(m[1].n for m in mats)
        M = max(sp["M"] for sp in specs)
        R = max(sp["R"] for sp in specs)
        config = {"workload": f"kpm_dos batch of {len(specs)}: cubic Anderson L16-96 + Kane-Mele N16k-1M, "
                              f"M 256-2048, R 1-16 (seed 20261205)",
                  "jobs": len(specs), "sum_N": n, "M": M, "R": R, "precision": "fp64",
                  "instances": [f"{sp['kind']}{sp['L']} M{sp['M']} R{sp['R']}" for sp in specs]}
    else:
        if args.precision != "fp64":
            sys.exit("bench.py: --precision applies to the 2D (bastin) workload")
        L = args.L or 256
        M = args.M or 1024
        R = args.R or 8
        nump = args.nump
        Hh = lattices.square_anderson(L, W=1.0, seed=20261202)
        Hg = lq.SparseMatrix.from_csr(ctx, Hh).Rescale(1 / 4.6, 0.0)
        n = Hh.n
        phi = lattices.random_phase_states(n, R, seed=20261202)
        dphi = lq.DeviceStates(ctx, phi)
        log(f"dos N={n} M={M} R={R}")

        algo = lq.LQ_ALGO_DOUBLING if args.algo == "doubling" else lq.LQ_ALGO_DIRECT

        def step():
            mu = lq.spectral_moments(Hg, None, dphi, M, algo=algo)
            muj = lq.jackson_1d(ctx, mu, 9.2, 1e-3)
            return lq.dos_fftgrid(ctx, muj, n, 9.2, 0.0, nump)

        moments_per_step = M * R
        config = {"workload": f"kpm_dos square {L}x{L} N={n} W=1 M={M} R={R} jackson dos_energies={nump}"
                              + (" doubling" if args.algo == "doubling" else ""),
                  "N": n, "M": M, "R": R, "nump": nump, "precision": "fp64", "algo": args.algo}

    for w in range(args.warmup):
        step()
        log(f"warmup {w + 1}/{args.warmup} done")
    ctx.synchronize()
    ctx.reset_timing()
    result = None
    lat = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        result = step()
        ctx.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
        log(f"step {k + 1}/{args.steps}: {lat[-1]:.1f} ms")
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    # one more step with HIP events around every launch (on the library's
    # stream) for the per-kernel figures; it is not part of `value`
    ctx.set_timing(True)
    ctx.reset_timing()
    step()
    ctx.synchronize()
    tm = ctx.timing()
    ctx.set_timing(False)
    tsteps = 1

    ms_per_step = elapsed * 1e3 / args.steps
    value = moments_per_step / (elapsed / args.steps)
    spmv_gbs = tm["spmv_bytes"] / (tm["spmv_ms"] * 1e-3) / 1e9 if tm["spmv_ms"] > 0 else 0.0
    # complex-equivalent rate (8 flops per complex multiply-add) and the rate of
    # the real MFMA flops the FP64 kernel executes: the three-multiplication
    # kernel (tables of at least 128 x 128) does 6 per complex multiply-add
    ctr_cplx_tfs = tm["contract_flops"] / (tm["contract_ms"] * 1e-3) / 1e12 if tm["contract_ms"] > 0 else 0.0
    ctr_mode = os.environ.get("LQ_CONTRACT", "")
    ctr_128x64 = ctr_mode in ("dl", "w1") or "LQ_CONTRACT_DL" in os.environ or "LQ_CONTRACT_DLBUF" in os.environ
    if ctr_mode == "4m":
        ctr_kernel, ctr_fpc = "k_contract_f64_lds", 8
    elif ctr_mode == "wave" or config.get("M", 0) < 128:
        ctr_kernel, ctr_fpc = "k_contract_f64", 8
    elif ctr_mode == "w1":
        ctr_kernel, ctr_fpc = "k_contract_f64_3m_w1", 6
    elif ctr_128x64 and os.environ.get("LQ_CONTRACT_DL", "") == "0":
        ctr_kernel, ctr_fpc = "k_contract_f64_3m", 6
    elif ctr_128x64:
        ctr_kernel, ctr_fpc = "k_contract_f64_3m_dl", 6
    else:
        ctr_kernel, ctr_fpc = "lq_contract_f64_3m_t128", 6
    ctr_tfs = ctr_cplx_tfs * ctr_fpc / 8
    kernels = {
        "spmv": {"ms_per_step": tm["spmv_ms"] / tsteps, "launches_per_step": tm["spmv_launches"] / tsteps,
                 "GB/s": spmv_gbs, "frac_hbm": spmv_gbs / HBM_PEAK_GBS},
        "contract": {"ms_per_step": tm["contract_ms"] / tsteps,
                     "launches_per_step": tm["contract_launches"] / tsteps,
                     "kernel": ctr_kernel, "TFLOP/s": ctr_tfs, "frac_fp64_mfma": ctr_tfs / FP64_MFMA_PEAK_TFS,
                     "complex_equiv_TFLOP/s": ctr_cplx_tfs},
        "post_ms_per_step": tm["post_ms"] / tsteps,
        "instrumented_step": "one extra step after the timed ones, HIP events around every launch",
    }
    # HBM bytes per launch of the dominant kernel, from the committed PMC summary of this workload,
    # used only when it was measured on the kernel sources of this tree
    traffic_tab = {}
    traffic_note = None
    try:
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")) as f:
            tab = json.load(f)
        if tab.get("source_sha16") != kernel_source_sha16():
            traffic_note = (f"PMC summary measured on kernel sources {tab.get('source_sha16')}, "
                            f"this tree is {kernel_source_sha16()}: traffic omitted")
        else:
            traffic_tab = tab["workloads"].get(config["workload"], {}).get("bytes_per_launch", {})
    except (OSError, ValueError, KeyError):
        traffic_tab = {}
    spmm_names = sorted((k for k in traffic_tab if k.startswith("k_spmm<")), key=lambda k: -traffic_tab[k])
    if spmm_names and tm["spmv_launches"] > 0:
        # physical bytes of the recursion step (PMC) beside the reference-CSR algorithmic bytes
        phys = traffic_tab[spmm_names[0]]
        avg_ms = tm["spmv_ms"] / tm["spmv_launches"]
        kernels["spmv"].update({"physical_kernel": spmm_names[0], "physical_bytes_per_launch": phys,
                                "physical_GB/s": phys / (avg_ms * 1e-3) / 1e9,
                                "frac_hbm_physical": phys / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS})
    kernels["spmv"]["frac_hbm_basis"] = ("frac_hbm: reference CSR bytes 20 nnz + 4(N+1) + 48 k N per launch "
                                         "(algorithmic); frac_hbm_physical: PMC FETCH_SIZE + WRITE_SIZE bytes")

    def traffic_of(name):
        # the summary names template kernels with their arguments (k_x<2>)
        if name in traffic_tab:
            return traffic_tab[name]
        hits = [v for k, v in traffic_tab.items() if k.startswith(name + "<")]
        return hits[0] if len(hits) == 1 else None
    sequential = None
    if args.workload == "batch":
        # the same jobs, one after another through the single-job solver (untimed in value)
        t0 = time.perf_counter()
        seq = [lq.spectral_moments(H, None, ds, sp["M"]) for sp, (H, _, ds) in zip(specs, mats)]
        ctx.synchronize()
        t_seq = time.perf_counter() - t0
        same = all(np.array_equal(a, b) for a, b in zip(seq, result))
        sequential = {"ms": t_seq * 1e3, "speedup_of_batch": t_seq / (elapsed / args.steps),
                      "bitwise_equal": bool(same)}
        log(f"sequential single-job solves: {t_seq * 1e3:.1f} ms, batch identical: {same}")
    accuracy = None
    split = args.workload == "bastin" and args.precision != "fp64"
    if split:
        # sigma of the split-precision path against the FP64 path (untimed)
        ref = step(0)
        accuracy = {}
        for o, comp in enumerate(("xx", "xy")):
            s_ref, s_got = np.asarray(ref[o][1]), np.asarray(result[o][1])
            accuracy[f"sigma_{comp}_rel_err"] = float(np.abs(s_got - s_ref).max() / np.abs(s_ref).max())
        log(f"accuracy vs fp64: {accuracy}")
    if split and PRECISIONS[args.precision] in (2, 3) and tm["contract_ms"] > tm["spmv_ms"]:
        S = PRECISIONS[args.precision]
        pieces = 3 if S == 2 else 6  # piece products per real product
        launches = max(tm["contract_launches"], 1)
        flops = 2.0 * M * M * n * R * 3 * pieces
        avg_ms = tm["contract_ms"] / launches
        tfs = flops / (avg_ms * 1e-3) / 1e12
        peak = FP32_MFMA_PEAK_TFS if S == 2 else BF16_MFMA_PEAK_TFS
        kname = f"k_contract_split<{S}>"
        roofline = {"bound": "mfma", "kernel": kname, "achieved": tfs, "peak": peak, "unit": "TFLOP/s",
                    "frac": tfs / peak, "traffic": traffic_of(kname),
                    "per_launch": f"{flops:.4g} flop (2 M^2 N R x 3 real products x {pieces} piece products)",
                    "avg_launch_ms": avg_ms}
    elif split and tm["contract_ms"] > tm["spmv_ms"]:
        S = PRECISIONS[args.precision]
        launches = max(tm["contract_launches"], 1)
        kp = -(-n * R // 4096) * 4096
        ops = 2.0 * M * M * kp * 4 * (S * (S + 1) // 2)  # INT8 MFMA ops per launch
        avg_ms = tm["contract_ms"] / launches
        tops = ops / (avg_ms * 1e-3) / 1e12
        kname = f"k_contract_i8<{S}, {2 if S <= 5 else 1}>"
        roofline = {"bound": "mfma", "kernel": kname + " (+ k_slice)", "achieved": tops,
                    "peak": INT8_MFMA_PEAK_TOPS, "unit": "TOP/s", "frac": tops / INT8_MFMA_PEAK_TOPS,
                    "traffic": traffic_of(kname), "traffic_note": traffic_note,
                    "per_launch": f"{ops:.4g} int8 op (8 M^2 ceil(N R / 4096) 4096 S(S+1)/2)",
                    "avg_launch_ms": avg_ms}
    elif tm["contract_ms"] > tm["spmv_ms"]:
        per_launch = tm["contract_flops"] / max(tm["contract_launches"], 1) * ctr_fpc / 8
        kname = ctr_kernel
        how = ("6*M^2*N*R: three real FP64 MFMA products per complex multiply-add" if ctr_fpc == 6
               else "8*M^2*N*R: four real FP64 MFMA products per complex multiply-add")
        roofline = {"bound": "mfma", "kernel": kname, "achieved": ctr_tfs, "peak": FP64_MFMA_PEAK_TFS,
                    "unit": "TFLOP/s", "frac": ctr_tfs / FP64_MFMA_PEAK_TFS, "traffic": traffic_of(kname),
                    "traffic_note": traffic_note,
                    "per_launch": f"{per_launch:.4g} flop ({how})",
                    "flop_executed_per_launch": per_launch,
                    "flop_complex_equiv_per_launch": tm["contract_flops"] / max(tm["contract_launches"], 1),
                    "complex_equiv_TFLOP/s": ctr_cplx_tfs,
                    "avg_launch_ms": tm["contract_ms"] / max(tm["contract_launches"], 1)}
    else:
        per_launch = tm["spmv_bytes"] / max(tm["spmv_launches"], 1)
        roofline = {"bound": "hbm", "kernel": "k_spmm", "achieved": spmv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": spmv_gbs / HBM_PEAK_GBS,
                    "traffic": traffic_tab[spmm_names[0]] if spmm_names else None,
                    "traffic_kernel": spmm_names[0] if spmm_names else None,
                    "traffic_note": traffic_note,
                    "per_launch": f"{per_launch:.4g} B (reference CSR bytes: 20 nnz + 4(N+1) + 48 k N)",
                    "avg_launch_ms": tm["spmv_ms"] / max(tm["spmv_launches"], 1)}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (bounded sample)")
        try:
            if args.workload == "bastin":
                cpu = cpu_baseline_bastin(host_ops, M, R, nump, phi)
            elif args.workload == "batch":
                cpu = cpu_baseline_batch(specs, mats)
            elif args.workload == "serve":
                cpu = None
            else:
                cpu = cpu_baseline_dos(Hh, M, R, phi)
        except Exception as e:  # the baseline is a report, not the measurement
            cpu = {"error": repr(e)}

    out = {
        "metric": "Chebyshev moments/sec",
        "value": value,
        "unit": "moments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128" if args.precision == "fp64" else f"complex128 in, {args.precision} MFMA contraction",
        "data": "synthetic",
        "config": config,
        "latency_ms": {"p50": _pct(lat, 50), "p99": _pct(lat, 99)},
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    if accuracy is not None:
        out["accuracy_vs_fp64"] = accuracy
    if sequential is not None:
        out["sequential_single_job"] = sequential
    if args.workload == "serve":
        out["serving"] = served[-1]
        out["latency_ms"] = served[-1]["latency_ms"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
