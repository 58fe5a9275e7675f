"""Benchmark of the batched Kronecker matvec on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Default workload (the headline): kronmult6_xbatched, FP64, d=6, n=4,
65,536 entries, distinct A factors per entry (uniform [-1,1]), X pointers
drawn uniformly from a pool of 8,192 vectors of 4,096, Y pointers drawn
uniformly from a pool of 2,048 vectors (about 32 atomic writers per Y), fixed
seed.  A step is one call over the whole batch, inputs resident in HBM.

Printed: one JSON line with entries/s, the roofline of the dominant kernel
and the reference CPU path timed on this host's cores.

Roofline: FP64 throughput.  Each entry costs 2*d*n^(d+1) = 196,608 flops
(n FMAs per element and mode).  The pools make the compulsory HBM footprint
about 0.46 GB per call (X pool 268 MB, Y pool read and written 134 MB, the
factors 50 MB, pointers 4 MB) against 12.9 GFLOP: 28 flop/B, above the FP64
ridge of 78.6 TFLOP/s / 8 TB/s = 9.8 flop/B.  roofline.achieved is the
launch's flops over the kernel's average duration (HIP events on its stream,
the update run through an xbatched plan so that one launch is one kernel).
The per-entry algorithmic bytes 8*(n^d [X] + 2*n^d [Y RMW] + d*n^2 [A]) +
(d+2)*8 [pointers] = 99,136 B ignore the pools' reuse; they are reported
under "hbm" with the compulsory bytes and the rocprofv3 counter bytes.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def _splitmix_uniform(count, seed):
    from helpers import uniform
    return uniform(count, seed)


def _splitmix_index(count, seed, mod):
    from helpers import splitmix64
    return (splitmix64(count, seed) % np.uint64(mod)).astype(np.int64)


def make_xbatched_inputs(d, n, batch, nx, ny, seed):
    """Host arrays of the headline workload (pools + maps)."""
    E = n ** d
    A = _splitmix_uniform(batch * d * n * n, seed * 7 + 1).reshape(batch * d, n * n)
    X = _splitmix_uniform(nx * E, seed * 7 + 2).reshape(nx, E)
    x_map = _splitmix_index(batch, seed * 7 + 3, nx)
    y_map = _splitmix_index(batch, seed * 7 + 4, ny)
    return A, X, x_map, y_map


def bytes_per_entry(d, n, ptr_layout=True, elem=8):
    b = elem * (n ** d + 2 * n ** d + d * n * n)
    return b + ((d + 2) * 8 if ptr_layout else 0)


def _cpu_rate(L, d, n, A, X, x_map, y_map, ny, sample, subbatch, min_seconds):
    """Median entries/s of the CPU path over repeated runs of `sample` entries."""
    import oracle
    E = n ** d
    Y = np.zeros((ny, E))
    W = np.zeros((2 * subbatch, E))
    a_map = np.arange(sample * d).reshape(sample, d)
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < min_seconds and len(times) < 50):
        t0 = time.perf_counter()
        oracle.xbatched(L, d, n, A, a_map, n, X, x_map[:sample], Y, y_map[:sample], W,
                        np.arange(2 * subbatch), sample, subbatch)
        times.append(time.perf_counter() - t0)
    return sample / float(np.median(times)), times


def cpu_baseline_xbatched(d, n, A, X, x_map, y_map, ny, sample):
    """The reference CPU path (oracle/_ref, compiled from the reference's own
    sources) or, where absent, the oracle's C restatement, on the first
    `sample` entries of the same workload (subbatchCount 1,024, as the
    reference's xbatched loop runs it), at every host core the process may
    use for about 10 s, then at one thread on a smaller sample for about 8 s;
    median rates.  Threads are left unbound: OMP_PROC_BIND=close with
    OMP_PLACES=cores put all eight threads of an 8-core container on one
    core (6.0e3 against 3.4e4 entries/s unbound)."""
    import oracle
    L = oracle.ref_lib()
    kind = "reference"
    if L is None:
        L = oracle.oracle_lib()
        kind = "port"
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    oracle.set_threads(L, threads)
    rate, times = _cpu_rate(L, d, n, A, X, x_map, y_map, ny, sample, 1024, 10.0)
    sample1 = max(1, sample // 16)
    oracle.set_threads(L, 1)
    rate1, times1 = _cpu_rate(L, d, n, A, X, x_map, y_map, ny, sample1, min(1024, sample1), 8.0)
    oracle.set_threads(L, threads)
    return {
        "value": rate,
        "unit": "entries/s",
        "cores": threads,
        "kind": kind,
        "sample": "the workload's first %d entries (same pools and maps), subbatchCount=1024, "
                  "%d OpenMP threads (OMP_PROC_BIND=%s, OMP_PLACES=%s), %d runs in %.1f s; "
                  "median %.3f s, best %.3f s per run"
                  % (sample, threads, os.environ.get("OMP_PROC_BIND"),
                     os.environ.get("OMP_PLACES"), len(times), sum(times),
                     float(np.median(times)), min(times)),
        "value_1thread": rate1,
        "sample_1thread": "the first %d entries at 1 thread, %d runs in %.1f s; median %.3f s"
                          % (sample1, len(times1), sum(times1), float(np.median(times1))),
    }


# The sources the headline kernel (kron_plane_kernel, pointer layout, FP64)
# is compiled from: its counter profile is used only for the build it
# measured.  (The Y-target grouping kernels are timed in the step but not in
# that profile.)
HEADLINE_SOURCES = ("kron_fast_f64_ptr.hip", "kron_fast.hpp", "kron_plane.hpp", "kron_dispatch.hpp",
                    "kron_device.hpp", "kron_internal.hpp")


def sources_sha16():
    """Hash of the headline kernel's sources."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "kronmult_amd", "csrc")
    for name in HEADLINE_SOURCES:
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def read_traffic(workload):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of
    this workload, or None when that summary measured other sources."""
    p = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        rec = json.load(f)
    if rec.get("sources_sha16") != sources_sha16():
        return None
    return rec.get("bytes_per_launch")


def _time_steps(step, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def preheat(step, seconds=0.25):
    """Run `step` back to back for about `seconds` before any warmup: the
    GPU's clocks ramp up over the first tens of milliseconds under load, and
    both timed regions should see the steady state."""
    import torch
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(20):
            step()
        torch.cuda.synchronize()


def bench_xbatched(args):
    import torch

    import kronmult_amd as K

    d, n, batch, nx, ny = 6, 4, 65536, 8192, 2048
    E = n ** d
    dev = "cuda:0"
    A, X, x_map, y_map = make_xbatched_inputs(d, n, batch, nx, ny, seed=2024)
    Ad = torch.from_numpy(A).to(dev)
    Xd = torch.from_numpy(X).to(dev)
    Yd = torch.zeros((ny, E), dtype=torch.float64, device=dev)
    nslot = batch  # pW: 2*subbatch slots (only touched by shapes larger than LDS)
    Wd = torch.empty((1, E), dtype=torch.float64, device=dev)

    def rows(t, idx):
        base = t.data_ptr()
        rb = t.stride(0) * t.element_size()
        return torch.from_numpy(base + np.asarray(idx, dtype=np.int64) * rb).to(dev)

    pA = rows(Ad, np.arange(batch * d))
    pX = rows(Xd, x_map)
    pY = rows(Yd, y_map)
    pW = rows(Wd, np.zeros(nslot, dtype=np.int64))
    stream = torch.cuda.current_stream()

    def step():
        K.kronmult_xbatched(d, n, pA, n, pX, pY, pW, batch, 0, dtype=torch.float64,
                            stream=stream)

    preheat(step, args.preheat)
    t0 = time.perf_counter()
    ms = _time_steps(step, args.steps, args.warmup, stream)
    wall = time.perf_counter() - t0
    value = batch / (ms * 1e-3)
    # The dominant kernel alone: the same update through an xbatched plan
    # (Y-target order computed once), so one launch per execute, timed with
    # HIP events on the stream it is launched on.
    plan = K.XbatchedPlan(d, n, pA, n, pX, pY, pW, batch, dtype=torch.float64, stream=stream)
    ms_kernel = _time_steps(lambda: plan.execute(stream), args.steps, args.warmup, stream)
    plan.close()
    # FP64 work: 2*d*n^(d+1) flops per entry (each mode: n FMAs per element),
    # the reference's own count (test_kronmult6_batched.cpp:324-335)
    flops = 2 * d * n ** (d + 1) * batch
    tflops = flops / (ms_kernel * 1e-3) / 1e12
    alg_bytes = bytes_per_entry(d, n) * batch
    # compulsory footprint: every X and Y vector the maps name (Y read and
    # written once), the factors and the pointer arrays
    compulsory = 8 * E * (len(np.unique(x_map)) + 2 * len(np.unique(y_map))) + \
        8 * d * n * n * batch + 8 * (d + 2) * batch
    traffic = read_traffic("xbatched_d6n4")
    sec = ms_kernel * 1e-3
    out = {
        "metric": "batch entries/sec (kronmult6_xbatched d=6 n=4 FP64, batchCount 65536)",
        "value": value,
        "unit": "entries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (splitmix64 uniform [-1,1), seed 2024)",
        "config": {"workload": "kronmult6_xbatched d=6 n=4 batchCount=65536 "
                               "Xpool=8192 Ypool=2048 lda=4",
                   "global_batch": batch},
        # Pooled X (8,192 vectors) and Y (2,048) make the workload's
        # compulsory HBM footprint 0.46 GB against 12.9 GFLOP: 28 flop/B,
        # above MI355X's FP64 ridge (78.6 TFLOP/s / 8 TB/s = 9.8), so the
        # binding roof is FP64 throughput (the kernel's FMAs run on the
        # vector units; 78.6 TFLOP/s is the FP64 peak of both the vector and
        # the matrix units)
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / FP64_PEAK_TFLOPS, "traffic": traffic},
        "kernel": "kron_plane_kernel<double,double,Ptr,true,false,true> (Y-target order)",
        "kernel_ms": ms_kernel,
        "flops_per_launch": flops,
        "hbm": {
            "peak_gbs": HBM_PEAK_GBS,
            "alg_bytes": alg_bytes, "alg_gbs": alg_bytes / sec / 1e9,
            "compulsory_bytes": compulsory, "compulsory_gbs": compulsory / sec / 1e9,
            "compulsory_frac": compulsory / sec / 1e9 / HBM_PEAK_GBS,
            "counter_bytes": traffic,
            "counter_gbs": (traffic / sec / 1e9) if traffic else None,
            "counter_frac": (traffic / sec / 1e9 / HBM_PEAK_GBS) if traffic else None,
        },
        "wall_s_timed": wall,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_xbatched(d, n, A, X, x_map, y_map, ny,
                                                    args.cpu_sample)
    print(json.dumps(out), flush=True)


def bench_batched(args):
    """kronmult6_batched at the same d=6, n=4 FP64 and 65,536 entries
    (contiguous layout, Y private per entry)."""
    import torch

    import kronmult_amd as K

    d, n, batch = 6, 4, 65536
    E = n ** d
    dev = "cuda:0"
    A = torch.from_numpy(_splitmix_uniform(batch * d * n * n, 11)).to(dev)
    X = torch.from_numpy(_splitmix_uniform(batch * E, 12)).to(dev)
    Y = torch.zeros(batch * E, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    ms = _time_steps(lambda: K.kronmult_batched(d, n, A, X, Y, None, batch, stream=stream),
                     args.steps, args.warmup, stream)
    alg = bytes_per_entry(d, n, ptr_layout=False) * batch
    print(json.dumps({"workload": "kronmult6_batched d=6 n=4 batchCount=65536", "dtype": "f64",
                      "ms_per_step": ms, "entries_per_s": batch / (ms * 1e-3),
                      "alg_gbs": alg / (ms * 1e-3) / 1e9,
                      "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}), flush=True)


def bench_adjoint(args):
    """Forward and hand-written adjoint, FP64, d=6, n=6, 4,096 entries
    (batched layout; A, X, dY uniform [-1,1])."""
    import torch

    import kronmult_amd as K

    d, n, batch = 6, 6, 4096
    E = n ** d
    dev = "cuda:0"
    A = torch.from_numpy(_splitmix_uniform(batch * d * n * n, 31)).to(dev)
    X = torch.from_numpy(_splitmix_uniform(batch * E, 32)).to(dev)
    dY = torch.from_numpy(_splitmix_uniform(batch * E, 33)).to(dev)
    Y = torch.zeros(batch * E, dtype=torch.float64, device=dev)
    W = torch.empty(2 * batch * E, dtype=torch.float64, device=dev)
    dX = torch.empty_like(X)
    dA = torch.empty_like(A)
    stream = torch.cuda.current_stream()
    ms_f = _time_steps(lambda: K.kronmult_batched(d, n, A, X, Y, W, batch, stream=stream),
                       args.steps, args.warmup, stream)
    ms_a = _time_steps(lambda: K.kronmult_batched_adjoint(d, n, A, X, dY, dX, dA, W, batch,
                                                          stream=stream),
                       args.steps, args.warmup, stream)
    # the pointer layout with private entries (identity maps): dX and dA are
    # added into zeroed buffers
    def rows(t, idx):
        base = t.data_ptr()
        rb = t.stride(0) * t.element_size()
        return torch.from_numpy(base + np.asarray(idx, dtype=np.int64) * rb).to(dev)

    A2, X2, dY2 = A.view(batch * d, -1), X.view(batch, E), dY.view(batch, E)
    dX2, dA2 = torch.zeros_like(X2), torch.zeros_like(A2)
    ids, fids = np.arange(batch), np.arange(batch * d)
    pA, pX, pdY = rows(A2, fids), rows(X2, ids), rows(dY2, ids)
    pdX, pdA = rows(dX2, ids), rows(dA2, fids)
    ms_x = _time_steps(lambda: K.kronmult_xbatched_adjoint(d, n, pA, n, pX, pdY, pdX, pdA, n, W,
                                                           batch, stream=stream),
                       args.steps, args.warmup, stream)
    # the same through the torch.autograd wrapper: Y = kron_matvec(A, X) (a
    # fresh zeroed Y, as the kernel accumulates), then Y.backward(dY)
    from kronmult_amd.torch_ops import kron_matvec

    Aw = A.view(batch, d, n, n).transpose(-1, -2).contiguous().requires_grad_(True)
    Xw = X.view(batch, E).clone().requires_grad_(True)
    dYw = dY.view(batch, E)

    def fwd_bwd():
        Aw.grad = None
        Xw.grad = None
        kron_matvec(Aw, Xw).backward(dYw)

    with torch.no_grad():
        ms_wf = _time_steps(lambda: kron_matvec(Aw, Xw), args.steps, args.warmup, stream)
    ms_wfb = _time_steps(fwd_bwd, args.steps, args.warmup, stream)
    fwd_bytes = bytes_per_entry(d, n, ptr_layout=False) * batch
    adj_bytes = 8 * (3 * E + 2 * d * n * n) * batch  # X, dY read, dX written, A read, dA written
    print(json.dumps({"workload": "forward + adjoint d=6 n=6 batchCount=4096", "dtype": "f64",
                      "forward_ms": ms_f, "adjoint_ms": ms_a, "adjoint_over_forward": ms_a / ms_f,
                      "xbatched_adjoint_ms": ms_x,
                      "forward_alg_gbs": fwd_bytes / (ms_f * 1e-3) / 1e9,
                      "adjoint_alg_gbs": adj_bytes / (ms_a * 1e-3) / 1e9,
                      "autograd_forward_ms": ms_wf, "autograd_forward_backward_ms": ms_wfb,
                      "autograd_backward_ms": ms_wfb - ms_wf}), flush=True)


def bench_xadjoint(args):
    """Forward and pointer-layout adjoint of the pooled d=6 n=4 workload:
    65,536 entries, X pool 8,192, Y pool 2,048, FP64.  The adjoint adds
    K^T dY into the X-pool gradient and each entry's factor gradients."""
    import torch

    import kronmult_amd as K

    d, n, batch, nx, ny = 6, 4, 65536, 8192, 2048
    E = n ** d
    dev = "cuda:0"
    A, X, x_map, y_map = make_xbatched_inputs(d, n, batch, nx, ny, seed=2024)
    Ad = torch.from_numpy(A).to(dev)
    Xd = torch.from_numpy(X).to(dev)
    Yd = torch.zeros((ny, E), dtype=torch.float64, device=dev)
    dYd = torch.from_numpy(_splitmix_uniform(ny * E, 77).reshape(ny, E)).to(dev)
    dXd = torch.zeros_like(Xd)
    dAd = torch.zeros_like(Ad)
    per = K.adjoint_workspace_elems(d, n, torch.float64)
    W = torch.empty(max(per, 1) * batch, dtype=torch.float64, device=dev)

    def rows(t, idx):
        base = t.data_ptr()
        rb = t.stride(0) * t.element_size()
        return torch.from_numpy(base + np.asarray(idx, dtype=np.int64) * rb).to(dev)

    pA, pdA = rows(Ad, np.arange(batch * d)), rows(dAd, np.arange(batch * d))
    pX, pdX = rows(Xd, x_map), rows(dXd, x_map)
    pY, pdY = rows(Yd, y_map), rows(dYd, y_map)
    Wd = torch.empty((1, E), dtype=torch.float64, device=dev)
    pW = rows(Wd, np.zeros(batch, dtype=np.int64))
    stream = torch.cuda.current_stream()

    def fwd():
        K.kronmult_xbatched(d, n, pA, n, pX, pY, pW, batch, 0, dtype=torch.float64,
                            stream=stream)

    def adj():
        K.kronmult_xbatched_adjoint(d, n, pA, n, pX, pdY, pdX, pdA, n, W if per else None,
                                    batch, dtype=torch.float64, stream=stream)

    preheat(fwd, args.preheat)
    ms_f = _time_steps(fwd, args.steps, args.warmup, stream)
    ms_a = _time_steps(adj, args.steps, args.warmup, stream)
    flops_f = 2 * d * n ** (d + 1) * batch
    print(json.dumps({"workload": "xbatched forward + adjoint d=6 n=4 batchCount=65536 "
                                  "Xpool=8192 Ypool=2048", "dtype": "f64",
                      "forward_ms": ms_f, "adjoint_ms": ms_a,
                      "adjoint_over_forward": ms_a / ms_f,
                      "forward_tflops": flops_f / (ms_f * 1e-3) / 1e12,
                      "adjoint_scratch_elems_per_entry": per}), flush=True)


def bench_mixed(args):
    """Mixed-shape workload: 8,192 entries with d uniform in 2..6 and n in
    3..8 (n^d <= 65,536 enforced by redrawing: 28 classes), A and X uniform
    [-1,1], one X per entry, Y targets from a pool of 1,024 per class, FP64.
    The persistent autotuned plan (50 applies per hipGraph replay) against
    the reference's way, one vbatched launch per class per apply: with this
    library's per-class kernels, and with the plain per-entry kernel (the
    reference's structure: one workgroup per entry, one pass through W per
    mode)."""
    import tempfile

    import torch

    import helpers as H
    import kronmult_amd as K

    dev = "cuda:0"
    counts = H.mixed_class_counts(8192, 2024)
    stream = torch.cuda.current_stream()
    keep, classes, per_class = [], [], []
    alg_bytes = 0
    for i, ((d, n), batch) in enumerate(counts.items()):
        E = n ** d
        A = torch.from_numpy(_splitmix_uniform(batch * d * n * n, 4000 + 3 * i)).to(dev)
        X = torch.from_numpy(_splitmix_uniform(batch * E, 4001 + 3 * i)).to(dev)
        Y = torch.zeros((1024, E), dtype=torch.float64, device=dev)
        y_map = _splitmix_index(batch, 4002 + 3 * i, 1024)
        pA = A.data_ptr() + np.arange(batch * d, dtype=np.int64) * (n * n * 8)
        pX = X.data_ptr() + np.arange(batch, dtype=np.int64) * (E * 8)
        pY = Y.data_ptr() + y_map * (E * 8)
        pA, pX, pY = (torch.from_numpy(p).to(dev) for p in (pA, pX, pY))
        # vbatched scratch: room for every entry's slot of 2*n^d elements
        W = torch.empty(2 * E * batch, dtype=torch.float64, device=dev)
        keep += [A, X, Y, W, pA, pX, pY]
        classes.append(K.MixedClass(d, n, pA, pX, pY, batch))
        per_class.append((d, n, pA, pX, pY, W, batch))
        alg_bytes += bytes_per_entry(d, n) * batch
    entries = sum(counts.values())

    def per_class_apply():
        for d, n, pA, pX, pY, W, batch in per_class:
            K.kronmult_vbatched([n] * d, [n] * d, pA, pX, pY, W, W.numel() * 8, batch,
                                dtype=torch.float64, stream=stream)

    def per_class_plain_apply():
        for d, n, pA, pX, pY, W, batch in per_class:
            K.kronmult_vbatched_plain([n] * d, [n] * d, pA, pX, pY, W, W.numel() * 8, batch,
                                      dtype=torch.float64, stream=stream)

    cache = os.path.join(tempfile.gettempdir(), "kronmult_amd_tune_%d.txt" % os.getpid())
    t0 = time.perf_counter()
    plan = K.MixedPlan(classes, tune_cache=cache, tune=True, stream=stream)
    tune_s = time.perf_counter() - t0
    if os.path.exists(cache):
        os.remove(cache)
    applies = 50
    ms_eager = _time_steps(lambda: plan.execute(stream), args.steps, args.warmup, stream)
    plan.capture(applies)
    reps = max(1, args.steps // 10)
    ms_graph = _time_steps(lambda: plan.replay(stream), reps, 1, stream) / applies
    ms_pc = _time_steps(per_class_apply, args.steps, args.warmup, stream)
    ms_plain = _time_steps(per_class_plain_apply, max(1, args.steps // 4), 1, stream)
    # the per-class launches captured in a graph as well (launch overhead removed)
    g = torch.cuda.CUDAGraph()
    per_class_apply()
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        for _ in range(applies):
            for d, n, pA, pX, pY, W, batch in per_class:
                K.kronmult_vbatched([n] * d, [n] * d, pA, pX, pY, W, W.numel() * 8, batch,
                                    dtype=torch.float64)
    ms_pc_graph = _time_steps(g.replay, reps, 1, stream) / applies
    kinds = {}
    for i, (d, n) in enumerate(counts):
        info = plan.class_info(i)
        kinds["%d,%d" % (d, n)] = "%s/v%d%s" % (info["kind"], info["variant"],
                                               "/atomic" if info["atomic"] else "")
    li = plan.launch_info()
    plan.close()
    out = {
        "workload": "mixed vbatched: 8192 entries, 28 (d,n) classes, d 2..6, n 3..8, "
                    "Ypool 1024/class, 50 applies per hipGraph",
        "dtype": "f64",
        "classes": len(counts),
        "entries": entries,
        "alg_bytes_per_apply": alg_bytes,
        "persistent_graph_ms_per_apply": ms_graph,
        "persistent_eager_ms_per_apply": ms_eager,
        "per_class_launch_ms_per_apply": ms_pc,
        "per_class_graph_ms_per_apply": ms_pc_graph,
        "plain_per_entry_ms_per_apply": ms_plain,
        "speedup_vs_plain_per_entry": ms_plain / ms_graph,
        "speedup_vs_per_class_launches": ms_pc / ms_graph,
        "speedup_vs_per_class_graph": ms_pc_graph / ms_graph,
        "entries_per_s": entries / (ms_graph * 1e-3),
        "alg_gbs": alg_bytes / (ms_graph * 1e-3) / 1e9,
        "frac": alg_bytes / (ms_graph * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "tune_s": tune_s,
        "launch": li,
        "mapping": kinds,
    }
    print(json.dumps(out), flush=True)


def bench_large_n(args):
    """Large-n dense path: d=3 n=32 and d=4 n=16, 2,048 entries each, FP64 and
    FP32 on the matrix cores (batched layout, A and X normal(0, 1/n), W the
    reference's (n^d, batch) workspace).  FP32 error stated against the FP64
    oracle on the first entries."""
    import torch

    import helpers as H
    import kronmult_amd as K
    import oracle

    dev = "cuda:0"
    stream = torch.cuda.current_stream()
    peak = {"f64": 78.6, "f32": 157.3}  # dense MFMA TFLOP/s
    rows = []
    for d, n in ((3, 32), (4, 16)):
        batch, E = 2048, n ** d
        s = 1.0 / np.sqrt(n)
        A64 = H.normal(batch * d * n * n, 501 + d, s)
        X64 = H.normal(batch * E, 502 + d, s)
        for dt, tdt in (("f64", torch.float64), ("f32", torch.float32)):
            A = torch.from_numpy(A64).to(dev, tdt)
            X = torch.from_numpy(X64).to(dev, tdt)
            Y = torch.zeros(batch * E, dtype=tdt, device=dev)
            W = torch.empty(batch * E, dtype=tdt, device=dev)
            ms = _time_steps(lambda: K.kronmult_batched(d, n, A, X, Y, W, batch, stream=stream),
                             args.steps, args.warmup, stream)
            w = 8 if dt == "f64" else 4
            alg = w * (3 * E + d * n * n) * batch
            flops = 2 * d * n ** (d + 1) * batch
            row = {"d": d, "n": n, "dtype": dt, "batch": batch, "ms": ms,
                   "entries_per_s": batch / (ms * 1e-3),
                   "alg_gbs": alg / (ms * 1e-3) / 1e9,
                   "hbm_frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "tflops": flops / (ms * 1e-3) / 1e12,
                   "mfma_frac": flops / (ms * 1e-3) / 1e12 / peak[dt]}
            if dt == "f32":
                # one apply from Y = 0 on 4 entries against the FP64 oracle
                k = 4
                Ys = torch.zeros(k * E, dtype=tdt, device=dev)
                K.kronmult_batched(d, n, A[:k * d * n * n], X[:k * E], Ys, W[:k * E], k,
                                   stream=stream)
                torch.cuda.synchronize()
                a32 = A64[:k * d * n * n].astype(np.float32).astype(np.float64)
                x32 = X64[:k * E].astype(np.float32).astype(np.float64)
                want = np.zeros(k * E)
                oracle.batched(oracle.oracle_lib(), d, n, a32, x32.copy(), want,
                               np.zeros(k * E), k)
                row["f32_rel_err_vs_f64_oracle"] = H.rel_err(Ys.cpu().numpy(), want)
            rows.append(row)
            del A, X, Y, W
    print(json.dumps({"workload": "large-n dense path (MFMA), batched layout",
                      "cases": rows}), flush=True)


def bench_fused_shapes(args):
    """Batched shapes of the fused kernels, one JSON line each: d = 2..6,
    n = 2..8, every n^d from 8 to 20,480 elements (40,960 in FP32, --dtype
    f32); 2^25 / n^d entries (at most 4M), uniform [0, 1) data; algorithmic
    bytes X + Y read-modify-write + A."""
    import torch
    import kronmult_amd as K
    dev = "cuda:0"
    stream = torch.cuda.current_stream()
    torch.manual_seed(2024)
    f32 = args.dtype == "f32"
    tdt, w = (torch.float32, 4) if f32 else (torch.float64, 8)
    for d in range(2, 7):
        for n in range(2, 9):
            E = n ** d
            if E > (40960 if f32 else 20480) or E < 8:
                continue
            batch = max(64, min(1 << 22, (1 << 25) // E))
            A = torch.rand(batch * d * n * n, dtype=tdt, device=dev)
            X = torch.rand(batch * E, dtype=tdt, device=dev)
            Y = torch.zeros(batch * E, dtype=tdt, device=dev)
            W = torch.empty(batch * E, dtype=tdt, device=dev)
            ms = _time_steps(lambda: K.kronmult_batched(d, n, A, X, Y, W, batch, stream=stream),
                             args.steps, args.warmup, stream)
            alg = w * (3 * E + d * n * n) * batch
            print(json.dumps({"workload": "kronmult_batched fused shape", "d": d, "n": n,
                              "dtype": args.dtype, "batch": batch, "ms": round(ms, 4),
                              "alg_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
                              "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}),
                  flush=True)
            del A, X, Y, W


def bench_kgemm(args):
    """kgemm_nn_batched / kgemm_nt_batched (FP64), one JSON line per case:
    C_b = A_b op(B_b) + C_b (alpha = 1, beta = 1) over contiguous per-entry
    operands, shapes of the recursion's steps at d = 6, n = 4 (mm = 1024,
    nn = kk = 4), a square 16^3 case and 4^3 entries; algorithmic bytes per entry
    8*(mm*kk + kk*nn + 2*mm*nn) (C read and written; once with beta = 0)."""
    import torch
    import kronmult_amd as K
    dev = "cuda:0"
    stream = torch.cuda.current_stream()
    torch.manual_seed(2024)
    # beta = 1 adds alpha*AB into C with device atomics (entries may share C,
    # kgemm_nn.hpp:392-394); beta = 0 stores it (C is written, not read)
    for tr, mm, nn, kk, batch, beta in (("nt", 1024, 4, 4, 65536, 1.0),
                                        ("nn", 1024, 4, 4, 65536, 1.0),
                                        ("nn", 1024, 4, 4, 65536, 0.0),
                                        ("nn", 16, 16, 16, 262144, 1.0),
                                        ("nt", 16, 16, 16, 262144, 1.0),
                                        ("nn", 16, 16, 16, 262144, 0.0),
                                        ("nn", 4, 4, 4, 1 << 20, 1.0)):
        A = torch.rand(batch * mm * kk, dtype=torch.float64, device=dev)
        B = torch.rand(batch * kk * nn, dtype=torch.float64, device=dev)
        C = torch.zeros(batch * mm * nn, dtype=torch.float64, device=dev)
        idx = torch.arange(batch, dtype=torch.int64, device=dev)
        pA = A.data_ptr() + idx * (mm * kk * 8)
        pB = B.data_ptr() + idx * (kk * nn * 8)
        pC = C.data_ptr() + idx * (mm * nn * 8)
        ldA = torch.full((batch,), mm, dtype=torch.int32, device=dev)
        ldB = torch.full((batch,), kk if tr == "nn" else nn, dtype=torch.int32, device=dev)
        ldC = torch.full((batch,), mm, dtype=torch.int32, device=dev)
        f = K.kgemm_nn_batched if tr == "nn" else K.kgemm_nt_batched
        ms = _time_steps(lambda: f(mm, nn, kk, 1.0, pA, ldA, pB, ldB, beta, pC, ldC, batch,
                                   stream=stream), args.steps, args.warmup, stream)
        alg = 8 * (mm * kk + kk * nn + (2 if beta != 0 else 1) * mm * nn) * batch
        print(json.dumps({"workload": "kgemm_%s_batched" % tr, "mm": mm, "nn": nn, "kk": kk,
                          "beta": beta, "dtype": "f64", "batch": batch, "ms": round(ms, 4),
                          "alg_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
                          "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}),
              flush=True)
        del A, B, C


def bench_shapes(args):
    """Batched shapes beyond the fused kernels, one JSON line each: d = 1 from
    n = 8 (matrix-vector kernel), d = 2 at n = 9..32 (a wave per entry), d = 3
    and 4 at n = 9..15 (generic kernel; row blocks of the factor in registers
    where they measured faster), FP64 and FP32.  2^25 / n^d entries per shape
    (about 800 MB of X, Y and W in FP64); uniform [0, 1) data from torch's
    generator; algorithmic bytes are X + Y read-modify-write + the factors."""
    import torch
    import kronmult_amd as K
    dev = "cuda:0"
    stream = torch.cuda.current_stream()
    shapes = [(1, 16), (1, 64), (2, 12), (2, 24), (2, 32)] + \
             [(3, n) for n in (9, 10, 11, 12, 13, 14, 15)] + \
             [(4, n) for n in (9, 10, 11, 12, 13, 14, 15)]
    torch.manual_seed(2024)
    for dt, tdt in (("f64", torch.float64), ("f32", torch.float32)):
        w = 8 if dt == "f64" else 4
        for d, n in shapes:
            E = n ** d
            batch = max(1, (1 << 25) // E)
            A = torch.rand(batch * d * n * n, dtype=tdt, device=dev)
            X = torch.rand(batch * E, dtype=tdt, device=dev)
            Y = torch.zeros(batch * E, dtype=tdt, device=dev)
            W = torch.empty(batch * E, dtype=tdt, device=dev)
            ms = _time_steps(lambda: K.kronmult_batched(d, n, A, X, Y, W, batch, stream=stream),
                             args.steps, args.warmup, stream)
            alg = w * (3 * E + d * n * n) * batch
            print(json.dumps({"workload": "kronmult_batched shape", "d": d, "n": n,
                              "dtype": dt, "batch": batch, "ms": round(ms, 4),
                              "entries_per_s": batch / (ms * 1e-3),
                              "alg_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
                              "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}),
                  flush=True)
            del A, X, Y, W


def bench_copy(args):
    """Device-to-device copy of 2 GiB: the achievable HBM read+write rate."""
    import torch
    n = 1 << 28  # doubles
    a = torch.empty(n, dtype=torch.float64, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    ms = _time_steps(lambda: b.copy_(a), args.steps, args.warmup, stream)
    gbs = 2 * n * 8 / (ms * 1e-3) / 1e9
    print(json.dumps({"workload": "copy 2 GiB (read + write)", "ms": ms, "gbs": gbs,
                      "frac": gbs / HBM_PEAK_GBS}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # the clocks ramp over the first tens of milliseconds under load (headline
    # calls 0.55 -> 0.40 ms over the first ~30): warm up for ~80 ms
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="xbatched_d6n4")
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"],
                    help="element type of the fused_shapes workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--preheat", type=float, default=0.25,
                    help="seconds of back-to-back headline calls before the warmup")
    ap.add_argument("--cpu-sample", type=int, default=65536)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "xbatched_d6n4":
        bench_xbatched(args)
    elif args.workload == "batched_d6n4":
        bench_batched(args)
    elif args.workload == "adjoint_d6n6":
        bench_adjoint(args)
    elif args.workload == "xadjoint_d6n4":
        bench_xadjoint(args)
    elif args.workload == "large_n":
        bench_large_n(args)
    elif args.workload == "mixed_sparse_grid":
        bench_mixed(args)
    elif args.workload == "copy":
        bench_copy(args)
    elif args.workload == "shapes":
        bench_shapes(args)
    elif args.workload == "fused_shapes":
        bench_fused_shapes(args)
    elif args.workload == "kgemm":
        bench_kgemm(args)
    else:
        print("unknown workload %s" % args.workload, file=sys.stderr)
        sys.exit(2)


if __name__ == "__main__":
    main()
