#!/usr/bin/env python3
"""Benchmark: batched MixDom2 K=3 forward+backward+counts on one MI355X.

One step = one launch of the batched kernel over the whole c1 workload
(1024 MixDom2 K=3 F=1 C=1 pairs, 17 states, ancestor lengths
round(LogNormal(ln 200, 0.5)) clipped to [30, 800], t ~ U(0.05, 2)) with all
inputs already resident in HBM.  Prints one JSON line.

  value       = sum_pairs (Lx+1)(Ly+1) S / step time         (cell-states/s)
  roofline    = HBM-bound: lattice traffic per cell-state of the pairs in
                the dominant kernel launch (the launch moving the most
                algorithmic bytes: 8 B per cell-state for a forward launch,
                which writes the lattice, or a backward launch, which reads
                it; 16 B for a launch doing both), divided by that launch's
                HIP-event duration on its own stream, against 8 TB/s
  cpu_baseline= the C restatement of the reference recursion (log space,
                OpenMP over pairs) on bounded samples of the same pairs, on
                all the host cores the job may use and on one core
  plain       = the per-pair drop-in (one launch per pair, as the reference's
                per-pair forward_backward_2d call) on a sample of pairs
  configs     = the other workloads, each timed on its own: c0 (TKF91
                forward-only, 64 pairs), c2 (MixDom2 K=8 F=4 C=4, 162 states,
                wide kernel), mid (the c1 shape at K=3 F=2, 32 states), c3 (gradients of the summed MixDom2 K=3 F=1 C=2
                log-likelihood over 200 pairs, forward + adjoint), fb1d (the
                1D alignment-constrained forward/backward + counts over the
                c1 pairs' generating alignments)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "tkf-mixdom_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "fwd+bwd+counts DP cell-states/s, MixDom2 K=3 batched vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STATE = 16.0
CHECKPOINT_BUDGET = 256 << 20  # the library's default per-pair lattice budget
# HBM bytes per launch, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this
# workload (KiB per dispatch, scaled by the copy-kernel calibration: fetch x2.00,
# write x1.00), keyed by (launch, pairs in it):
# profiles/fbc_kernels_c1_mixdom2_k3_long_one_launch.md (the default class
# configuration: the 6 longest pairs on the one-launch schedule),
# profiles/fbc_kernels_c1_mixdom2_k3.md (no long class) and
# profiles/fbv2_kernel_c1_mixdom2_k3.md (one-workgroup-per-pair kernel).
TRAFFIC = {("fbc_bwd_kernel<LMix<3,1>,3>", 817): 8.50e9, ("fbc_fwd_kernel<LMix<3,1>,3>", 817): 7.72e9,
           ("fbc_bidi_kernel<LMix<3,1>,3> long", 6): 0.90e9,
           ("fbc_bwd_kernel<LMix<3,1>,3>", 823): 8.69e9, ("fbc_fwd_kernel<LMix<3,1>,3>", 823): 8.17e9,
           ("fbc_bwd_kernel<LMix<3,1>,2>", 189): 3.25e8, ("fbc_fwd_kernel<LMix<3,1>,2>", 189): 2.74e8,
           ("fbv2_kernel<LMix<3,1>,8>", 33): 3.13e9, ("fbv2_kernel<LMix<3,1>,4>", 269): 8.28e9,
           ("fbv2_kernel<LMix<3,1>,2>", 506): 5.14e9, ("fbv2_kernel<LMix<3,1>,1>", 216): 0.665e9}


def _bytes_per_cell_state(kernel_name):
    """Algorithmic lattice bytes per cell-state of one launch: the forward
    writes the lattice (8 B), the backward reads it (8 B); a kernel that does
    both in one launch moves 16 B."""
    if kernel_name.startswith(("fbc_fwd", "fbc_bwd")):
        return BYTES_PER_CELL_STATE / 2
    return BYTES_PER_CELL_STATE


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_cores():
    """Cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (the share a GPU box gives a job)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(w, seconds=15.0, single_seconds=8.0):
    """The C oracle on bounded samples of the workload's pairs: once on all
    the host cores this process may use, once on one core."""
    cores = _host_cores()
    allc = _cpu_sample(w, seconds, cores)
    one = _cpu_sample(w, single_seconds, 1)
    allc.update(cpu_model=_cpu_model(), nproc=os.cpu_count(),
                single_core={"value": one["value"], "unit": one["unit"], "cores": 1,
                             "sample": one["sample"]})
    return allc


def _cpu_sample(w, seconds, threads):
    """Time the C oracle on a bounded sample of the workload's pairs."""
    import numpy as np
    from oracle import cfb2d
    rng = np.random.default_rng(123)
    order = rng.permutation(w.B)
    lt = w.log_trans.numpy()
    tab = w.log_tab.numpy()
    # calibrate on a few pairs, then size the sample to ~`seconds`
    done_cells, elapsed, used = 0, 0.0, []
    for chunk in np.array_split(order, max(1, w.B // threads)):
        if elapsed > seconds:
            break
        idx = list(chunk)
        t0 = time.perf_counter()
        cfb2d.fb2d_batch(lt[idx], w.state_types, w.state_group, tab[idx],
                         [w.xs[i] for i in idx], [w.ys[i] for i in idx], nthreads=threads)
        elapsed += time.perf_counter() - t0
        done_cells += int(sum((len(w.xs[i]) + 1) * (len(w.ys[i]) + 1) for i in idx)) * w.S
        used += idx
    return {"value": done_cells / elapsed, "unit": "cell-states/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(used)} of {w.B} c1 pairs ({done_cells:.3e} cell-states, "
                      f"{elapsed:.1f} s, C log-space restatement, OpenMP over pairs)"}


def plain_per_pair(w, n_pairs=48, emit=True):
    """Per-pair drop-in: one launch per pair (the reference calls
    forward_backward_2d per pair), timed over a sample of pairs."""
    import numpy as np
    import torch
    from tkfmi import _abi
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(7)
    idx = rng.choice(w.B, size=min(n_pairs, w.B), replace=False)
    runners = []
    for i in idx:
        pp = _abi.PackedPairs([w.xs[i]], [w.ys[i]], dev)
        runners.append(_abi.BatchRunner(w.log_trans[i].to(dev), w.state_types, w.state_group,
                                        w.log_tab[i].to(dev), pp, want_emit_counts=emit))
    for r in runners[:4]:
        r.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in runners:
        r.run()
        torch.cuda.synchronize()
    el = time.perf_counter() - t0
    cells = int(sum((len(w.xs[i]) + 1) * (len(w.ys[i]) + 1) for i in idx)) * w.S
    return {"value": cells / el, "unit": "cell-states/s", "pairs": len(idx),
            "ms_per_pair": 1e3 * el / len(idx)}


def serve_bench(loads, n_requests, max_seconds=10.0):
    """Serving: a Poisson stream of ``n_requests`` scoring / E-step requests,
    50/50 TKF92 (5 states) / MixDom2 K=3 (17 states), lengths
    round(LogNormal(ln 250, 0.5)) clipped to [20, 2000], at each offered
    load (requests/s), served in real time by the continuous-batching loop
    after a warm-up stream.  At low loads the stream is cut to about
    ``max_seconds`` of arrivals (e.g. 1000 requests at 100 req/s).  Reports
    p50 / p99 latency per load."""
    import numpy as np
    from tkfmi import serve, synth
    rng = np.random.default_rng(4)
    Q, pi, _ = synth.random_Q_pi(rng)
    eng = serve.ServingEngine(dict(lam=0.04, mu=0.05, ext=float(rng.uniform(0.5, 0.95)), Q=Q, pi=pi),
                              synth.mixdom2_params(rng, 3, 1, 1))
    eng.serve(serve.poisson_requests(99, 300, rate=1e6), realtime=False)  # warm-up
    S_of = {"tkf92": 5, "mixdom2": 17}
    out = []
    for load in loads:
        n = int(min(n_requests, max(200, load * max_seconds)))
        reqs = serve.poisson_requests(4, n, rate=load)
        t0 = time.perf_counter()
        res = eng.serve(reqs, realtime=True)
        wall = time.perf_counter() - t0
        s = serve.summarize(reqs, res, wall, S_of)
        s["offered_req_per_s"] = load
        s["requests"] = n
        _log(f"serving at {load} req/s: p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms, "
             f"{s['achieved_req_per_s']:.0f} req/s, {s['cell_states_per_s']:.3e} cell-states/s")
        out.append(s)
    return {"workload": f"up to {n_requests} Poisson requests per load (about {max_seconds:.0f} s of "
                        "arrivals at low loads), 50/50 TKF92 (S=5) / MixDom2 K=3 (S=17), "
                        "50/50 scoring / E-step, lengths LogNormal(ln250,0.5) in [20,2000]",
            "loads": out}


def _time_calls(fn, steps, warmup=1):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return 1e3 * (time.perf_counter() - t0) / steps


def _kernel_ms(runner_fn):
    from tkfmi import _abi
    runner_fn()
    return {k["name"]: round(k["ms"], 4) for k in _abi.last_kernel_times()}


def fb1d_bench(w, steps):
    """1D E-step DP over the alignments the c1 pairs were generated from:
    per-pair chi and per-class tables of the c1 workload, emissions looked up
    on the device, transition + emission counts."""
    import torch
    from tkfmi import _abi, synth
    dev = torch.device("cuda", 0)
    cts, ancs, descs = synth.alignment_columns(w)
    cols = _abi.PackedColumns(cts, ancs, descs, dev)
    r = _abi.Runner1D(w.log_trans.to(dev), init_state=0, final_state=1, columns=cols,
                      state_types=w.state_types, state_group=w.state_group,
                      log_tab=w.log_tab.to(dev), want_emit_counts=True)
    ms = _time_calls(r.run, steps)
    r.flags |= _abi.TIME_KERNELS
    kern = _kernel_ms(r.run)
    r.flags &= ~_abi.TIME_KERNELS
    col_states = cols.rows * w.S
    return {"value": col_states / (ms / 1e3), "unit": "column-states/s", "ms": ms,
            "pairs": w.B, "columns": cols.rows, "states": w.S, "kernels_ms": kern,
            "workload": "c1 pairs' generating alignments, MixDom2 K=3 (17 states), "
                        "fwd+bwd+transition and emission counts"}


def c0_bench(steps):
    """c0: TKF91 forward-only log-likelihood of 64 pairs (lengths U{50..150},
    t = 0.5)."""
    import torch
    from tkfmi import _abi, synth
    w = synth.make_tkf(seed=0, B=64)
    dev = torch.device("cuda", 0)
    r = _abi.BatchRunner(w.log_trans.to(dev), w.state_types, w.state_group, w.log_tab.to(dev),
                         _abi.PackedPairs(w.xs, w.ys, dev), forward_only=True)
    ms = _time_calls(r.run, steps)
    cs = w.cell_states()
    return {"value": cs / (ms / 1e3), "unit": "cell-states/s (forward only)", "ms": ms,
            "pairs": w.B, "cell_states": cs}


def c2_bench(pairs, steps=2, budget=0):
    """c2: MixDom2 K=8 F=4 C=4 (162 states) forward + backward + transition
    and emission counts, lengths U{100..1000}, on the banded kernel: with
    full lattices (checkpoint budget 0: 121 GiB of workspace for the 256
    pairs), or with the pairs whose lattice exceeds ``budget`` bytes
    checkpointed (each backward band recomputes its lattice).  One warm-up
    call, then the mean of ``steps`` timed calls."""
    import numpy as np
    import torch
    from tkfmi import _abi, synth
    rng = np.random.default_rng(2)
    w = synth.make_mixdom2(seed=2, B=pairs, K=8, F=4, C=4, lengths=rng.integers(100, 1001, pairs),
                           name="c2")
    dev = torch.device("cuda", 0)
    _abi.set_checkpoint_budget(budget)
    try:
        r = _abi.BatchRunner(w.log_trans.to(dev), w.state_types, w.state_group, w.log_tab.to(dev),
                             _abi.PackedPairs(w.xs, w.ys, dev), want_emit_counts=True)
        ms = _time_calls(r.run, steps)
        r.flags |= _abi.TIME_KERNELS
        kern = _kernel_ms(r.run)
        ws = r.ws_bytes
        del r
    finally:
        _abi.set_checkpoint_budget(_abi.CHECKPOINT_BUDGET_DEFAULT)
    torch.cuda.empty_cache()
    cs = w.cell_states()
    # 6 S flops per cell-state (forward, backward, count products)
    tf = 6.0 * w.S * cs / (ms / 1e3) / 1e12
    return {"value": cs / (ms / 1e3), "unit": "cell-states/s", "ms": ms, "pairs": w.B,
            "cell_states": cs, "fp64_tflops": tf, "frac_of_fp64_mfma_peak": tf / 78.6,
            "emission_counts": True, "checkpoint_budget": budget, "kernels_ms": kern,
            "workspace_GiB": ws / 2**30}


def mid_bench(steps, K=3, F=2):
    """The c1 workload's shape (1024 pairs, LogNormal(ln200,0.5) lengths in
    [30,800]) on a mid-size MixDom2 layout: K=3 F=2 (32 states), K=5 F=2 (52)
    or K=8 F=2 (82), the sizes of the reference's trained models:
    fwd+bwd+transition and emission counts in one call (banded kernel)."""
    import torch
    from tkfmi import _abi, synth
    w = synth.make_mixdom2(seed=1, B=1024, K=K, F=F, C=1, name=f"c1_k{K}f{F}")
    dev = torch.device("cuda", 0)
    r = _abi.BatchRunner(w.log_trans.to(dev), w.state_types, w.state_group, w.log_tab.to(dev),
                         _abi.PackedPairs(w.xs, w.ys, dev), want_emit_counts=True)
    ms = _time_calls(r.run, steps)
    r.flags |= _abi.TIME_KERNELS
    kern = _kernel_ms(r.run)
    cs = w.cell_states()
    out = {"value": cs / (ms / 1e3), "unit": "cell-states/s", "ms": ms, "pairs": w.B,
           "states": w.S, "cell_states": cs, "kernels_ms": kern}
    del r
    torch.cuda.empty_cache()
    return out


def critical_path(w, emit):
    """The longest pair of the batch alone, its launch times (HIP events): a
    call cannot end before its longest pair's chain of antidiagonal steps
    does, whatever the batch around it.  A lone pair fits the GPU, so it runs
    on the chained kernels' one-launch schedule (forward and backward
    together, meeting at an antidiagonal)."""
    import numpy as np
    import torch
    from tkfmi import _abi
    dev = torch.device("cuda", 0)
    b = int(np.argmax([(len(x) + 1) * (len(y) + 1) for x, y in zip(w.xs, w.ys)]))
    r = _abi.BatchRunner(w.log_trans[b].to(dev), w.state_types, w.state_group, w.log_tab[b].to(dev),
                         _abi.PackedPairs([w.xs[b]], [w.ys[b]], dev), want_emit_counts=emit,
                         time_kernels=True)
    for _ in range(3):
        r.run()
    ks = []
    for _ in range(5):
        r.run()
        kt = _abi.last_kernel_times()
        ks.append(sum(k["ms"] for k in kt))
    torch.cuda.synchronize()
    Lx, Ly = len(w.xs[b]), len(w.ys[b])
    return {"pair": [Lx, Ly], "ms": float(np.median(ks)),
            "antidiagonal_steps": Lx + Ly + 1, "kernels": sorted({k["name"] for k in kt})}


def c3_bench(steps):
    """c3: gradient of the summed MixDom2 K=3 F=1 C=2 log-likelihood of 200
    pairs (LogNormal median 200) w.r.t. every model parameter and t:
    torch builders + the forward/backward kernel + autograd composition.
    Timed with the builders replayed from HIP graphs (graph=True, the
    fixed-minibatch training step) and, for reference, launched eagerly."""
    import numpy as np
    import torch
    from tkfmi import synth, vjp
    w = synth.make_mixdom2(seed=3, B=200, K=3, F=1, C=2, name="c3")
    pr = w.params
    dev = torch.device("cuda", 0)
    p = {"lam": pr["main_lam"], "mu": pr["main_mu"], "dlam": pr["dom_lam"], "dmu": pr["dom_mu"],
         "dw": pr["dom_w"], "frag_w": pr["frag_w"], "ext": pr["ext"], "pis": pr["class_pi"],
         "S": pr["class_S"], "cd": pr["class_dist"]}
    tp = {k: torch.tensor(np.asarray(v, dtype=np.float64), device=dev, requires_grad=True)
          for k, v in p.items()}
    tt = torch.tensor(w.ts, dtype=torch.float64, device=dev, requires_grad=True)

    def step(graph):
        for v in list(tp.values()) + [tt]:
            v.grad = None
        lp = vjp.mixdom2_log_prob_batch(tp["lam"], tp["mu"], tt, tp["dlam"], tp["dmu"], tp["dw"],
                                        tp["frag_w"], tp["ext"], tp["pis"], tp["S"], tp["cd"],
                                        w.xs, w.ys, graph=graph)
        lp.sum().backward()

    ms_eager = _time_calls(lambda: step(False), steps, warmup=5)
    ms = _time_calls(lambda: step(True), steps, warmup=5)
    cs = w.cell_states()
    return {"value": cs / (ms / 1e3), "unit": "cell-states/s (value + full gradient)", "ms": ms,
            "ms_eager_builders": ms_eager, "builders": "HIP graphs (graph=True)",
            "pairs": w.B, "cell_states": cs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serve-loads", default="100,300,1000,3000,10000",
                    help="offered loads (requests/s) for the serving measurement; empty to skip")
    ap.add_argument("--serve-requests", type=int, default=5000)
    ap.add_argument("--no-emit", action="store_true",
                    help="transition counts only (no expected emission counts)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the c0 / c2 / c3 / fb1d measurements")
    ap.add_argument("--configs", default="fb1d,c0,c3,mid,c2",
                    help="which of the fb1d / c0 / c3 / mid / c2 measurements to run (comma list)")
    ap.add_argument("--c2-pairs", type=int, default=256,
                    help="pairs of the c2 (162-state) measurement; 0 skips it")
    ap.add_argument("--kernel", default="chain", choices=["chain", "class", "generic"],
                    help="2D kernel: chained workgroups per pair (default), one workgroup per "
                         "pair, or the any-layout kernel")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the chained kernels' autotuning (library defaults)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (x5), to calibrate "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE against known traffic")
    args = ap.parse_args()
    sel = set() if args.no_configs else {c.strip() for c in args.configs.split(",") if c.strip()}
    known = {"fb1d", "c0", "c3", "mid", "c2"}
    if sel - known:
        raise SystemExit(f"unknown --configs entries: {sorted(sel - known)}")
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(pairs would be dealt out to GPUs by their DP cost, one process per GPU)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from tkfmi import _abi, synth

    if args.copy_probe_mib:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8, "launches": 5}), flush=True)
        return

    _log("generating c1 workload")
    w = synth.make_mixdom2(seed=1, B=args.pairs)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    pairs = _abi.PackedPairs(w.xs, w.ys, dev)
    lt_dev, tab_dev = w.log_trans.to(dev), w.log_tab.to(dev)

    def make_runner():
        return _abi.BatchRunner(lt_dev, w.state_types, w.state_group, tab_dev, pairs,
                                want_emit_counts=not args.no_emit, kernel=args.kernel)

    tuned = None
    runner = make_runner()
    if args.kernel == "chain" and not args.no_tune:
        # untimed set-up: libtkfmi's autotuner picks the chained kernels'
        # workgroup widths per length class for this batch's key
        from tkfmi import autotune
        _log("autotuning the chained kernels on this batch")
        tuned = autotune.tune(runner)
        _log(f"autotune: key {tuned['key']} best {tuned['best']}: " +
             ", ".join(f"{k}: {v:.2f} ms" for k, v in tuned["ms"].items()))
    cell_states = w.cell_states()
    _log(f"pairs={w.B} S={w.S} cell-states={cell_states:.4e} workspace={runner.ws_bytes/2**30:.2f} GiB")
    for _ in range(args.warmup):
        runner.run()
    torch.cuda.synchronize()
    # untimed set-up beyond the W warm-up steps: calls until at least 0.5 s
    # has passed (the first ~10 calls of a process run up to 1.7x slower while
    # the device and the library's pinned upload buffers warm up; the
    # autotuner's calls do the same when it runs)
    t_warm = time.perf_counter()
    while time.perf_counter() - t_warm < 0.5:
        runner.run()
    torch.cuda.synchronize()
    # per-kernel durations: HIP events recorded by the library on the stream
    # each kernel runs on (one kernel per wave-count class)
    base_flags = runner.flags
    runner.flags = base_flags | _abi.TIME_KERNELS
    per_kernel = {}
    for i in range(args.steps):
        runner.run()
        for k in _abi.last_kernel_times():
            per_kernel.setdefault(k["name"], []).append(k)
    runner.flags = base_flags
    torch.cuda.synchronize()
    # whole-step wall time (outputs zeroed inside the step, as a caller would)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        runner.run()
    torch.cuda.synchronize()
    step_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    # the same call captured once into a HIP graph and replayed (no host work)
    graph = runner.capture()
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    graph_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    logp = runner.log_p.cpu().numpy()
    n = runner.n.cpu().numpy()
    assert np.all(np.isfinite(logp)), "non-finite log-likelihood"
    value = cell_states / (step_ms / 1e3)
    kernels = []
    for name, recs in per_kernel.items():
        ms = float(np.mean([r["ms"] for r in recs]))
        cs = recs[0]["cells"] * w.S
        bpc = _bytes_per_cell_state(name)
        kernels.append({"name": name, "pairs": recs[0]["n_pairs"], "cell_states": cs, "ms": ms,
                        "bytes_per_cell_state": bpc,
                        "achieved_GBs": bpc * cs / (ms / 1e3) / 1e9})
    kernels.sort(key=lambda k: -k["ms"])
    # the dominant launch: the one that moves the most algorithmic bytes (the
    # main class's backward; the long class's one-launch kernel runs longer on
    # a handful of pairs)
    dom = max(kernels, key=lambda k: (k["bytes_per_cell_state"] * k["cell_states"], k["ms"]))
    crit = critical_path(w, not args.no_emit) if args.kernel == "chain" else None
    _log(f"step {step_ms:.3f} ms; kernels: " +
         ", ".join(f"{k['name']} {k['ms']:.3f} ms ({k['pairs']} pairs)" for k in kernels))

    out = {
        "metric": METRIC, "value": value, "unit": "cell-states/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "MixDom2 K=3 F=1 C=1 (17 states) forward+backward+expected "
                               "transition and emission counts, 1024 uneven pairs, "
                               "LogNormal(ln200,0.5) lengths in [30,800], t~U(0.05,2), one call",
                   "pairs": w.B, "states": w.S, "cell_states": cell_states,
                   "emission_counts": not args.no_emit, "kernel": args.kernel,
                   "graph_replay_ms_per_step": graph_ms,
                   "chain_config": (None if tuned is None else
                                    {"long_w": tuned["best"][0], "main_w": tuned["best"][1],
                                     "long_frac": tuned["best"][2], "long_one_launch": tuned["best"][3],
                                     "tuned_ms": {str(k): round(v, 3) for k, v in tuned["ms"].items()}}),
                   "sum_logp": float(logp.sum()), "sum_counts": float(n.sum())},
        "roofline": {"bound": "hbm", "achieved": dom["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": TRAFFIC.get((dom["name"], dom["pairs"])),
                     "kernel": dom["name"], "kernel_ms": dom["ms"],
                     "bytes_per_cell_state": dom["bytes_per_cell_state"], "kernels": kernels},
    }
    if crit is not None:
        # the step's floor: its longest pair alone (forward + backward)
        out["roofline"]["critical_path"] = {
            "longest_pair": crit["pair"], "antidiagonal_steps": crit["antidiagonal_steps"],
            "ms": crit["ms"], "us_per_step": 1e3 * crit["ms"] / crit["antidiagonal_steps"],
            "frac_of_step": crit["ms"] / step_ms, "kernels": crit["kernels"]}
        _log(f"critical path: longest pair {crit['pair']} alone {crit['ms']:.3f} ms "
             f"({crit['ms'] / step_ms:.2f} of the step)")
    if not args.no_plain:
        _log("timing the per-pair drop-in")
        pl = plain_per_pair(w, emit=not args.no_emit)
        out["plain_per_pair"] = pl
        out["batched_over_plain"] = value / pl["value"]
    if sel:
        cfg = {}
        if "fb1d" in sel:
            _log("1D alignment-constrained E-step DP")
            cfg["fb1d"] = fb1d_bench(w, args.steps)
            _log(f"fb1d: {cfg['fb1d']['ms']:.3f} ms, {cfg['fb1d']['value']:.3e} column-states/s")
        if "c0" in sel:
            _log("c0 TKF91 forward-only")
            cfg["c0"] = c0_bench(args.steps)
        if "c3" in sel:
            _log("c3 gradients, 200 pairs")
            cfg["c3"] = c3_bench(max(1, args.steps // 2))
            _log(f"c3: {cfg['c3']['ms']:.1f} ms per gradient "
                 f"(builders launched eagerly: {cfg['c3']['ms_eager_builders']:.1f} ms)")
        if "mid" in sel:
            for K in (3, 5, 8):
                _log(f"c1-shaped batch at K={K} F=2 ({2 + 10 * K} states)")
                key = f"mid_k{K}f2"
                cfg[key] = mid_bench(max(1, args.steps // 2), K=K, F=2)
                _log(f"mid K={K} F=2: {cfg[key]['ms']:.1f} ms, {cfg[key]['value']:.3e} cell-states/s")
        if "c2" in sel and args.c2_pairs > 0:
            _log(f"c2 162-state wide kernel, {args.c2_pairs} pairs")
            cfg["c2"] = c2_bench(args.c2_pairs)
            _log(f"c2: {cfg['c2']['ms']:.0f} ms, {cfg['c2']['fp64_tflops']:.2f} TF/s")
            _log("c2 with the default per-pair checkpoint budget (banded recomputation)")
            cfg["c2_checkpointed"] = c2_bench(args.c2_pairs, budget=CHECKPOINT_BUDGET)
            _log(f"c2 checkpointed: {cfg['c2_checkpointed']['ms']:.0f} ms in "
                 f"{cfg['c2_checkpointed']['workspace_GiB']:.1f} GiB of workspace")
        out["configs"] = cfg
    loads = [float(v) for v in args.serve_loads.split(",") if v.strip()]
    if loads:
        _log("serving measurement")
        out["serving"] = serve_bench(loads, args.serve_requests)
    if not args.no_cpu:
        _log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(w, seconds=args.cpu_seconds)
        _log(f"cpu baseline: {out['cpu_baseline']['value']:.3e} cell-states/s on "
             f"{out['cpu_baseline']['cores']} cores, {out['cpu_baseline']['single_core']['value']:.3e} on one "
             f"({out['cpu_baseline']['cpu_model']})")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
