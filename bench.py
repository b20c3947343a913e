#!/usr/bin/env python3
"""Benchmark: batched MixDom2 K=3 forward+backward+counts on one MI355X.

One step = one launch of the batched kernel over the whole c1 workload
(1024 MixDom2 K=3 F=1 C=1 pairs, 17 states, ancestor lengths
round(LogNormal(ln 200, 0.5)) clipped to [30, 800], t ~ U(0.05, 2)) with all
inputs already resident in HBM.  Prints one JSON line.

  value       = sum_pairs (Lx+1)(Ly+1) S / step time         (cell-states/s)
  roofline    = HBM-bound: 16 B of lattice traffic per cell-state
                (write + read of the forward lattice) of the pairs in the
                dominant kernel launch (the longest of the per-wave-count
                class launches), divided by that launch's HIP-event duration
                on its own stream, against 8 TB/s
  cpu_baseline= the C restatement of the reference recursion (log space,
                OpenMP over pairs) on a bounded sample of the same pairs
  plain       = the per-pair drop-in (one launch per pair, as the reference's
                per-pair forward_backward_2d call) on a sample of pairs
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
# HBM bytes per launch, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this
# workload (KiB per dispatch, scaled by the copy-kernel calibration: fetch x2.00,
# write x1.00); profiles/fbv2_kernel_c1_mixdom2_k3.md.
TRAFFIC = {"fbv2_kernel<LMix<3,1>,8>": 3.13e9, "fbv2_kernel<LMix<3,1>,4>": 8.28e9,
           "fbv2_kernel<LMix<3,1>,2>": 5.14e9, "fbv2_kernel<LMix<3,1>,1>": 0.665e9}


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, seconds=15.0, threads=None):
    """Time the C oracle on a bounded sample of the workload's pairs."""
    import numpy as np
    from oracle import cfb2d
    threads = threads or min(16, os.cpu_count() or 1)
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


def serve_bench(loads, n_requests):
    """Serving: a Poisson stream of ``n_requests`` scoring / E-step requests,
    50/50 TKF92 (5 states) / MixDom2 K=3 (17 states), lengths
    round(LogNormal(ln 250, 0.5)) clipped to [20, 2000], at each offered
    load (requests/s), served in real time by the continuous-batching loop
    after a warm-up stream.  Reports p50 / p99 latency per load."""
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
        reqs = serve.poisson_requests(4, n_requests, rate=load)
        t0 = time.perf_counter()
        res = eng.serve(reqs, realtime=True)
        wall = time.perf_counter() - t0
        s = serve.summarize(reqs, res, wall, S_of)
        s["offered_req_per_s"] = load
        _log(f"serving at {load} req/s: p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms, "
             f"{s['achieved_req_per_s']:.0f} req/s, {s['cell_states_per_s']:.3e} cell-states/s")
        out.append(s)
    return {"workload": f"{n_requests} Poisson requests, 50/50 TKF92 (S=5) / MixDom2 K=3 (S=17), "
                        "50/50 scoring / E-step, lengths LogNormal(ln250,0.5) in [20,2000]",
            "loads": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serve-loads", default="1000,3000,10000",
                    help="offered loads (requests/s) for the serving measurement; empty to skip")
    ap.add_argument("--serve-requests", type=int, default=5000)
    ap.add_argument("--no-emit", action="store_true",
                    help="transition counts only (no expected emission counts)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (x5), to calibrate "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE against known traffic")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(pairs shard across GPUs by cost; see DESIGN.md)", file=sys.stderr)
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
    runner = _abi.BatchRunner(w.log_trans.to(dev), w.state_types, w.state_group,
                              w.log_tab.to(dev), pairs, want_emit_counts=not args.no_emit)
    cell_states = w.cell_states()
    _log(f"pairs={w.B} S={w.S} cell-states={cell_states:.4e} workspace={runner.ws_bytes/2**30:.2f} GiB")
    for _ in range(args.warmup):
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
    logp = runner.log_p.cpu().numpy()
    n = runner.n.cpu().numpy()
    assert np.all(np.isfinite(logp)), "non-finite log-likelihood"
    value = cell_states / (step_ms / 1e3)
    kernels = []
    for name, recs in per_kernel.items():
        ms = float(np.mean([r["ms"] for r in recs]))
        cs = recs[0]["cells"] * w.S
        kernels.append({"name": name, "pairs": recs[0]["n_pairs"], "cell_states": cs, "ms": ms,
                        "achieved_GBs": BYTES_PER_CELL_STATE * cs / (ms / 1e3) / 1e9})
    kernels.sort(key=lambda k: -k["ms"])
    dom = kernels[0]
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
                   "emission_counts": not args.no_emit,
                   "sum_logp": float(logp.sum()), "sum_counts": float(n.sum())},
        "roofline": {"bound": "hbm", "achieved": dom["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": TRAFFIC.get(dom["name"]),
                     "kernel": dom["name"], "kernel_ms": dom["ms"],
                     "bytes_per_cell_state": BYTES_PER_CELL_STATE, "kernels": kernels},
    }
    if not args.no_plain:
        _log("timing the per-pair drop-in")
        pl = plain_per_pair(w, emit=not args.no_emit)
        out["plain_per_pair"] = pl
        out["batched_over_plain"] = value / pl["value"]
    loads = [float(v) for v in args.serve_loads.split(",") if v.strip()]
    if loads:
        _log("serving measurement")
        out["serving"] = serve_bench(loads, args.serve_requests)
    if not args.no_cpu:
        _log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(w, seconds=args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
