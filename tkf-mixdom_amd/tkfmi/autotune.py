"""Autotuner for the chained 2D forward/backward kernels.

The chained kernels split a batch into length classes (the longest pairs,
the other pairs of several bands, 2-band and 1-band pairs) and run each
class as chains of workgroups.  How many waves a workgroup of the two long
classes has trades per-pair parallelism against occupancy (a 4-wave backward
workgroup needs ~92 KB of LDS, so one fits a compute unit; a 3-wave one needs
~76 KB and a 2-wave one ~58 KB, so two fit).  The best choice depends on the batch's length mix, so :func:`tune`
(libtkfmi's tkfmi_autotune) times a few settings on the caller's own batch
and keeps the fastest for that batch's key -- its state layout and the
length bin of its longest sequence -- which later calls with the same key
use.  :func:`tune_chain` is the process-wide variant through
``tkfmi_chain_config``.
"""
import time

import torch

from . import _abi

# (long_w, main_w, long_frac, long_one): waves per workgroup of the longest
# pairs' class, of the other pairs' class, the cell fraction of the batch's
# largest pair from which a pair counts as long (0: no long class), and 1 to
# run that long class on the one-launch schedule (forward and backward of a
# pair together, see set_chain_schedule).
DEFAULT = (3, 3, 0.7, 1)  # the library's own defaults
CANDIDATES = ((3, 3, 0.7, 1), (3, 3, 0.5, 1), (4, 3, 0.6, 1), (3, 3, 0.0, 0), (3, 3, 0.3, 0), (4, 3, 0.3, 0),
              (4, 4, 0.3, 0), (3, 3, 0.15, 0), (4, 4, 0.0, 0), (3, 2, 0.0, 0), (2, 2, 0.0, 0))


def tune(runner, reps=3):
    """libtkfmi's own autotuner (tkfmi_autotune) on a BatchRunner's batch:
    times the candidate configurations of the chained kernels on that batch
    and keeps the fastest for the batch's key (state layout, length bin of
    its longest sequence), which later calls with that key then use.
    Resizes the runner's workspace for the chosen configuration.  Returns
    {"key": (layout, length_bin), "best": (long_w, main_w, long_frac, long_one),
    "ms": {config: ms per call}} (empty "ms" when the layout has a single
    configuration)."""
    L = _abi.lib()
    flags = runner.flags & ~(_abi.TIME_KERNELS | _abi.REUSE_DESCRIPTORS)
    wsb = int(L.tkfmi_autotune_workspace_bytes(_abi.ctypes.byref(runner.prob), flags))
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=runner.dev)
    res = _abi.TuneResult()
    stream = torch.cuda.current_stream(runner.dev)
    with torch.cuda.stream(stream):
        runner.status.zero_()
    _abi.check(L.tkfmi_autotune(_abi.ctypes.byref(runner.prob), _abi.ctypes.byref(runner.out),
                                _abi._ptr(ws), _abi.ctypes.c_size_t(wsb),
                                _abi.ctypes.c_void_p(stream.cuda_stream), flags, int(reps),
                                _abi.ctypes.byref(res)))
    runner.resize_workspace()
    cfgs = [(res.long_w[c], res.main_w[c], res.long_frac[c], res.long_one[c]) for c in range(res.n)]
    return {"key": (res.layout, res.length_bin),
            "best": cfgs[res.best] if res.n else None,
            "ms": {c: res.ms[i] for i, c in enumerate(cfgs)}}


def set_chain_config(long_w, main_w, long_frac, long_one=0):
    """Set the library's chained-kernel configuration (process-wide)."""
    L = _abi.lib()
    _abi.check(L.tkfmi_chain_config(int(long_w), int(main_w), float(long_frac), int(long_one)))


def _time_runner(runner, steps):
    runner.run()
    torch.cuda.synchronize(runner.dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        runner.run()
    torch.cuda.synchronize(runner.dev)
    return 1e3 * (time.perf_counter() - t0) / steps


def tune_chain(make_runner, candidates=CANDIDATES, steps=3, margin=0.03):
    """Time each configuration on the caller's batch and keep the fastest.

    make_runner: callable returning a fresh :class:`tkfmi._abi.BatchRunner`
    for the batch (a runner's workspace is sized for the configuration in
    force when it is built).  The first candidate (the library default) is
    kept unless another is faster by more than ``margin`` (a fraction), so
    that timing noise does not flip the choice.  Returns {"best": (long_w,
    main_w, long_frac, long_one), "ms": {config: ms per call}}; the library is left set
    to the best.
    """
    timings = {}
    for cfg in candidates:
        set_chain_config(*cfg)
        r = make_runner()
        timings[cfg] = _time_runner(r, steps)
        del r
    best = candidates[0]
    fastest = min(timings, key=timings.get)
    if timings[fastest] < (1.0 - margin) * timings[best]:
        best = fastest
    set_chain_config(*best)
    return {"best": best, "ms": timings}


# The chained kernels' one-launch schedule: where a pair's forward and
# backward meet (fraction of Lx + Ly; 0: never one launch) and the largest
# call (in workgroups) that uses it (-1: what the GPU holds at once).
DEFAULT_SCHEDULE = (0.45, -1)  # the library's own defaults


def set_chain_schedule(cut, max_workgroups=-1):
    """Set the chained kernels' schedule (process-wide): calls of at most
    ``max_workgroups`` workgroups (-1: twice the device's compute units) run
    each pair's forward and backward in one launch, meeting at antidiagonal
    round(cut (Lx + Ly)); larger calls, and every call with cut = 0, run the
    forward launch, then the backward launch."""
    L = _abi.lib()
    _abi.check(L.tkfmi_chain_schedule(float(cut), int(max_workgroups)))
