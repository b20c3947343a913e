"""Parity at the benchmark configurations' full sizes, against long-double
reference values (tests/golden/full_pairs.npz, made by
tests/golden/make_full_pairs.py with the oracle in 80-bit long double):
log P, transition counts, emission counts and posteriors at 200 sampled
cells, each at 1e-10 relative with a floor of 1e-12 x the largest entry of
its kind, on every kernel that accepts the layout:

  c1max_a/b  (17 states, 800 x 793 and 790 x 858): chained, one workgroup
             per pair, generic
  c2_300, c2_1000 (162 states, up to 1000 x 1055), best_d3f2 (32 states,
             the reference's trained bw_d3f2_fullseed set, 700 x 695),
             best_d8f2 (82 states, bw_d8f2_fullseed, 620 x 632): the
             16..160-state kernels -- banded (chained 16-row bands, full
             lattice), one workgroup per pair sweeping antidiagonals, and
             the latter with antidiagonal checkpointing.  Rows this long
             also take the antidiagonal kernel's 64-cell residue-code reload
             path (antidiagonals longer than 512 cells).
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
RTOL = 1e-10
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "full_pairs.npz")


def _load(name):
    d = np.load(GOLDEN, allow_pickle=False)
    return {k.split("/", 1)[1]: d[k] for k in d.files if k.startswith(name + "/")}


def _close(a, b, rtol=RTOL, floor=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b) / (np.abs(b) + floor * scale)
    return bool(err.max() <= rtol), float(err.max())


def _check(name, kernel=None, checkpoint=False):
    from tkfmi import _abi, hmm
    p = _load(name)
    G = int(p["log_tab"].shape[0])
    try:
        _abi.set_checkpoint_budget(0)  # full lattices unless checkpoint=True
        res = hmm.fb2d_batch(torch.as_tensor(p["log_trans"]), p["state_types"], [p["x"]], [p["y"]],
                             torch.as_tensor(p["log_tab"]), p["state_group"], want_posteriors=True,
                             want_emit_counts=True, kernel=kernel, checkpoint=checkpoint)
    finally:
        _abi.set_checkpoint_budget(_abi.CHECKPOINT_BUDGET_DEFAULT)
    _abi.check_status(res)
    lp = float(res["log_p"][0])
    print(f"{name} kernel={kernel} ckpt={checkpoint}: logP {lp!r} vs {float(p['logp'])!r}", flush=True)
    assert abs(lp - float(p["logp"])) <= RTOL * abs(float(p["logp"])), (lp, float(p["logp"]))
    post = res["posteriors"][0]
    s = torch.as_tensor(p["samples"], dtype=torch.long, device=post.device)
    worst = {}
    for what, got, want in (("n", res["n"][0].cpu().numpy(), p["n"]),
                            ("emission counts", res["emit_counts"][0].cpu().numpy().reshape(G, -1), p["ecnt"]),
                            ("posteriors", post[s[:, 0], s[:, 1]].cpu().numpy(), p["post"])):
        ok, worst[what] = _close(got, want)
        assert ok, (what, worst[what])
    print(f"   relative error: logP {abs(lp / float(p['logp']) - 1.0):.1e}, "
          + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()), flush=True)


@pytest.mark.parametrize("kernel", ["chain", "class", "generic"])
@pytest.mark.parametrize("name", ["c1max_a", "c1max_b"])
def test_c1_max_pairs_against_long_double(name, kernel):
    _check(name, kernel=kernel)


@pytest.mark.parametrize("mode", ["band", "antidiagonal", "checkpoint"])
@pytest.mark.parametrize("name", ["c2_300", "c2_1000", "best_d3f2", "best_d8f2"])
def test_wide_pairs_against_long_double(name, mode):
    _check(name, kernel="class" if mode == "antidiagonal" else None, checkpoint=mode == "checkpoint")
