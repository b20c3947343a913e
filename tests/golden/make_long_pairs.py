"""Writes tests/golden/long_pairs.npz: long MixDom2 K=3 pairs (rows longer
than the kernel's LDS row buffer) with expected outputs from the oracle run
in 80-bit long double.  For pairs this long the float64 log-space recursion
(the reference's own numerics) drifts by ~1e-10 relative in the counts, so
the long-double run is the reference value the GPU is held to.

Run from the repository root:  python tests/golden/make_long_pairs.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tkf-mixdom_amd")]

from oracle import dp2d, emissions  # noqa: E402
from tkfmi import synth  # noqa: E402

POST_STRIDE = 37


def main():
    w = synth.make_mixdom2(seed=21, B=3, lengths=[40, 33, 25])
    rng = np.random.default_rng(2)
    w.ys[0] = rng.integers(0, 20, size=1200)
    w.ys[1] = rng.integers(0, 20, size=2000)
    w.ys[2] = w.ys[2][:20]
    p = w.params
    out = {"state_types": np.asarray(w.state_types, np.int32),
           "state_group": np.asarray(w.state_group, np.int32),
           "log_trans": w.log_trans.numpy(), "log_tab": w.log_tab.numpy(),
           "post_stride": np.int64(POST_STRIDE)}
    for b in range(w.B):
        x, y = np.asarray(w.xs[b]), np.asarray(w.ys[b])
        em = emissions.pair_hmm_emissions_per_class(np.asarray(w.state_types), x, y,
                                                    p["class_subs"][b], p["class_pi"],
                                                    p["class_dist"], p["K"], p["F"])
        lp, post, n, _ = dp2d.forward_backward(w.log_trans[b].numpy(), np.asarray(w.state_types),
                                               em, dtype=np.longdouble)
        out[f"x{b}"] = x.astype(np.uint8)
        out[f"y{b}"] = y.astype(np.uint8)
        out[f"logp{b}"] = np.float64(lp)
        out[f"n{b}"] = n
        out[f"post{b}"] = post[:, ::POST_STRIDE, :]
        print(f"pair {b}: Lx={len(x)} Ly={len(y)} logP={lp:.12f}", flush=True)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "long_pairs.npz"), **out)


if __name__ == "__main__":
    main()
