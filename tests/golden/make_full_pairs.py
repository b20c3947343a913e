"""Writes tests/golden/full_pairs.npz: pairs at the benchmark configurations'
full sizes with reference outputs from the oracle in 80-bit long double
(oracle/c/fb2d_scaled.c built with -DREAL_LONG_DOUBLE; at these lengths the
FP64 log-space recursion itself drifts by ~1e-10, so the long-double run is
the value the GPU is held to).

Pairs (name: states, lengths):
  c1max_a, c1max_b  MixDom2 K=3 F=1 C=1 (17 states), ancestors of 800 and
                    790 residues (the 1024-pair configuration's clip limit)
  c2_300, c2_1000   MixDom2 K=8 F=4 C=4 (162 states), ancestors of 300 and
                    1000 residues (the 256-pair configuration's range)
  best_d3f2         the reference's trained bw_d3f2_fullseed_15iter set
                    (K=3 F=2, 32 states, per-domain emissions), 700 residues
  best_d8f2         bw_d8f2_fullseed_15iter (K=8 F=2, 82 states), 620 residues
Descendants are simulated from each pair's own model (tkfmi.synth).  Stored
per pair: the inputs the kernels take (codes, log transitions, per-group log
emission tables, state types and groups) and the outputs: log P, transition
counts (S,S), emission counts (G,483) and the posteriors at 200 random cells.

Run once from the repository root (~2 minutes on 8 cores; reads the
reference's parameter files through tests/golden/params_best.npz):
    python tests/golden/make_full_pairs.py
"""
import os
import sys

import numpy as np
import scipy.linalg

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tkf-mixdom_amd")]

from oracle import cfb2d, emissions, mixdom as omix, tkf  # noqa: E402
from tkfmi import synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
N_SAMPLES = 200


def _samples(rng, Lx, Ly):
    i = rng.integers(0, Lx + 1, N_SAMPLES)
    j = rng.integers(0, Ly + 1, N_SAMPLES)
    i[:2], j[:2] = (0, Lx), (0, Ly)          # the origin and the corner
    return np.stack([i, j], axis=1).astype(np.int32)


def _from_workload(w, b):
    tab = w.log_tab[b].numpy() if w.log_tab.dim() == 3 else w.log_tab.numpy()
    return dict(x=np.asarray(w.xs[b]), y=np.asarray(w.ys[b]), log_trans=w.log_trans[b].numpy(),
                log_tab=tab, state_types=np.asarray(w.state_types, np.int32),
                state_group=np.asarray(w.state_group, np.int32))


def _from_params_best(name, Lx, seed):
    """A pair under one of the reference's trained parameter sets, per-domain
    emissions (the 2D E-step's model, train_pfam.py:1613-1741), t = t_rep."""
    d = np.load(os.path.join(HERE, "params_best.npz"), allow_pickle=False)
    g = lambda k: d[f"{name}/{k}"]  # noqa: E731
    K, F = g("ext_rates").shape
    t = float(g("t_rep"))
    chi = omix.build_nested_trans(float(g("main_ins")), float(g("main_del")), t, g("dom_ins"),
                                  g("dom_del"), g("dom_weights"), g("frag_weights"), g("ext_rates"))
    st = np.asarray(omix.state_types(K, F), np.int32)
    dom = np.array([0, 0] + [(s - 2) // (5 * F) for s in range(2, len(st))])
    Qs, pis = g("dom_Qs"), g("dom_pis")
    subs = np.stack([scipy.linalg.expm(Qs[k] * t) for k in range(K)])
    rng = np.random.default_rng(seed)
    wx = pis.mean(axis=0)
    x = rng.choice(20, size=Lx, p=wx / wx.sum())
    y = synth.sample_descendant(rng, x, chi, st, lambda s, a: pis[dom[s], a] * subs[dom[s], a],
                                lambda s: pis[dom[s]], wx[x] / wx.sum())
    log_subs, log_pis = emissions.log_tables_per_domain(subs, pis)
    tab = np.concatenate([(log_pis[:, :, None] + log_subs).reshape(K, -1), log_pis, log_pis], axis=1)
    return dict(x=x, y=np.asarray(y), log_trans=tkf.floor_log(chi), log_tab=tab, state_types=st,
                state_group=dom.astype(np.int32))


def main():
    pairs = {}
    w1 = synth.make_mixdom2(seed=101, B=2, K=3, F=1, C=1, lengths=[800, 790])
    pairs["c1max_a"], pairs["c1max_b"] = _from_workload(w1, 0), _from_workload(w1, 1)
    w2 = synth.make_mixdom2(seed=102, B=2, K=8, F=4, C=4, lengths=[300, 1000])
    pairs["c2_300"], pairs["c2_1000"] = _from_workload(w2, 0), _from_workload(w2, 1)
    pairs["best_d3f2"] = _from_params_best("bw_d3f2_fullseed_15iter", 700, 103)
    pairs["best_d8f2"] = _from_params_best("bw_d8f2_fullseed_15iter", 620, 104)
    out = {}
    rng = np.random.default_rng(7)
    for name, p in pairs.items():
        Lx, Ly = len(p["x"]), len(p["y"])
        samp = _samples(rng, Lx, Ly)
        lp, n, ec, post = cfb2d.fb2d_scaled(p["log_trans"], p["state_types"], p["state_group"],
                                            p["log_tab"], p["x"], p["y"], samp, long_double=True,
                                            nthreads=8)
        for k, v in p.items():
            out[f"{name}/{k}"] = v.astype(np.uint8) if k in ("x", "y") else v
        out[f"{name}/logp"] = np.float64(lp)
        out[f"{name}/n"] = n
        out[f"{name}/ecnt"] = ec
        out[f"{name}/samples"] = samp
        out[f"{name}/post"] = post
        print(f"{name}: S={len(p['state_types'])} Lx={Lx} Ly={Ly} logP={lp:.10f}", flush=True)
    np.savez_compressed(os.path.join(HERE, "full_pairs.npz"), **out)


if __name__ == "__main__":
    main()
