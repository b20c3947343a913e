"""Writes tests/golden/params_best.npz: the model parameters of the six
trained parameter sets the reference ships (python/params/best/*.npz), as
data for parity tests on realistic models.

Only the numeric model arrays are kept (main_ins, main_del, dom_ins,
dom_del, dom_weights, frag_weights, ext_rates (K, F: MixDom1 rates, put on
the diagonal of (K, F, F) by models/mixdom.py:284-287), dom_Qs, dom_pis,
dom_S_exch, t_rep), each under the key "<set>/<name>".  The files are read
with numpy.load(allow_pickle=False): nothing in them is executed.

Run once in the build container (the reference is not present where the
tests run):  python tests/golden/make_params_best.py
"""
import glob
import os

import numpy as np

SRC = "/root/reference/python/params/best"
KEYS = ("main_ins", "main_del", "dom_ins", "dom_del", "dom_weights", "frag_weights",
        "ext_rates", "dom_Qs", "dom_pis", "dom_S_exch", "t_rep")


def main():
    out = {}
    for f in sorted(glob.glob(os.path.join(SRC, "*.npz"))):
        name = os.path.basename(f)[:-4]
        with np.load(f, allow_pickle=False) as d:
            for k in KEYS:
                out[f"{name}/{k}"] = np.asarray(d[k], dtype=np.float64)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params_best.npz")
    np.savez_compressed(dst, **out)
    print(dst, len(out), "arrays")


if __name__ == "__main__":
    main()
