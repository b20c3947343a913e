"""Writes tests/golden/pairs_shard.jsonl and tests/golden/pairs_invalid.jsonl:
precompiled pair records in the reference's X/A/Y format
(python/tkfmixdom/jax/util/pair_format.py), one JSON object per line.

pairs_shard.jsonl: 48 pairs simulated under a MixDom2 model (tkfmi.synth),
each encoded from the alignment it was generated from, plus edge cases --
an empty pair, insert-only and delete-only pairs, a 2,400-column pair with
multi-digit runs, and an alignment string with text the reference's run
regex skips.  pairs_invalid.jsonl: one record per validation failure (an
invalid character in X and in Y, M+D != |X|, M+I != |Y|).

Run from the repository root:  python tests/golden/make_pair_shard.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tkf-mixdom_amd")]

from tkfmi import pairstore, synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    w = synth.make_mixdom2(seed=111, B=48, K=3, F=1, C=1, median=120, lo=5, hi=400)
    recs = []
    for b, (ct, anc, desc) in enumerate(zip(*synth.alignment_columns(w))):
        ct = np.asarray(ct)
        x = np.asarray(w.xs[b])
        y = np.asarray(w.ys[b])[:int(np.isin(ct, (1, 2)).sum())]
        recs.append(pairstore.encode_pair(x, y, ct.tolist(), w.ts[b], f"SYN{b % 5:05d}", f"r{b}", f"s{b}"))
    rng = np.random.default_rng(3)
    recs.append(pairstore.encode_pair([], [], [], 0.3, "EDGE", "empty", ""))
    recs.append(pairstore.encode_pair([], rng.integers(0, 20, 7), [2] * 7, 0.4, "EDGE", "ins", ""))
    recs.append(pairstore.encode_pair(rng.integers(0, 20, 9), [], [3] * 9, 0.5, "EDGE", "del", ""))
    st = [1] * 700 + [2] * 35 + [1] * 500 + [3] * 120 + [1] * 1000 + [2] * 45
    nx, ny = sum(s != 2 for s in st), sum(s != 3 for s in st)
    recs.append(pairstore.encode_pair(rng.integers(0, 20, nx), rng.integers(0, 20, ny), st, 1.25, "EDGE", "long", ""))
    junk = pairstore.encode_pair(rng.integers(0, 20, 5), rng.integers(0, 20, 3), [1, 1, 3, 1, 3], 0.7,
                                 "EDGE", "junk", "")
    junk["a"] = "M2 D1x M1*D1"   # the run regex sees M2 D1 M1 D1 = M M D M D
    recs.append(junk)
    for r in recs:
        ok, why = pairstore.validate_record(r)
        assert ok, (r, why)
    with open(os.path.join(HERE, "pairs_shard.jsonl"), "w") as f:
        for r in recs:
            f.write(json.dumps(r) + "\n")
    bad = [dict(recs[0], x="AC*DE" + recs[0]["x"][5:], id="bad_x"),
           dict(recs[1], y=recs[1]["y"][:-1] + "b", id="bad_y"),
           dict(recs[2], x=recs[2]["x"] + "A", id="len_x"),
           dict(recs[3], y=recs[3]["y"][1:], id="len_y")]
    with open(os.path.join(HERE, "pairs_invalid.jsonl"), "w") as f:
        for r in bad:
            f.write(json.dumps(r) + "\n")
    print(len(recs), "records,", len(bad), "invalid")


if __name__ == "__main__":
    main()
