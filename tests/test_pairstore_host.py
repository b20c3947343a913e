"""Host codec of the X/A/Y precompiled-pair format (tkfmi.pairstore,
reference python/tkfmixdom/jax/util/pair_format.py) on the golden shard
tests/golden/pairs_shard.jsonl: the reference's own doctest examples
(pair_format.py:47-52, 78-80), encode / decode round trips, the run regex's
skipping of other text, and validate_record's four failure kinds."""
import json
import os

import numpy as np

from tkfmi import pairstore

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _records(name):
    with open(os.path.join(GOLD, name)) as f:
        return [json.loads(line) for line in f if line.strip()]


def test_reference_doctest_vectors():
    assert pairstore.rle_encode([1, 1, 1, 2, 1, 3, 3]) == "M3I1M1D2"
    assert pairstore.rle_encode([1]) == "M1"
    assert pairstore.rle_decode("M3I1M1D2") == [1, 1, 1, 2, 1, 3, 3]
    assert pairstore.rle_encode([]) == "" and pairstore.rle_decode("") == []
    assert pairstore.rle_decode("M2 D1x M1*D1") == [1, 1, 3, 1, 3]
    assert pairstore.rle_decode("M12I3") == [1] * 12 + [2] * 3


def test_shard_round_trip_and_validation():
    recs = _records("pairs_shard.jsonl")
    assert len(recs) == 53
    for r in recs:
        assert pairstore.validate_record(r) == (True, "")
        x, y, st, anc, desc, t, _, _ = pairstore.decode_pair(r)
        st = np.asarray(st)
        assert len(x) == (st != 2).sum() and len(y) == (st != 3).sum()
        assert anc == list(x) and desc == list(y)
        if "junk" not in r["a"] and " " not in r["a"]:
            again = pairstore.encode_pair(x, y, list(st), t, r["fam"])
            assert (again["x"], again["a"], again["y"], again["t"]) == (r["x"], r["a"], r["y"], r["t"])
    assert pairstore.read_jsonl(os.path.join(GOLD, "pairs_shard.jsonl")) == recs


def test_invalid_records_fail_validation():
    want = {"bad_x": "Invalid AA in X", "bad_y": "Invalid AA in Y", "len_x": "M+D=", "len_y": "M+I="}
    for r in _records("pairs_invalid.jsonl"):
        ok, why = pairstore.validate_record(r)
        assert not ok and why.startswith(want[r["id"]]), (r["id"], why)
