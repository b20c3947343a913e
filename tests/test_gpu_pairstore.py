"""The device-resident pair store (tkfmi_pairs_decode, tkfmi.pairstore.PairStore)
on the golden shard tests/golden/pairs_shard.jsonl: every pair decoded on
the device equals the host codec's decode_pair exactly (residue codes, the
M/I/D state of every column, ancestor and descendant characters); the 2D
E-step delta computed straight from the store equals the one from the
reference's tuple lists; the 1D alignment-constrained forward/backward on
the store's columns equals the one on host-packed columns; invalid records
are refused with their reason."""
import json
import os
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SHARD = os.path.join(GOLD, "pairs_shard.jsonl")


def _tuples(recs):
    from tkfmi import pairstore
    return [pairstore.decode_pair(r)[:6] for r in recs]


def test_device_decode_equals_host_codec():
    from tkfmi import pairstore
    store = pairstore.PairStore.from_jsonl(SHARD)
    recs = pairstore.read_jsonl(SHARD)
    assert store.B == len(recs) == 53
    assert not store.status.any()
    for b, (x, y, st, anc, desc, t) in enumerate(_tuples(recs)):
        gx, gy, gst, ganc, gdesc, gt = store.decoded(b)
        assert np.array_equal(gx, x) and np.array_equal(gy, y), b
        assert gst == st and ganc == anc and gdesc == desc, b
        assert gt == t and store.n_cols[b] == len(st)


def test_estep_2d_from_store_equals_tuples():
    from tkfmi import estep, mixdom, pairstore
    rng = np.random.default_rng(4)
    K, F = 3, 1
    Qs, pis = [], []
    for _ in range(K):
        S = rng.gamma(1.0, 1.0, (20, 20)); S = (S + S.T) / 2
        pi = rng.dirichlet(np.ones(20))
        Q = S * pi[None]; np.fill_diagonal(Q, 0.0); np.fill_diagonal(Q, -Q.sum(1))
        Qs.append(Q); pis.append(pi)
    Qs, pis = np.stack(Qs), np.stack(pis)
    cp = dict(main_ins=0.04, main_del=0.05, dom_ins=np.array([0.03, 0.045, 0.02]),
              dom_del=np.array([0.05, 0.06, 0.04]), dom_weights=np.array([0.3, 0.3, 0.4]),
              frag_weights=np.ones((3, 1)), ext_rates=np.array([[0.7], [0.85], [0.6]]))
    st = mixdom.state_types(K, F)
    store = pairstore.PairStore.from_jsonl(SHARD)
    d1 = estep.process_pairs_batched_2d(store, cp, st, Qs[0], pis[0], len(st), dom_Qs=Qs, dom_pis=pis,
                                        n_dom=K, n_frag=F)
    d2 = estep.process_pairs_batched_2d(_tuples(pairstore.read_jsonl(SHARD)), cp, st, Qs[0], pis[0],
                                        len(st), dom_Qs=Qs, dom_pis=pis, n_dom=K, n_frag=F)
    assert d1["n_pairs"] == d2["n_pairs"] == 53
    assert d1["total_ll"] == d2["total_ll"]
    for k in ("agg_n_chi", "dom_match_counts", "dom_insert_counts", "dom_delete_counts", "dom_W", "dom_U"):
        a, b = np.asarray(d1[k]), np.asarray(d2[k])
        assert np.allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(b).max()), k


def test_fb1d_on_store_columns_equals_packed_columns():
    from tkfmi import _abi, hmm, mixdom, pairstore, synth, tables
    w = synth.make_mixdom2(seed=12, B=1, K=2, F=2, C=1, lengths=[10])
    store = pairstore.PairStore.from_jsonl(SHARD)
    recs = pairstore.read_jsonl(SHARD)
    lt = w.log_trans[0].to(hmm.device())
    tab = w.log_tab[0].to(hmm.device())
    cols = store.columns()
    common = dict(init_state=0, final_state=1, state_types=w.state_types, state_group=w.state_group,
                  log_tab=tab, want_emit_counts=True)
    r1 = _abi.Runner1D(lt, columns=cols, **common)
    r1.run(zero_outputs=False)
    cts, ancs, descs = [], [], []
    for r in recs:
        x, y, st, anc, desc, _ = pairstore.decode_pair(r)[:6]
        st = np.asarray(st, np.int64)
        a = np.zeros(len(st), np.int64); d = np.zeros(len(st), np.int64)
        a[st != 2] = anc; d[st != 3] = desc
        cts.append(st); ancs.append(a); descs.append(d)
    r2 = _abi.Runner1D(lt, columns=_abi.PackedColumns(cts, ancs, descs, hmm.device()), **common)
    r2.run(zero_outputs=False)
    assert np.array_equal(r1.log_p.cpu().numpy(), r2.log_p.cpu().numpy())
    assert np.allclose(r1.n.cpu().numpy(), r2.n.cpu().numpy(), rtol=1e-13, atol=0)
    assert np.allclose(r1.ecnt.cpu().numpy(), r2.ecnt.cpu().numpy(), rtol=1e-13, atol=0)


@pytest.mark.parametrize("which,reason", [("bad_x", "invalid character in X"),
                                          ("bad_y", "invalid character in Y"),
                                          ("len_x", "M+D run lengths"), ("len_y", "M+I run lengths")])
def test_invalid_records_refused(which, reason):
    from tkfmi import pairstore
    with open(os.path.join(GOLD, "pairs_invalid.jsonl")) as f:
        rec = [json.loads(line) for line in f if json.loads(line)["id"] == which]
    good = pairstore.read_jsonl(SHARD)[:3]
    with pytest.raises(ValueError, match=re.escape(reason)):
        pairstore.PairStore(good + rec)
