"""Dense log-emission input (the reference's log_emit_table,
dp/hmm.py:1834-1873) on the generic kernel (<= 15 emitting states) and the
16..160-state kernel, against the FP64 oracle fed the same dense lattice:
log P, transition counts and posteriors at 1e-10 relative (1e-12 x the
largest entry as a floor).

The tables are the model's own emissions perturbed by N(0, 0.3) per entry
(so that they are no longer table-structured) with about 5% of the emitting
entries set to NEG_INF = -1e30 (impossible emissions), as a caller's
rate-annotated or soft-observation table would be.  The reference surface
forward_backward_2d is checked with sequences padded beyond their real
lengths (real_Lx / real_Ly) and garbage in the table beyond the real
region, which the reference masks to NEG_INF."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
RTOL = 1e-10
NEG_INF = -1e30


def _close(a, b, rtol=RTOL, floor=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b)
    return bool(np.all(err <= rtol * np.abs(b) + floor * scale)), float((err / (np.abs(b) + floor * scale)).max())


def _dense_tables(w, rng, holes=0.05):
    from oracle import emissions
    st = np.asarray(w.state_types)
    emitting = (st == 1) | (st == 2) | (st == 3)
    out = []
    for b in range(w.B):
        tab = w.log_tab[b].numpy() if w.log_tab.dim() == 3 else w.log_tab.numpy()
        e = emissions.emit_from_tables(st, w.state_group, tab, w.xs[b], w.ys[b])
        e = e + rng.normal(0.0, 0.3, e.shape)
        hole = (rng.random(e.shape) < holes) & emitting[None, None, :]
        e[hole] = NEG_INF
        out.append(e)
    return out


CASES = [pytest.param(1, 1, 1, id="K1F1-S7"), pytest.param(3, 1, 1, id="K3F1-S17"),
         pytest.param(3, 2, 1, id="K3F2-S32"), pytest.param(8, 2, 2, id="K8F2-S82")]


@pytest.mark.parametrize("K,F,C", CASES)
def test_dense_emissions_match_oracle(K, F, C):
    from oracle import dp2d
    from tkfmi import hmm, synth
    w = synth.make_mixdom2(seed=70 + K + F, B=5, K=K, F=F, C=C, lengths=[0, 9, 47, 120, 33])
    w.ys[2] = w.ys[2][:0]                                   # an empty descendant
    rng = np.random.default_rng(5)
    dense = _dense_tables(w, rng)
    res = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, None, w.state_group,
                         want_posteriors=True, log_emit=[torch.as_tensor(e) for e in dense])
    from tkfmi import _abi
    _abi.check_status(res)
    st = np.asarray(w.state_types)
    for b in range(w.B):
        lp, post, n, _ = dp2d.forward_backward(w.log_trans[b].numpy(), st, dense[b])
        got = float(res["log_p"][b])
        assert abs(got - lp) <= RTOL * abs(lp), (b, got, lp)
        ok, worst = _close(res["n"][b].cpu().numpy(), n)
        assert ok, ("n", b, worst)
        ok, worst = _close(res["posteriors"][b].cpu().numpy(), post)
        assert ok, ("posteriors", b, worst)
    # forward-only gives the same log P
    fo = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, None, w.state_group,
                        forward_only=True, log_emit=[torch.as_tensor(e) for e in dense])
    assert np.array_equal(fo["log_p"].cpu().numpy(), res["log_p"].cpu().numpy())


@pytest.mark.parametrize("K,F", [(1, 1), (3, 2)])
def test_forward_backward_2d_dense_table_with_real_lengths(K, F):
    """The reference surface with padded sequences: real_Lx / real_Ly cut
    the table, entries beyond them (here +5, which would change the result
    if used) are ignored, posteriors come back at the _pad_to_bin shape and
    are zero outside the real region."""
    from oracle import dp2d
    from tkfmi import hmm, synth
    w = synth.make_mixdom2(seed=90 + K, B=1, K=K, F=F, C=1, lengths=[37])
    rng = np.random.default_rng(9)
    e = _dense_tables(w, rng)[0]
    Lx, Ly = len(w.xs[0]), len(w.ys[0])
    pad_x, pad_y = Lx + 6, Ly + 3
    x_pad = np.concatenate([w.xs[0], np.zeros(pad_x - Lx, np.int64)])
    y_pad = np.concatenate([w.ys[0], np.zeros(pad_y - Ly, np.int64)])
    big = np.full((pad_x + 1, pad_y + 1, w.S), 5.0)
    big[:Lx + 1, :Ly + 1] = e
    lt = w.log_trans[0]
    lp, post, n = hmm.forward_backward_2d(lt, w.state_types, x_pad, y_pad, None, None,
                                          log_emit_table=big, real_Lx=Lx, real_Ly=Ly)
    lp_o, post_o, n_o, _ = dp2d.forward_backward(lt.numpy(), np.asarray(w.state_types), e)
    assert abs(float(lp) - lp_o) <= RTOL * abs(lp_o)
    ok, worst = _close(n.cpu().numpy(), n_o)
    assert ok, worst
    post = post.cpu().numpy()
    assert post.shape == (hmm._pad_to_bin(pad_x) + 1, hmm._pad_to_bin(pad_y) + 1, w.S)
    ok, worst = _close(post[:Lx + 1, :Ly + 1], post_o)
    assert ok, worst
    assert not post[Lx + 1:].any() and not post[:, Ly + 1:].any()
    lp_f = hmm.forward_backward_2d(lt, w.state_types, x_pad, y_pad, None, None,
                                   log_emit_table=big, real_Lx=Lx, real_Ly=Ly, forward_only=True)
    assert float(lp_f) == float(lp)


def test_dense_refuses_emission_counts():
    from tkfmi import _abi, hmm, synth
    w = synth.make_mixdom2(seed=3, B=1, lengths=[12])
    e = _dense_tables(w, np.random.default_rng(1))
    with pytest.raises(ValueError):
        hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, None, w.state_group,
                       want_emit_counts=True, log_emit=[torch.as_tensor(x) for x in e])
