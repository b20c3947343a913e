"""GPU parity: libtkfmi batched 1D alignment-constrained forward/backward
(tkfmi_fb1d_batched) vs the FP64 oracle (oracle/fb1d.py).

Tolerance: 1e-10 relative (logP, transition counts, posteriors, emission
counts); entries tiny compared with the largest of their kind get an
absolute floor of 1e-12 x that maximum.
"""
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-10
A = 20


def _close(a, b, rtol=RTOL, floor=1e-12):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if b.size == 0:
        return True, 0.0
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b)
    ok = err <= rtol * np.abs(b) + floor * scale
    return bool(ok.all()), float((err / (np.abs(b) + floor * scale)).max())


def _model(name):
    from oracle import mixdom, tkf
    if name == "tkf92":
        return tkf.tkf92_trans(0.04, 0.05, 0.7, 0.6), list(tkf.TKF_STATE_TYPES), 1, 1, 0, 4
    K, F = {"k3f1": (3, 1), "k3f3": (3, 3), "k2f2": (2, 2)}[name]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    dw = rng.dirichlet(np.ones(K))
    fw = rng.dirichlet(np.ones(F), size=K)
    ext = rng.uniform(0.5, 0.95, (K, F))
    trans = mixdom.build_nested_trans(0.04, 0.05, 0.8, [0.03] * K, [0.05] * K, dw, fw, ext)
    return trans, list(mixdom.state_types(K, F)), K, F, 0, 1


def _groups(st, F):
    if len(st) == 5:
        return [0] * 5
    return [0, 0] + [(s - 2) // (5 * F) for s in range(2, len(st))]


def _tables(rng, G):
    from oracle import emissions
    tabs = []
    for _ in range(G):
        pi = rng.dirichlet(np.ones(A))
        P = rng.dirichlet(np.ones(A), size=A)
        ls, lp = emissions.log_tables_per_domain(P[None], pi[None])
        tabs.append(np.concatenate([(lp[0][:, None] + ls[0]).ravel(), lp[0], lp[0]]))
    return np.stack(tabs)


def _alignments(rng, lengths):
    cts, ancs, descs = [], [], []
    for L in lengths:
        cts.append(rng.integers(1, 4, L))
        ancs.append(rng.integers(0, A, L))
        descs.append(rng.integers(0, A, L))
    return cts, ancs, descs


def _random_stochastic(rng, S, s_idx, e_idx):
    """A dense random transition matrix over S states with a start and an
    end state (no transitions into the start, none out of the end)."""
    T = rng.dirichlet(np.ones(S) * 0.5, size=S)
    T[:, s_idx] = 0.0
    T[e_idx] = 0.0
    T /= np.maximum(T.sum(1, keepdims=True), 1e-300)
    T[e_idx] = 0.0
    return T


@pytest.mark.parametrize("model", ["tkf92", "k3f1", "k2f2", "k3f3"])
def test_constrained_matches_oracle(model):
    import torch
    from oracle import fb1d, tkf
    from tkfmi import hmm1d
    trans, st, K, F, s0, e0 = _model(model)
    rng = np.random.default_rng(zlib.crc32(f"c-{model}".encode()))
    g = _groups(st, F)
    G = max(g) + 1
    lengths = [0, 1, 2, 5, 17, 64, 65, 130, 400]
    cts, ancs, descs = _alignments(rng, lengths)
    B = len(lengths)
    tabs = np.stack([_tables(rng, G) for _ in range(B)])              # per-pair tables
    lt = tkf.safe_log(trans)
    print(f"[fb1d constrained] model={model} S={len(st)} lengths={lengths}", flush=True)
    res = hmm1d.fb1d_constrained_batch(torch.as_tensor(lt), st, g, torch.as_tensor(tabs), cts, ancs,
                                       descs, init_state=s0, final_state=e0, want_posteriors=True,
                                       want_emit_counts=True)
    lp = res["log_p"].cpu().numpy()
    n = res["n"].cpu().numpy()
    ec = res["emit_counts"].cpu().numpy()
    posts = [p.cpu().numpy() for p in res["posteriors"]]
    for b in range(B):
        le = fb1d.constrained_emissions(st, g, tabs[b], cts[b], ancs[b], descs[b])
        lpo, posto, no = fb1d.forward_backward_1d(lt, le, s0, e0)
        eco = fb1d.emission_counts_1d(posto, st, g, cts[b], ancs[b], descs[b], G)
        assert abs(lp[b] - lpo) <= RTOL * max(1.0, abs(lpo)), (b, lp[b], lpo)
        for name, got, want in (("n", n[b], no), ("post", posts[b], posto), ("ecnt", ec[b], eco)):
            ok, err = _close(got, want)
            assert ok, (model, b, name, err)


@pytest.mark.parametrize("S", [5, 33, 64, 65, 129, 162])
def test_dense_generic_states_match_oracle(S):
    """Dense log emissions over S states: one wave (S <= 64), two waves with
    T in LDS (65..128), three waves with T streamed from HBM (> 128)."""
    import torch
    from oracle import fb1d, tkf
    from tkfmi import hmm1d
    rng = np.random.default_rng(S)
    s0, e0 = 0, S - 1
    T = _random_stochastic(rng, S, s0, e0)
    lengths = [3, 40, 257]
    emits = []
    for L in lengths:
        e = np.log(rng.uniform(0.01, 1.0, (L, S)))
        e[rng.random((L, S)) < 0.2] = -1e30                            # sparse emission pattern
        keep = rng.integers(1, S - 1, L)                                # >= 1 emitting state per column
        e[np.arange(L), keep] = np.log(rng.uniform(0.01, 1.0, L))
        e[:, [s0, e0]] = -1e30
        emits.append(e)
    lt = tkf.safe_log(T)
    print(f"[fb1d dense] S={S} lengths={lengths}", flush=True)
    res = hmm1d.fb1d_batch(torch.as_tensor(lt), [torch.as_tensor(e) for e in emits],
                           init_state=s0, final_state=e0)
    for b, e in enumerate(emits):
        lpo, posto, no = fb1d.forward_backward_1d(lt, e, s0, e0)
        got = float(res["log_p"][b])
        assert abs(got - lpo) <= RTOL * max(1.0, abs(lpo)), (S, b, got, lpo)
        for name, g, w in (("n", res["n"][b].cpu().numpy(), no),
                           ("post", res["posteriors"][b].cpu().numpy(), posto)):
            ok, err = _close(g, w)
            assert ok, (S, b, name, err)


def test_forward_only_and_padded_wrapper():
    import torch
    from oracle import fb1d, tkf
    from tkfmi import hmm1d
    trans, st, K, F, s0, e0 = _model("k3f1")
    rng = np.random.default_rng(11)
    g = _groups(st, F)
    tab = _tables(rng, 3)
    ct, a, d = [x[0] for x in _alignments(rng, [90])]
    le = fb1d.constrained_emissions(st, g, tab, ct, a, d)
    lt = tkf.safe_log(trans)
    lpo, posto, no = fb1d.forward_backward_1d(lt, le, s0, e0)
    padded = np.vstack([le, np.full((38, len(st)), -1e30)])          # _pad_to_bin-style padding
    print("[fb1d padded wrapper] L=90 padded=128", flush=True)
    lp, post, n = hmm1d.forward_backward_1d_padded(torch.as_tensor(lt), torch.as_tensor(padded), 90,
                                                   init_state=s0, final_state=e0)
    assert abs(float(lp) - lpo) <= RTOL * abs(lpo)
    assert _close(n.cpu().numpy(), no)[0]
    assert _close(post[:90].cpu().numpy(), posto)[0] and float(post[90:].abs().max()) == 0.0
    lpf = hmm1d.forward_backward_1d_padded(torch.as_tensor(lt), torch.as_tensor(padded), 90,
                                           init_state=s0, final_state=e0, forward_only=True)
    assert abs(float(lpf) - lpo) <= RTOL * abs(lpo)
    lp0, post0, n0 = hmm1d.forward_backward_1d_padded(torch.as_tensor(lt), torch.as_tensor(padded), 0,
                                                      init_state=s0, final_state=e0)
    assert float(lp0) == pytest.approx(lt[s0, e0], rel=1e-15)
    assert float(n0[s0, e0]) == 1.0 and float(n0.sum()) == 1.0 and float(post0.abs().max()) == 0.0


def test_long_alignments_invariants():
    """Lengths up to 6000 columns: every column is emitted by exactly one
    state, entries per state type equal the column counts, and the start /
    end rows of n sum to one."""
    import torch
    from oracle import tkf
    from tkfmi import hmm1d
    trans, st, K, F, s0, e0 = _model("k3f1")
    rng = np.random.default_rng(12)
    g = _groups(st, F)
    tab = _tables(rng, 3)
    lengths = [6000, 2500, 1]
    cts, ancs, descs = _alignments(rng, lengths)
    print(f"[fb1d long] lengths={lengths}", flush=True)
    res = hmm1d.fb1d_constrained_batch(torch.as_tensor(tkf.safe_log(trans)), st, g, torch.as_tensor(tab),
                                       cts, ancs, descs, want_posteriors=True, want_emit_counts=True)
    stn = np.asarray(st)
    for b, L in enumerate(lengths):
        n = res["n"][b].cpu().numpy()
        post = res["posteriors"][b].cpu().numpy()
        assert np.isfinite(float(res["log_p"][b]))
        np.testing.assert_allclose(post.sum(1), 1.0, rtol=1e-10)
        assert n[s0].sum() == pytest.approx(1.0, rel=1e-10)
        assert n[:, e0].sum() == pytest.approx(1.0, rel=1e-10)
        for ty in (1, 2, 3):
            assert n[:, stn == ty].sum() == pytest.approx((cts[b] == ty).sum(), rel=1e-10)
        assert float(res["emit_counts"][b].sum()) == pytest.approx(L, rel=1e-10)


def test_impossible_pair_returns_neg_inf_and_zero_counts():
    import torch
    from oracle import tkf
    from tkfmi import hmm1d
    trans, st, K, F, s0, e0 = _model("tkf92")
    T = trans.copy()
    T[s0, 1:4] = 0.0                    # the start state can only end
    T[s0, e0] = 1.0
    rng = np.random.default_rng(13)
    cts, ancs, descs = _alignments(rng, [7, 0])
    print("[fb1d impossible pair]", flush=True)
    res = hmm1d.fb1d_constrained_batch(torch.as_tensor(tkf.safe_log(T)), st, [0] * 5,
                                       torch.as_tensor(_tables(rng, 1)), cts, ancs, descs,
                                       init_state=s0, final_state=e0, want_posteriors=True)
    assert float(res["log_p"][0]) == -1e30
    assert float(res["n"][0].abs().sum()) == 0.0 and float(res["emit_counts"][0].abs().sum()) == 0.0
    assert float(res["log_p"][1]) == pytest.approx(0.0, abs=1e-15)   # log T[S,E] = log 1
    assert float(res["n"][1][s0, e0]) == 1.0
