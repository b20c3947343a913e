"""Pins the 1D alignment-constrained oracle (oracle/fb1d.py): brute-force
state-path enumeration, the alignment-sum identity against the 2D oracle
(sum over every alignment of the 1D probability = the 2D pair-HMM
probability, and likewise for transition and emission counts), finite
differences, column invariants and the empty closed form.  CPU only."""
import zlib

import numpy as np
import pytest

from oracle import dp2d, emissions, fb1d, mixdom, tkf

A = 20

MODELS = {
    "tkf91": lambda: (tkf.tkf91_trans(0.04, 0.05, 0.5), tkf.TKF_STATE_TYPES, 1),
    "tkf92": lambda: (tkf.tkf92_trans(0.04, 0.05, 0.7, 0.6), tkf.TKF_STATE_TYPES, 1),
    "mixdom_k2f2": lambda: (mixdom.build_nested_trans(
        0.04, 0.05, 0.5, [0.03, 0.05], [0.05, 0.06], [0.4, 0.6],
        [[0.3, 0.7], [0.5, 0.5]], [[0.8, 0.6], [0.7, 0.9]]), mixdom.state_types(2, 2), 2),
    "mixdom_k3f1": lambda: (mixdom.build_nested_trans(
        0.04, 0.05, 1.3, [0.04] * 3, [0.05] * 3, [0.2, 0.5, 0.3], [[1.0]] * 3,
        [[0.6], [0.8], [0.9]]), mixdom.state_types(3, 1), 1),
}


def _ends(st):
    st = np.asarray(st)
    return int(np.argmax(st == tkf.S)), int(np.argmax(st == tkf.E))


def _groups(st, F):
    """Per-domain table group of each state: (s - 2) // (5F) for MixDom
    body states, 0 for TKF and the SS / EE states."""
    st = np.asarray(st)
    if len(st) == 5:
        return np.zeros(5, dtype=np.int64)
    g = np.zeros(len(st), dtype=np.int64)
    g[2:] = np.arange(len(st) - 2) // (5 * F)
    return g


def _tables(rng, n_groups):
    tabs = []
    for _ in range(n_groups):
        pi = rng.dirichlet(np.ones(A))
        P = rng.dirichlet(np.ones(A), size=A)
        ls, lp = emissions.log_tables_per_domain(P[None], pi[None])
        tabs.append(np.concatenate([(lp[0][:, None] + ls[0]).ravel(), lp[0], lp[0]]))
    return np.stack(tabs)


def _random_alignment(rng, L):
    ct = rng.integers(1, 4, L)
    return ct, rng.integers(0, A, L), rng.integers(0, A, L)


@pytest.mark.parametrize("model", ["tkf91", "tkf92", "mixdom_k2f2", "mixdom_k3f1"])
@pytest.mark.parametrize("L", [0, 1, 2, 3])
def test_fb1d_matches_brute_force(model, L):
    rng = np.random.default_rng(zlib.crc32(f"1d-{model}-{L}".encode()))
    trans, st, F = MODELS[model]()
    s0, e0 = _ends(st)
    g = _groups(st, F)
    tab = _tables(rng, int(g.max()) + 1)
    ct, a, b = _random_alignment(rng, L)
    le = fb1d.constrained_emissions(st, g, tab, ct, a, b)
    lt = tkf.safe_log(trans)
    lp, post, n = fb1d.forward_backward_1d(lt, le, s0, e0)
    Pb, nb, postb = fb1d.brute_1d(trans, np.exp(le), s0, e0)
    assert abs(lp - np.log(Pb)) <= 1e-12 * max(1.0, abs(lp))
    np.testing.assert_allclose(n, nb, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(post, postb, rtol=1e-10, atol=1e-13)
    assert fb1d.forward_backward_1d(lt, le, s0, e0, forward_only=True) == pytest.approx(lp, rel=1e-14)


@pytest.mark.parametrize("model", ["tkf92", "mixdom_k2f2"])
@pytest.mark.parametrize("Lx,Ly", [(0, 0), (1, 0), (0, 2), (2, 1), (2, 3), (3, 3)])
def test_alignment_sum_equals_2d(model, Lx, Ly):
    """sum_aln P_1D(aln) = P_2D, sum_aln P_1D(aln) n_1D(aln) = P_2D n_2D and
    the same for emission counts: the 2D DP sums over every alignment path
    what the 1D DP computes for one alignment."""
    rng = np.random.default_rng(zlib.crc32(f"sum-{model}-{Lx}-{Ly}".encode()))
    trans, st, F = MODELS[model]()
    s0, e0 = _ends(st)
    g = _groups(st, F)
    G = int(g.max()) + 1
    tab = _tables(rng, G)
    x = rng.integers(0, A, Lx)
    y = rng.integers(0, A, Ly)
    lt = tkf.safe_log(trans)
    em2 = emissions.emit_from_tables(st, g, tab, x, y)
    lp2, post2, n2, _ = dp2d.forward_backward(lt, st, em2)
    ec2 = emissions.emission_counts(post2, st, g, x, y, G)
    tot, nacc, ecacc = 0.0, np.zeros_like(n2), np.zeros_like(ec2)
    for aln in fb1d.alignments(Lx, Ly):
        ct, a, b = fb1d.alignment_columns(aln, x, y)
        le = fb1d.constrained_emissions(st, g, tab, ct, a, b)
        lp, post, n = fb1d.forward_backward_1d(lt, le, s0, e0)
        p = np.exp(lp)
        tot += p
        nacc += p * n
        ecacc += p * fb1d.emission_counts_1d(post, st, g, ct, a, b, G)
    assert np.log(tot) == pytest.approx(lp2, rel=1e-12, abs=1e-12)
    np.testing.assert_allclose(nacc / tot, n2, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(ecacc / tot, ec2, rtol=1e-10, atol=1e-13)


def test_column_invariants_and_gradients():
    rng = np.random.default_rng(5)
    trans, st, F = MODELS["mixdom_k2f2"]()
    s0, e0 = _ends(st)
    g = _groups(st, F)
    tab = _tables(rng, 2)
    ct, a, b = _random_alignment(rng, 37)
    le = fb1d.constrained_emissions(st, g, tab, ct, a, b)
    lt = tkf.safe_log(trans)
    lp, post, n = fb1d.forward_backward_1d(lt, le, s0, e0)
    np.testing.assert_allclose(post.sum(1), 1.0, rtol=1e-12)       # one state per column
    assert n[s0].sum() == pytest.approx(1.0, rel=1e-12)
    assert n[:, e0].sum() == pytest.approx(1.0, rel=1e-12)
    st_a = np.asarray(st)
    for ty in (1, 2, 3):                                            # entries per type = columns
        assert n[:, st_a == ty].sum() == pytest.approx((ct == ty).sum(), rel=1e-12)
    # |logP| ~ 150: central differences with h = 1e-5 carry ~1e-9 roundoff
    h = 1e-5
    for (k, k2) in [(0, 2), (2, 2), (2, 7), (5, 1), (11, 13), (20, 21)]:
        if trans[k, k2] <= 0:
            continue
        d = np.zeros_like(lt)
        d[k, k2] = h
        fd = (fb1d.forward_backward_1d(lt + d, le, s0, e0, True)
              - fb1d.forward_backward_1d(lt - d, le, s0, e0, True)) / (2 * h)
        assert fd == pytest.approx(n[k, k2], rel=1e-6, abs=1e-8)
    for (t, k) in [(0, 2), (5, 3), (36, 4), (20, 12)]:
        if le[t, k] <= -1e29:
            continue
        d = np.zeros_like(le)
        d[t, k] = h
        fd = (fb1d.forward_backward_1d(lt, le + d, s0, e0, True)
              - fb1d.forward_backward_1d(lt, le - d, s0, e0, True)) / (2 * h)
        assert fd == pytest.approx(post[t, k], rel=1e-6, abs=1e-8)


def test_empty_alignment_closed_form():
    trans, st, _ = MODELS["mixdom_k3f1"]()
    s0, e0 = _ends(st)
    lp, post, n = fb1d.forward_backward_1d(tkf.safe_log(trans), np.zeros((0, len(st))), s0, e0)
    assert lp == pytest.approx(np.log(trans[s0, e0]), rel=1e-14)
    assert n[s0, e0] == 1.0 and n.sum() == 1.0 and post.shape == (0, len(st))


def test_columns_from_alignment_fills_like_reference():
    states = [1, 2, 3, 1, 2]
    ct, a, b = fb1d.columns_from_alignment(states, [4, 7, 9], [5, 6, 8, 10])
    assert list(a) == [4, 0, 7, 9, 0]
    assert list(b) == [5, 6, 0, 8, 10]
