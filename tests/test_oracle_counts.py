"""Pins the oracle's expected emission counts (oracle/emissions.py
emission_counts): they are the gradient of logP with respect to the log
emission tables (central differences), they satisfy exact identities with
the transition counts, and they agree with a brute-force tally over every
alignment path on tiny pairs.  CPU only."""
import itertools

import numpy as np
import pytest

from oracle import brute, dp2d, emissions, mixdom, tkf

A = 20
A_PAD = emissions.A_PAD
TAB_SZ = emissions.TAB_SZ


def _mix_model(rng, K=2, F=1):
    chi = mixdom.build_nested_trans(0.04, 0.05, 0.9, rng.uniform(0.03, 0.05, K),
                                    rng.uniform(0.05, 0.07, K), rng.dirichlet(np.ones(K)),
                                    rng.dirichlet(np.ones(F), size=K),
                                    rng.uniform(0.5, 0.9, (K, F)))
    st = mixdom.state_types(K, F)
    grp = [0, 0] + [(s - 2) // (5 * F) for s in range(2, len(st))]
    tab = rng.normal(-3.0, 1.0, (K, TAB_SZ))
    return chi, st, grp, tab


def test_emission_counts_are_table_gradient():
    """d logP / d log_tab[g, e] == emission_counts[g, e]."""
    rng = np.random.default_rng(11)
    chi, st, grp, tab = _mix_model(rng)
    x = np.array([3, 7, 3, 20, 1])
    y = np.array([7, 3, 9, 20])
    lt = tkf.safe_log(chi)
    em = emissions.emit_from_tables(st, grp, tab, x, y)
    _, post, _, _ = dp2d.forward_backward(lt, st, em)
    ec = emissions.emission_counts(post, st, grp, x, y, 2)
    h = 1e-6
    entries = [(0, 3 * A_PAD + 7), (1, 3 * A_PAD + 3), (0, 20 * A_PAD + 20), (1, 7 * A_PAD + 9),
               (0, A_PAD * A_PAD + 9), (1, A_PAD * A_PAD + 20), (0, A_PAD * A_PAD + A_PAD + 3),
               (1, A_PAD * A_PAD + A_PAD + 1), (0, 5 * A_PAD + 5)]
    for g, e in entries:
        tp, tm = tab.copy(), tab.copy()
        tp[g, e] += h
        tm[g, e] -= h
        lp_p, _ = dp2d.forward(lt, st, emissions.emit_from_tables(st, grp, tp, x, y))
        lp_m, _ = dp2d.forward(lt, st, emissions.emit_from_tables(st, grp, tm, x, y))
        assert (lp_p - lp_m) / (2 * h) == pytest.approx(ec[g, e], rel=1e-6, abs=1e-8), (g, e)
    assert ec[0, 5 * A_PAD + 5] == 0.0  # (5, 5) never aligned


def test_emission_count_identities():
    """Totals per type equal the expected visits in the transition counts:
    sum(match) == sum_k n[:, M-type k], and likewise for I and D; and every
    ancestor residue is emitted exactly once (M or D)."""
    rng = np.random.default_rng(12)
    chi, st, grp, tab = _mix_model(rng, K=3)
    x, y = rng.integers(0, A, 19), rng.integers(0, A, 23)
    lt = tkf.safe_log(chi)
    em = emissions.emit_from_tables(st, grp, tab, x, y)
    _, post, n, _ = dp2d.forward_backward(lt, st, em)
    ec = emissions.emission_counts(post, st, grp, x, y, 3)
    st = np.asarray(st)
    nm = A_PAD * A_PAD
    assert ec[:, :nm].sum() == pytest.approx(n[:, st == 1].sum(), rel=1e-11)
    assert ec[:, nm:nm + A_PAD].sum() == pytest.approx(n[:, st == 2].sum(), rel=1e-11)
    assert ec[:, nm + A_PAD:].sum() == pytest.approx(n[:, st == 3].sum(), rel=1e-11)
    # per ancestor residue a: M-row sums + D counts == occurrences of a in x
    match = ec[:, :nm].reshape(3, A_PAD, A_PAD).sum(axis=(0, 2))
    dele = ec[:, nm + A_PAD:].sum(axis=0)
    np.testing.assert_allclose(match + dele, np.bincount(x, minlength=A_PAD), rtol=1e-10, atol=1e-12)
    ins = ec[:, nm:nm + A_PAD].sum(axis=0)
    mcol = ec[:, :nm].reshape(3, A_PAD, A_PAD).sum(axis=(0, 1))
    np.testing.assert_allclose(mcol + ins, np.bincount(y, minlength=A_PAD), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("Lx,Ly", [(2, 2), (3, 1), (1, 3), (0, 2), (2, 0)])
def test_emission_counts_match_brute_force(Lx, Ly):
    """Brute force: enumerate every state path, tally the table entries each
    emission uses, weight by the path probability."""
    rng = np.random.default_rng(100 + 10 * Lx + Ly)
    chi, st, grp, tab = _mix_model(rng)
    x, y = rng.integers(0, A, Lx), rng.integers(0, A, Ly)
    lt = tkf.safe_log(chi)
    em = emissions.emit_from_tables(st, grp, tab, x, y)
    _, post, _, _ = dp2d.forward_backward(lt, st, em)
    ec = emissions.emission_counts(post, st, grp, x, y, 2)
    # brute: posterior per cell from path enumeration, then the entry each
    # (cell, state) emission uses, looked up independently of the oracle
    Pb, _, postb = brute.enumerate_paths(chi, st, np.exp(em), Lx, Ly)
    want = np.zeros_like(ec)
    for i, j, s in itertools.product(range(Lx + 1), range(Ly + 1), range(len(st))):
        w = postb[i, j, s]
        if w == 0.0:
            continue
        if st[s] == 1:
            want[grp[s], x[i - 1] * A_PAD + y[j - 1]] += w
        elif st[s] == 2:
            want[grp[s], A_PAD * A_PAD + y[j - 1]] += w
        elif st[s] == 3:
            want[grp[s], A_PAD * A_PAD + A_PAD + x[i - 1]] += w
    np.testing.assert_allclose(ec, want, rtol=1e-10, atol=1e-14)
