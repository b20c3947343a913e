"""Pins the oracle DP (oracle/dp2d.py, oracle/c/fb2d_oracle.c) against
brute-force path enumeration, exact count identities, finite differences
and the empty-pair closed form.  CPU only."""
import zlib

import numpy as np
import pytest

from oracle import brute, cfb2d, dp2d, emissions, mixdom, tkf

A = 20


def _rand_sub(rng, A_=A):
    pi = rng.dirichlet(np.ones(A_))
    P = rng.dirichlet(np.ones(A_), size=A_)
    return P, pi


MODELS = {
    "tkf91": lambda: (tkf.tkf91_trans(0.04, 0.05, 0.5), tkf.TKF_STATE_TYPES),
    "tkf92": lambda: (tkf.tkf92_trans(0.04, 0.05, 0.7, 0.6), tkf.TKF_STATE_TYPES),
    "tkf91_cond": lambda: (tkf.tkf91_trans_cond(0.05, 0.05, 0.3), tkf.TKF_STATE_TYPES),
    "mixdom_k2f2": lambda: (mixdom.build_nested_trans(
        0.04, 0.05, 0.5, [0.03, 0.05], [0.05, 0.06], [0.4, 0.6],
        [[0.3, 0.7], [0.5, 0.5]], [[0.8, 0.6], [0.7, 0.9]]), mixdom.state_types(2, 2)),
    "mixdom_k3f1": lambda: (mixdom.build_nested_trans(
        0.04, 0.05, 1.3, [0.04] * 3, [0.05] * 3, [0.2, 0.5, 0.3], [[1.0]] * 3,
        [[0.6], [0.8], [0.9]]), mixdom.state_types(3, 1)),
}


@pytest.mark.parametrize("model", sorted(MODELS))
@pytest.mark.parametrize("Lx,Ly", [(0, 0), (1, 0), (0, 2), (2, 1), (2, 2), (3, 2)])
def test_dp_matches_brute_force(model, Lx, Ly):
    rng = np.random.default_rng(zlib.crc32(f"{model}-{Lx}-{Ly}".encode()))
    trans, st = MODELS[model]()
    P, pi = _rand_sub(rng)
    x = rng.integers(0, A, Lx)
    y = rng.integers(0, A, Ly)
    em = emissions.pair_hmm_emissions(st, x, y, P, pi)
    lp, post, n, lpb = dp2d.forward_backward(tkf.safe_log(trans), st, em)
    Pb, nb, postb = brute.enumerate_paths(trans, st, np.exp(em), Lx, Ly)
    assert abs(lp - np.log(Pb)) <= 1e-12 * max(1.0, abs(lp))
    assert abs(lpb - lp) <= 1e-12 * max(1.0, abs(lp))      # backward logP
    np.testing.assert_allclose(n, nb, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(post, postb, rtol=1e-10, atol=1e-12)


def test_empty_pair_closed_form():
    trans, st = MODELS["tkf91"]()
    em = emissions.pair_hmm_emissions(st, [], [], np.eye(A), np.ones(A) / A)
    lp, post, n, _ = dp2d.forward_backward(tkf.safe_log(trans), st, em)
    assert lp == pytest.approx(np.log(trans[0, 4]), rel=1e-14)
    assert n[0, 4] == pytest.approx(1.0, rel=1e-14)
    assert post[0, 0, 0] == pytest.approx(1.0, rel=1e-14)


@pytest.mark.parametrize("model", ["tkf92", "mixdom_k2f2"])
def test_count_identities(model):
    rng = np.random.default_rng(7)
    trans, st = MODELS[model]()
    st = np.asarray(st)
    P, pi = _rand_sub(rng)
    x = rng.integers(0, A, 23)
    y = rng.integers(0, A, 31)
    em = emissions.pair_hmm_emissions(st, x, y, P, pi)
    lp, post, n, lpb = dp2d.forward_backward(tkf.safe_log(trans), st, em)
    assert abs(lpb - lp) < 1e-11 * abs(lp)
    assert n[:, np.isin(st, (1, 3))].sum() == pytest.approx(23, rel=1e-12)
    assert n[:, np.isin(st, (1, 2))].sum() == pytest.approx(31, rel=1e-12)
    assert n[0].sum() == pytest.approx(1.0, rel=1e-12)
    e = int(np.argmax(st == 4))
    assert n[:, e].sum() == pytest.approx(1.0, rel=1e-12)
    body = np.isin(st, (1, 2, 3))
    np.testing.assert_allclose(n[:, body].sum(0), n[body, :].sum(1), rtol=1e-10)
    # posterior column sums: every x position i>=1 is consumed exactly once
    mdp = post[1:, :, np.isin(st, (1, 3))].sum(axis=(1, 2))
    np.testing.assert_allclose(mdp, 1.0, rtol=1e-10)


def test_counts_are_log_transition_gradient():
    """d logP / d log T[k,k'] == n[k,k'] (central differences)."""
    rng = np.random.default_rng(3)
    trans, st = MODELS["mixdom_k2f2"]()
    P, pi = _rand_sub(rng)
    x = rng.integers(0, A, 6)
    y = rng.integers(0, A, 5)
    em = emissions.pair_hmm_emissions(st, x, y, P, pi)
    lt = tkf.safe_log(trans)
    _, _, n, _ = dp2d.forward_backward(lt, st, em)
    h = 1e-6
    for (k, k2) in [(0, 2), (2, 2), (2, 7), (5, 1), (11, 13), (20, 21), (3, 4)]:
        if trans[k, k2] <= 0:
            continue
        lp_p, _ = dp2d.forward(lt + h * np.eye(len(st))[k][:, None] * np.eye(len(st))[k2][None, :], st, em)
        lp_m, _ = dp2d.forward(lt - h * np.eye(len(st))[k][:, None] * np.eye(len(st))[k2][None, :], st, em)
        assert (lp_p - lp_m) / (2 * h) == pytest.approx(n[k, k2], rel=1e-6, abs=1e-7)


def test_posterior_is_log_emission_gradient():
    """d logP / d e(i,j,k) == posterior(i,j,k)."""
    rng = np.random.default_rng(4)
    trans, st = MODELS["tkf92"]()
    P, pi = _rand_sub(rng)
    x = rng.integers(0, A, 5)
    y = rng.integers(0, A, 4)
    em = emissions.pair_hmm_emissions(st, x, y, P, pi)
    lt = tkf.safe_log(trans)
    _, post, _, _ = dp2d.forward_backward(lt, st, em)
    h = 1e-6
    for (i, j, k) in [(1, 1, 1), (2, 3, 1), (0, 2, 2), (3, 0, 3), (5, 4, 1), (4, 4, 2)]:
        ep = em.copy(); ep[i, j, k] += h
        emn = em.copy(); emn[i, j, k] -= h
        d = (dp2d.forward(lt, st, ep)[0] - dp2d.forward(lt, st, emn)[0]) / (2 * h)
        assert d == pytest.approx(post[i, j, k], rel=1e-6, abs=1e-7)


def _tab_from_logs(log_sub, log_pi):
    return np.concatenate([(log_pi[:, None] + log_sub).ravel(), log_pi, log_pi])


@pytest.mark.parametrize("model", ["tkf91", "mixdom_k2f2"])
def test_c_oracle_matches_numpy(model):
    rng = np.random.default_rng(9)
    trans, st = MODELS[model]()
    P, pi = _rand_sub(rng)
    log_sub, log_pi = emissions.log_tables_shared(P, pi)
    tab = _tab_from_logs(log_sub, log_pi)[None]
    xs = [rng.integers(0, A, n) for n in (0, 3, 17, 40)]
    ys = [rng.integers(0, A, n) for n in (2, 0, 21, 33)]
    lt = tkf.safe_log(trans)
    lp, n, posts = cfb2d.fb2d_batch(lt, st, np.zeros(len(st)), tab, xs, ys, want_posteriors=True,
                                    nthreads=2)
    for b, (x, y) in enumerate(zip(xs, ys)):
        em = emissions.pair_hmm_emissions(st, x, y, P, pi)
        lp_o, post_o, n_o, _ = dp2d.forward_backward(lt, st, em)
        assert lp[b] == pytest.approx(lp_o, rel=1e-12)
        np.testing.assert_allclose(n[b], n_o, rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(posts[b], post_o, rtol=1e-10, atol=1e-13)
