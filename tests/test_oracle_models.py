"""Pins the oracle's model builders (oracle/tkf.py, oracle/mixdom.py,
oracle/emissions.py) with closed forms, invariants and high-precision
evaluation, and checks the package's torch builders (tkfmi.params,
tkfmi.mixdom, tkfmi.tables) against them.  CPU only."""
import numpy as np
import pytest
import torch

from oracle import emissions, mixdom, tkf

A = 20
GRID = [(0.04, 0.05, 0.5), (0.04, 0.05, 0.05), (0.03, 0.09, 2.0), (0.1, 0.1, 0.7),
        (0.05, 0.05 * (1 + 3e-5), 1.3), (0.02, 0.3, 0.01)]


@pytest.mark.parametrize("lam,mu,t", GRID)
def test_tkf_rows_sum_to_one(lam, mu, t):
    """core/params.py:29-54, 92-108: rows S, M, I, D are distributions."""
    t91 = tkf.tkf91_trans(lam, mu, t)
    np.testing.assert_allclose(t91[:4].sum(axis=1), 1.0, rtol=1e-13)
    assert np.all(t91 >= 0)
    for ext in (0.0, 0.3, 0.9):
        t92 = tkf.tkf92_trans(lam, mu, t, ext)
        np.testing.assert_allclose(t92[:4].sum(axis=1), 1.0, rtol=1e-13)
    np.testing.assert_array_equal(tkf.tkf92_trans(lam, mu, t, 0.0), t91)


@pytest.mark.parametrize("lam,mu,t", GRID)
def test_tkf91_cond_divides_out_length_prior(lam, mu, t):
    """core/params.py:57-89: kappa removed from M/D entries, (1-kappa) from E."""
    k = lam / mu
    if abs(1 - k) < 1e-3:
        return  # (1 - kappa) ~ 0: the relation is ill-conditioned there
    t91, tc = tkf.tkf91_trans(lam, mu, t), tkf.tkf91_trans_cond(lam, mu, t)
    for r in (tkf.S, tkf.M, tkf.I, tkf.D):
        assert tc[r, tkf.M] == pytest.approx(t91[r, tkf.M] / k, rel=1e-12)
        assert tc[r, tkf.D] == pytest.approx(t91[r, tkf.D] / k, rel=1e-12)
        assert tc[r, tkf.I] == pytest.approx(t91[r, tkf.I], rel=1e-12)
        assert tc[r, tkf.E] == pytest.approx(t91[r, tkf.E] / (1 - k), rel=1e-10)


def _mp_beta_gamma(lam, mu, t):
    """TKF beta/gamma from their defining formulas at 50 digits
    (core/bdi.py:45-93 general branch; at lam == mu the formulas' limits)."""
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 50
    lam, mu, t = mpmath.mpf(lam), mpmath.mpf(mu), mpmath.mpf(t)
    if lam == mu:
        s = mu * t
        beta = s / (1 + s)
        gamma = 1 - 1 / ((1 + s) * ((1 - mpmath.e ** (-s)) / s))
        return float(beta), float(gamma)
    eta, alpha = mpmath.e ** (-lam * t), mpmath.e ** (-mu * t)
    beta = lam * (eta - alpha) / (mu * eta - lam * alpha)
    gamma = 1 - mu * beta / (lam * (1 - alpha))
    return float(beta), float(gamma)


@pytest.mark.parametrize("kgap", [0.0, 5e-5, -5e-5, 1.5e-4, -1.5e-4, 1e-2])
def test_beta_gamma_vs_mpmath_across_lhopital_switch(kgap):
    """Equal-rate (L'Hopital) branch for |1-kappa| < 1e-4 (core/bdi.py:14,
    64-66, 89-91): exact at kappa = 1, within O(|1-kappa|) just inside the
    switch, and the general branch accurate just outside it."""
    mu, t = 0.05, 0.8
    lam = mu * (1 - kgap)
    b, g = float(tkf.tkf_beta(lam, mu, t)), float(tkf.tkf_gamma(lam, mu, t))
    bm, gm = _mp_beta_gamma(lam, mu, t)
    tol = 1e-13 if kgap == 0.0 else (2 * abs(kgap) if abs(kgap) < 1e-4 else 1e-8)
    assert b == pytest.approx(bm, rel=tol, abs=1e-15)
    assert g == pytest.approx(gm, rel=tol, abs=1e-15)


MIX_CASES = [
    dict(K=3, F=1, t=0.7, ext=[[0.6], [0.8], [0.9]]),
    dict(K=2, F=2, t=1.3, ext=[[[0.5, 0.2], [0.1, 0.6]], [[0.3, 0.3], [0.0, 0.85]]]),
    dict(K=1, F=3, t=0.1, ext=[[0.7, 0.5, 0.9]]),
]


def _mix_chi(case, rng):
    K, F = case["K"], case["F"]
    dom_w = rng.dirichlet(np.ones(K))
    frag_w = rng.dirichlet(np.ones(F), size=K)
    return mixdom.build_nested_trans(0.04, 0.05, case["t"], rng.uniform(0.02, 0.06, K),
                                     rng.uniform(0.05, 0.08, K), dom_w, frag_w, case["ext"])


@pytest.mark.parametrize("case", MIX_CASES, ids=lambda c: f"K{c['K']}F{c['F']}")
def test_mixdom_chi_rows(case):
    """Collapsed chi (models/mixdom.py:252-423): SS and body rows are
    distributions, nothing enters SS, EE has no out-transitions."""
    chi = _mix_chi(case, np.random.default_rng(case["K"] * 10 + case["F"]))
    st = mixdom.state_types(case["K"], case["F"])
    assert chi.shape == (len(st), len(st))
    assert np.all(chi >= -1e-15)
    np.testing.assert_allclose(chi[0].sum(), 1.0, rtol=1e-12)
    np.testing.assert_allclose(chi[2:].sum(axis=1), 1.0, rtol=1e-12)
    assert np.all(chi[1] == 0) and np.all(chi[:, 0] == 0)


def test_mixdom_k1f1_shape_matches_tkf_structure():
    """K=1, F=1 MixDom with a single domain: the body is one M/I/D triple
    plus the two extra (II, DD) slots; state types follow mixdom.py:69-77."""
    st = mixdom.state_types(1, 1)
    np.testing.assert_array_equal(st, [0, 4, 1, 2, 3, 2, 3])
    st = mixdom.state_types(2, 2)
    assert list(st[2:12]) == [1, 1, 2, 2, 3, 3, 2, 2, 3, 3]


def _rand_sub(rng, n=None):
    pi = rng.dirichlet(np.ones(A))
    P = rng.dirichlet(np.ones(A), size=A if n is None else (n, A))
    return P, pi


def test_shared_emissions_direct_cells():
    """pair_hmm_emissions (dp/hmm.py:628-670) evaluated cell by cell."""
    rng = np.random.default_rng(1)
    P, pi = _rand_sub(rng)
    P[3, 5] = 0.0
    x, y = rng.integers(0, A, 6), rng.integers(0, A, 5)
    x[2], y[1] = 3, 5
    x[4] = 20  # wildcard
    st = tkf.TKF_STATE_TYPES
    em = emissions.pair_hmm_emissions(st, x, y, P, pi)
    lpi = lambda a: 0.0 if a >= A else np.log(pi[a] + 1e-30)
    lsub = lambda a, b: 0.0 if (a >= A or b >= A) else np.log(P[a, b] + 1e-30)
    for i in range(7):
        for j in range(6):
            a = x[i - 1] if i else 0
            b = y[j - 1] if j else 0
            assert em[i, j, 1] == pytest.approx(lpi(a) + lsub(a, b), rel=1e-14)
            assert em[i, j, 2] == pytest.approx(lpi(b), rel=1e-14)
            assert em[i, j, 3] == pytest.approx(lpi(a), rel=1e-14)
            assert em[i, j, 0] == 0.0 and em[i, j, 4] == 0.0
    assert em[3, 2, 1] == pytest.approx(np.log(pi[3] + 1e-30) + np.log(1e-30), rel=1e-14)


def test_per_class_single_class_equals_per_domain():
    """C = 1 per-class emissions == per-domain emissions (class weight 1)."""
    rng = np.random.default_rng(2)
    K, F = 2, 2
    st = mixdom.state_types(K, F)
    subs, pi = _rand_sub(rng, n=K)
    pis = np.stack([pi, rng.dirichlet(np.ones(A))])
    x, y = rng.integers(0, A, 7), rng.integers(0, A, 9)
    per_dom = emissions.pair_hmm_emissions_per_domain(st, x, y, subs, pis, K, F)
    # per-class with one class per domain, weight 1 on the domain's class
    cd = np.zeros((K, F, K))
    for d in range(K):
        cd[d, :, d] = 1.0
    per_cls = emissions.pair_hmm_emissions_per_class(st, x, y, subs, pis, cd, K, F)
    np.testing.assert_allclose(per_cls, per_dom, rtol=1e-13, atol=1e-13)


def _groups_per_class(S, F):
    out = [0, 0]
    for s in range(2, S):
        b = s - 2
        out.append((b // (5 * F)) * F + (b % (5 * F)) % F)
    return out


def test_tables_reproduce_emission_builders():
    """The kernel's per-group table format reproduces all three reference
    emission builders cell by cell."""
    rng = np.random.default_rng(3)
    K, F, C = 2, 2, 3
    st = mixdom.state_types(K, F)
    S = len(st)
    x, y = rng.integers(0, A, 8), rng.integers(0, A, 6)
    x[5] = 20
    # shared
    P, pi = _rand_sub(rng)
    ls, lp = emissions.log_tables_shared(P, pi)
    tab = np.concatenate([(lp[:, None] + ls).ravel(), lp, lp])[None]
    np.testing.assert_allclose(emissions.emit_from_tables(st, [0] * S, tab, x, y),
                               emissions.pair_hmm_emissions(st, x, y, P, pi), rtol=1e-14)
    # per domain
    subs, _ = _rand_sub(rng, n=K)
    pis = rng.dirichlet(np.ones(A), size=K)
    lss, lps = emissions.log_tables_per_domain(subs, pis)
    tab = np.stack([np.concatenate([(lps[d][:, None] + lss[d]).ravel(), lps[d], lps[d]])
                    for d in range(K)])
    grp = [0, 0] + [(s - 2) // (5 * F) for s in range(2, S)]
    np.testing.assert_allclose(emissions.emit_from_tables(st, grp, tab, x, y),
                               emissions.pair_hmm_emissions_per_domain(st, x, y, subs, pis, K, F),
                               rtol=1e-14)
    # per class
    csubs, _ = _rand_sub(rng, n=C)
    cpis = rng.dirichlet(np.ones(A), size=C)
    cd = rng.dirichlet(np.ones(C), size=(K, F))
    cd[1, 0, 2] = 0.0
    cd[1, 0] /= cd[1, 0].sum()
    m, sgl = emissions.log_mixture_tables_per_class(csubs, cpis, cd)
    tab = np.concatenate([m.reshape(K * F, -1), sgl.reshape(K * F, -1), sgl.reshape(K * F, -1)], 1)
    np.testing.assert_allclose(
        emissions.emit_from_tables(st, _groups_per_class(S, F), tab, x, y),
        emissions.pair_hmm_emissions_per_class(st, x, y, csubs, cpis, cd, K, F),
        rtol=1e-13, atol=1e-13)


# ------------------------------------------------ package builders vs oracle
@pytest.mark.parametrize("lam,mu,t", GRID)
def test_torch_tkf_builders_match_oracle(lam, mu, t):
    from tkfmi import params
    np.testing.assert_allclose(params.tkf91_trans(lam, mu, t).numpy(),
                               tkf.tkf91_trans(lam, mu, t), rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(params.tkf91_trans_cond(lam, mu, t).numpy(),
                               tkf.tkf91_trans_cond(lam, mu, t), rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(params.tkf92_trans(lam, mu, t, 0.6).numpy(),
                               tkf.tkf92_trans(lam, mu, t, 0.6), rtol=1e-13, atol=1e-300)


@pytest.mark.parametrize("case", MIX_CASES, ids=lambda c: f"K{c['K']}F{c['F']}")
def test_torch_mixdom_chi_matches_oracle(case):
    from tkfmi import mixdom as tmix
    rng = np.random.default_rng(5)
    K, F = case["K"], case["F"]
    dl, dm = rng.uniform(0.02, 0.06, K), rng.uniform(0.05, 0.08, K)
    dw, fw = rng.dirichlet(np.ones(K)), rng.dirichlet(np.ones(F), size=K)
    ts = np.array([case["t"], 0.3, 1.9])
    chi_t = tmix.build_nested_trans(0.04, 0.05, torch.tensor(ts), dl, dm, dw, fw, case["ext"])
    for b, t in enumerate(ts):
        chi_o = mixdom.build_nested_trans(0.04, 0.05, t, dl, dm, dw, fw, case["ext"])
        np.testing.assert_allclose(chi_t[b].numpy(), chi_o, rtol=1e-12, atol=1e-15)


def test_torch_tables_match_oracle():
    from tkfmi import tables
    rng = np.random.default_rng(6)
    K, F, C = 3, 1, 2
    csubs, _ = _rand_sub(rng, n=C)
    cpis = rng.dirichlet(np.ones(A), size=C)
    cd = rng.dirichlet(np.ones(C), size=(K, F))
    m, sgl = emissions.log_mixture_tables_per_class(csubs, cpis, cd)
    want = np.concatenate([m.reshape(K * F, -1), sgl.reshape(K * F, -1), sgl.reshape(K * F, -1)], 1)
    got = tables.per_class_tables(torch.tensor(csubs), torch.tensor(cpis), torch.tensor(cd)).numpy()
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
    P, pi = _rand_sub(rng)
    ls, lp = emissions.log_tables_shared(P, pi)
    want = np.concatenate([(lp[:, None] + ls).ravel(), lp, lp])[None]
    np.testing.assert_allclose(tables.shared_tables(torch.tensor(P), torch.tensor(pi)).numpy(),
                               want, rtol=1e-14)
