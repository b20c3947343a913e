"""Pins the restoration / BDI oracle (oracle/suffstats.py, oracle/bdi.py):

* eliminating the null groups of the exploded machine reproduces the
  collapsed chi of oracle/mixdom.py (two independent constructions of the
  same model);
* the restored counts satisfy the score identity
      d/dtheta sum n_chi log chi(theta) = sum n_full d log T_full / dtheta
  for every structural parameter (finite differences), for arbitrary n_chi;
* restored expected counts conserve flow through every null state;
* BDI: the conservation law E_B - E_D = sum n[:,I] - sum n[:,D], the scores
  against double-precision finite differences, and continuity of E_S across
  the equal-rate switch;
* the TKF92 self-loop split against the ext score identity.
CPU only."""
import numpy as np
import pytest

from oracle import bdi, dp2d, emissions, mixdom, suffstats, tkf

CASES = [
    dict(K=3, F=1, ext=[[0.6], [0.8], [0.9]]),
    dict(K=2, F=2, ext=[[[0.5, 0.2], [0.1, 0.6]], [[0.3, 0.3], [0.0, 0.85]]]),
    dict(K=1, F=3, ext=[[0.7, 0.5, 0.9]]),
]


def _params(case, seed):
    rng = np.random.default_rng(seed)
    K, F = case["K"], case["F"]
    return dict(main_ins=0.04, main_del=0.05, dom_ins=rng.uniform(0.02, 0.06, K),
                dom_del=rng.uniform(0.05, 0.08, K), dom_w=rng.dirichlet(np.ones(K)),
                frag_w=rng.dirichlet(np.ones(F), size=K), ext=np.asarray(case["ext"], float))


def _chi(p, t):
    return mixdom.build_nested_trans(p["main_ins"], p["main_del"], t, p["dom_ins"], p["dom_del"],
                                     p["dom_w"], p["frag_w"], p["ext"])


def _full(p, t):
    return suffstats.build_exploded(p["main_ins"], p["main_del"], t, p["dom_ins"], p["dom_del"],
                                    p["dom_w"], p["frag_w"], p["ext"])


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"K{c['K']}F{c['F']}")
@pytest.mark.parametrize("t", [0.05, 0.7, 2.0])
def test_elimination_reproduces_chi(case, t):
    p = _params(case, 1)
    K, F = case["K"], case["F"]
    T, names = _full(p, t)
    assert T.shape[0] == 8 + 8 * K + 15 * K * F
    rows = T.sum(axis=1)
    end = names.index("End")
    np.testing.assert_allclose(np.delete(rows, end), 1.0, rtol=1e-12)
    Tr, kept, _ = suffstats.eliminate_chain(T, names, K, F)
    assert kept[:2] == ["Start", "End"]
    emit = [n for n in kept[2:]]
    want = [f"{e}[{k},{f}]" for k in range(K) for e in suffstats._EMIT_KINDS for f in range(F)]
    assert emit == want
    np.testing.assert_allclose(Tr, _chi(p, t), rtol=1e-11, atol=1e-15)


def _perturb(p, key, idx, h):
    q = {k: np.array(v, dtype=np.float64, copy=True) for k, v in p.items()}
    q[key][idx] += h
    return q


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"K{c['K']}F{c['F']}")
def test_restored_counts_score_identity(case):
    """Restoration distributes each collapsed count over the exploded
    transitions in proportion to their posterior usage, so the gradient of
    the collapsed complete-data score equals the exploded one."""
    K, F = case["K"], case["F"]
    p = _params(case, 2)
    t = 0.9
    rng = np.random.default_rng(3)
    chi = _chi(p, t)
    n_chi = np.where(chi > 0, rng.uniform(0.1, 3.0, chi.shape), 0.0)
    ss = suffstats.exact_suffstats(n_chi, t=t, **p)
    n_full = ss["n_full"]
    h = 1e-6
    checks = [("t", None), ("main_ins", ()), ("main_del", ()), ("dom_ins", (0,)), ("dom_del", (K - 1,)),
              ("dom_w", (0,)), ("frag_w", (0, F - 1)), ("ext", (0, F - 1) if p["ext"].ndim == 2 else (0, 1, 0))]
    for key, idx in checks:
        if key == "t":
            cp, cm = _chi(p, t + h), _chi(p, t - h)
            Tp, Tm = _full(p, t + h)[0], _full(p, t - h)[0]
        else:
            qp, qm = _perturb(p, key, idx, h), _perturb(p, key, idx, -h)
            cp, cm = _chi(qp, t), _chi(qm, t)
            Tp, Tm = _full(qp, t)[0], _full(qm, t)[0]
        pos = chi > 0
        lhs = (n_chi[pos] * (np.log(cp[pos]) - np.log(cm[pos]))).sum() / (2 * h)
        T0 = _full(p, t)[0]
        posf = T0 > 0
        rhs = (n_full[posf] * (np.log(Tp[posf]) - np.log(Tm[posf]))).sum() / (2 * h)
        assert rhs == pytest.approx(lhs, rel=1e-6, abs=1e-7), key


def test_restored_expected_counts_conserve_flow():
    """n_chi from a real forward-backward pass: every null state of the
    exploded machine has in-flow == out-flow, Start emits one unit, End
    absorbs one, and the extracted groups agree with each other."""
    case = CASES[0]
    K, F = case["K"], case["F"]
    p = _params(case, 4)
    t = 0.6
    chi = _chi(p, t)
    st = mixdom.state_types(K, F)
    rng = np.random.default_rng(5)
    P = rng.dirichlet(np.ones(20), size=20)
    pi = rng.dirichlet(np.ones(20))
    x, y = rng.integers(0, 20, 14), rng.integers(0, 20, 17)
    em = emissions.pair_hmm_emissions(st, x, y, P, pi)
    _, _, n_chi, _ = dp2d.forward_backward(tkf.safe_log(chi), st, em)
    ss = suffstats.exact_suffstats(n_chi, t=t, **p)
    n = ss["n_full"]
    names = suffstats.state_names(K, F)
    for u, name in enumerate(names):
        if name in ("Start", "End"):
            continue
        assert n[u].sum() == pytest.approx(n[:, u].sum(), rel=1e-10, abs=1e-12), name
    assert n[names.index("Start")].sum() == pytest.approx(1.0, rel=1e-12)
    assert n[:, names.index("End")].sum() == pytest.approx(1.0, rel=1e-12)
    top = ss["top_5x5"]
    assert ss["dom_w"].sum() == pytest.approx(top[:, 1:4].sum(), rel=1e-10)
    assert top[:, 4].sum() == pytest.approx(1.0, rel=1e-12)
    # ancestor residues: M-type and D-type emissions of the exploded machine
    emits = [names.index(f"{e}[{k},{f}]") for k in range(K) for e in ("MatEmit", "DelEmit", "DEmit")
             for f in range(F)]
    assert n[:, emits].sum() == pytest.approx(len(x), rel=1e-10)


def test_bdi_conservation_and_scores():
    rng = np.random.default_rng(6)
    n = rng.uniform(0.0, 5.0, (5, 5))
    n[:, 0] = 0
    n[4] = 0
    lam, mu, t, T = 0.04, 0.05, 0.7, 3.1
    EB, ED, ES = bdi.bdi_stats(n, lam, mu, t, T)
    cons = n[:, 2].sum() - n[:, 3].sum()
    assert EB - ED == pytest.approx(cons, rel=1e-12, abs=1e-12)
    # scores vs double-precision central differences of the WFST logP
    import mpmath
    g = bdi.count_groups(n)
    f = lambda a, b: float(bdi._logP(mpmath.mpf(a), mpmath.mpf(b), t, g))
    h = 1e-7
    dl = (f(lam + h, mu) - f(lam - h, mu)) / (2 * h)
    dm = (f(lam, mu + h) - f(lam, mu - h)) / (2 * h)
    ES_chk = (cons + mu * dm - lam * dl - lam * T) / (lam - mu)
    assert ES == pytest.approx(ES_chk, rel=1e-6)


def _dp_counts(lam, mu, t, seed):
    """5x5 expected counts of a TKF91 pair from the oracle DP."""
    rng = np.random.default_rng(seed)
    P = rng.dirichlet(np.ones(20), size=20)
    pi = rng.dirichlet(np.ones(20))
    x, y = rng.integers(0, 20, 30), rng.integers(0, 20, 25)
    em = emissions.pair_hmm_emissions(tkf.TKF_STATE_TYPES, x, y, P, pi)
    return dp2d.forward_backward(tkf.safe_log(tkf.tkf91_trans(lam, mu, t)),
                                 tkf.TKF_STATE_TYPES, em)[2]


def test_bdi_limit_branch_continuous():
    """For expected counts of a real pair (observation time T = t * sum of
    the E column), E_S of the equal-rate branch at kappa = 1 equals the
    limit of the general formula, evaluated at 40 digits just off kappa = 1
    (the general numerator vanishes there)."""
    mu, t = 0.05, 0.9
    n = _dp_counts(0.04, mu, t, 7)
    T = t * n[:, 4].sum()
    lim = bdi.bdi_stats(n, mu, mu, t, T)
    for d in (1e-6, -1e-6):
        gen = bdi.bdi_stats(n, mu * (1 - d), mu, t, T)
        for a, b in zip(lim, gen):
            assert a == pytest.approx(b, rel=1e-5, abs=1e-6)


def test_tkf92_resolve_ext_score_identity():
    """d/d ext sum n log tau92(ext) == ext_count/ext - notext/(1-ext)."""
    rng = np.random.default_rng(8)
    lam, mu, t, ext = 0.04, 0.05, 0.8, 0.6
    tau = tkf.tkf92_trans(lam, mu, t, ext)
    n = np.where(tau > 0, rng.uniform(0.2, 4.0, (5, 5)), 0.0)
    res, ec, ne = bdi.tkf92_resolve(n, ext, tau)
    h = 1e-7
    pos = tau > 0
    lp = (n[pos] * np.log(tkf.tkf92_trans(lam, mu, t, ext + h)[pos])).sum()
    lm = (n[pos] * np.log(tkf.tkf92_trans(lam, mu, t, ext - h)[pos])).sum()
    assert (lp - lm) / (2 * h) == pytest.approx(ec / ext - ne / (1 - ext), rel=1e-7)
    np.testing.assert_allclose(res.sum() + ec, n.sum(), rtol=1e-13)
