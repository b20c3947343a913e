"""Device parameter builder (tkfmi_build_params_batched) vs the oracle
(oracle.tkf, oracle.mixdom, oracle.emissions) with substitution matrices
from scipy.linalg.expm, for every model and emission mode, and on the six
trained parameter sets the reference ships (tests/golden/params_best.npz,
python/params/best/*.npz).

Tolerances: log transitions 1e-12 (absolute on values of magnitude <= 1,
relative above) where the TKF closed forms are well conditioned, and their
measured FP64 conditioning (``_chi_tol``) at short times; log emission
tables 1e-12 likewise; P(t) recovered from
the per-domain tables against scipy.linalg.expm 1e-12 relative to its
largest entry.  The device evaluates expm by Pade scaling and squaring (the
rule of jax.scipy.linalg.expm, which the reference calls), for any Q --
bw_d8f2_fullseed_15iter's Q is 1.4e-8 away from reversible and its rows sum
to 6e-8, and it must be taken as given."""
import os

import numpy as np
import pytest
import scipy.linalg
import torch

pytestmark = pytest.mark.gpu
A = 20
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "params_best.npz")
SETS = ("bw_d3f2_fullseed_15iter", "bw_d3f2_pfam100_20iter", "bw_d5f2_pfam100_20iter",
        "bw_d8f2_fullseed_15iter", "bw_d8f2_pfam100_15iter", "svi_d3f2_fullseed_15iter")
TS = np.array([0.0, 0.01, 0.05, 0.3, 0.9, 1.7, 2.0, 4.0, 30.0])


def _gtr(rng, n):
    Qs, pis = [], []
    for _ in range(n):
        S = rng.gamma(1.0, 1.0, (A, A))
        S = (S + S.T) / 2
        pi = rng.dirichlet(np.ones(A))
        Q = S * pi[None, :]
        np.fill_diagonal(Q, 0.0)
        np.fill_diagonal(Q, -Q.sum(axis=1))
        Qs.append(Q); pis.append(pi)
    return np.stack(Qs), np.stack(pis)


def _cmp(got, want, tol=1e-12, what=""):
    got = got.cpu().numpy() if isinstance(got, torch.Tensor) else np.asarray(got)
    want = np.asarray(want, np.float64)
    err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
    assert err.max() <= tol, (what, float(err.max()))


def _chi_tol(fn):
    """Tolerance for the log transitions ``fn()`` (oracle, probability
    space) evaluates.  The TKF closed forms (core/bdi.py:44-97) subtract
    exp(-lam t) - exp(-mu t) and then 1 - mu beta / (lam (1 - alpha)), which
    cancel when (mu - lam) t and mu t are small (the equal-rate branch takes
    over only at |1 - lam/mu| < 1e-4), so one ulp of difference between two
    exp implementations (the GPU's and the host's) moves chi by far more
    than one ulp.  The tolerance is 16 x the largest change of log chi when
    every exp of oracle.tkf is perturbed by one ulp with alternating signs
    (both sign patterns), and at least 1e-12."""
    import oracle.tkf as ot
    base = ot.floor_log(fn())
    spread = 0.0
    for start in (1.0, -1.0):
        calls = [0]

        class _Np:
            def __getattr__(self, k):
                return getattr(np, k)

            def exp(self, x):
                calls[0] += 1
                return np.exp(x) * (1.0 + start * (-1.0) ** calls[0] * 2.0 ** -52)

        saved, ot.np = ot.np, _Np()
        try:
            pert = ot.floor_log(fn())
        finally:
            ot.np = saved
        spread = max(spread, float((np.abs(pert - base) / np.maximum(1.0, np.abs(base))).max()))
    return max(1e-12, 16.0 * spread)


def _tab_shared_or_domain(log_subs, log_pis):
    """Oracle (G,21,21)+(G,21) log tables -> the kernel's (G,483) layout."""
    G = log_subs.shape[0]
    m = log_pis[:, :, None] + log_subs
    return np.concatenate([m.reshape(G, -1), log_pis, log_pis], axis=1)


def _tab_per_class(match, single):
    K, F = match.shape[:2]
    return np.concatenate([match.reshape(K * F, -1), single.reshape(K * F, -1),
                           single.reshape(K * F, -1)], axis=1)


def _load_set(name):
    d = np.load(GOLDEN, allow_pickle=False)
    g = lambda k: d[f"{name}/{k}"]
    return dict(main_ins=float(g("main_ins")), main_del=float(g("main_del")), dom_ins=g("dom_ins"),
                dom_del=g("dom_del"), dom_w=g("dom_weights"), frag_w=g("frag_weights"),
                ext=g("ext_rates"), Qs=g("dom_Qs"), pis=g("dom_pis"), S_exch=g("dom_S_exch"),
                t_rep=float(g("t_rep")))


@pytest.mark.parametrize("model", ["tkf91", "tkf91_cond", "tkf92"])
@pytest.mark.parametrize("floor", [False, True])
def test_tkf_builder(model, floor):
    from oracle import emissions, tkf
    from tkfmi import build
    rng = np.random.default_rng(1)
    Qs, pis = _gtr(rng, 1)
    ts = TS[TS > 0]
    b = build.DeviceBuilder(model, "shared", Qs, pis, lam=0.04, mu=0.05, ext=0.6, log_floor=floor)
    lt, tab, _ = b.build(ts)
    for i, t in enumerate(ts):
        fn = {"tkf91": lambda: tkf.tkf91_trans(0.04, 0.05, t),
              "tkf91_cond": lambda: tkf.tkf91_trans_cond(0.04, 0.05, t),
              "tkf92": lambda: tkf.tkf92_trans(0.04, 0.05, t, 0.6)}[model]
        tau = fn()
        want = tkf.floor_log(tau) if floor else tkf.safe_log(tau)
        _cmp(lt[i], want, _chi_tol(fn), what=f"trans t={t}")
        ls, lp = emissions.log_tables_shared(scipy.linalg.expm(Qs[0] * t), pis[0])
        _cmp(tab[i], _tab_shared_or_domain(ls[None], lp[None]), what=f"tab t={t}")


@pytest.mark.parametrize("K,F,C", [(3, 1, 1), (3, 1, 2), (2, 2, 3), (1, 3, 1)])
def test_mixdom_builder_per_class_and_per_domain(K, F, C):
    from oracle import emissions, mixdom, tkf
    from tkfmi import build
    rng = np.random.default_rng(K * 10 + F + C)
    mix = dict(main_ins=0.04, main_del=0.05, dom_ins=rng.uniform(0.02, 0.06, K),
               dom_del=rng.uniform(0.05, 0.08, K), dom_w=rng.dirichlet(np.ones(K)),
               frag_w=rng.dirichlet(np.ones(F), size=K),
               ext=rng.uniform(0.5, 0.9, (K, F, 1)) * rng.dirichlet(np.ones(F), size=(K, F)))
    Qs, pis = _gtr(rng, C)
    cd = rng.dirichlet(np.ones(C), size=(K, F))
    cd[0, 0, 0] = 0.0
    cd[0, 0] /= cd[0, 0].sum() if cd[0, 0].sum() > 0 else 1.0
    ts = TS[TS > 0]
    b = build.DeviceBuilder("mixdom", "per_class", Qs, pis, mix=mix, class_dist=cd, log_floor=True)
    lt, tab, _ = b.build(ts)
    Qd, pd_ = _gtr(rng, K)
    b2 = build.DeviceBuilder("mixdom", "per_domain", Qd, pd_, mix=mix, log_floor=False)
    lt2, tab2, _ = b2.build(ts)
    for i, t in enumerate(ts):
        chi = mixdom.build_nested_trans(0.04, 0.05, t, mix["dom_ins"], mix["dom_del"], mix["dom_w"],
                                        mix["frag_w"], mix["ext"])
        tol = _chi_tol(lambda: mixdom.build_nested_trans(0.04, 0.05, t, mix["dom_ins"], mix["dom_del"],
                                                         mix["dom_w"], mix["frag_w"], mix["ext"]))
        _cmp(lt[i], tkf.floor_log(chi), tol, what=f"chi t={t}")
        _cmp(lt2[i], tkf.safe_log(chi), tol, what=f"chi safe_log t={t}")
        subs = np.stack([scipy.linalg.expm(Qs[c] * t) for c in range(C)])
        _cmp(tab[i], _tab_per_class(*emissions.log_mixture_tables_per_class(subs, pis, cd)),
             what=f"per-class tab t={t}")
        subs_d = np.stack([scipy.linalg.expm(Qd[k] * t) for k in range(K)])
        _cmp(tab2[i], _tab_shared_or_domain(*emissions.log_tables_per_domain(subs_d, pd_)),
             what=f"per-domain tab t={t}")


@pytest.mark.parametrize("name", SETS)
def test_params_best_expm_and_tables(name):
    """P(t) for every domain rate matrix of the reference's trained sets,
    recovered from the per-domain tables (match entry minus log pi), against
    scipy.linalg.expm; the tables and chi against the oracle."""
    from oracle import emissions, mixdom, tkf
    from tkfmi import build
    p = _load_set(name)
    K, F = p["ext"].shape
    ts = np.concatenate([TS, [p["t_rep"]]])
    mix = dict(main_ins=p["main_ins"], main_del=p["main_del"], dom_ins=p["dom_ins"],
               dom_del=p["dom_del"], dom_w=p["dom_w"], frag_w=p["frag_w"], ext=p["ext"])
    b = build.DeviceBuilder("mixdom", "per_domain", p["Qs"], p["pis"], mix=mix, log_floor=True)
    lt, tab, _ = b.build(ts)
    tab = tab.cpu().numpy()
    for i, t in enumerate(ts):
        P_ref = np.stack([scipy.linalg.expm(p["Qs"][k] * t) for k in range(K)])
        logP = tab[i, :, :441].reshape(K, 21, 21)[:, :20, :20] - np.log(p["pis"])[:, :, None]
        P_dev = np.exp(logP)
        err = np.abs(P_dev - P_ref).max() / np.abs(P_ref).max()
        assert err <= 1e-12, (name, t, err)
        _cmp(tab[i], _tab_shared_or_domain(*emissions.log_tables_per_domain(P_ref, p["pis"])),
             what=f"{name} tab t={t}")
        if t > 0:
            chi = mixdom.build_nested_trans(p["main_ins"], p["main_del"], t, p["dom_ins"], p["dom_del"],
                                            p["dom_w"], p["frag_w"], p["ext"])
            tol = _chi_tol(lambda: mixdom.build_nested_trans(
                p["main_ins"], p["main_del"], t, p["dom_ins"], p["dom_del"], p["dom_w"], p["frag_w"],
                p["ext"]))
            _cmp(lt[i], tkf.floor_log(chi), tol, what=f"{name} chi t={t}")


def test_builder_takes_nonreversible_rate_matrix():
    """A rate matrix with no detailed balance (the reference's
    transition_matrix: 'Need not be reversible') is exponentiated as is."""
    from oracle import emissions
    from tkfmi import build
    rng = np.random.default_rng(3)
    Q = rng.uniform(0.1, 1.0, (A, A))
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    pi = rng.dirichlet(np.ones(A))
    b = build.DeviceBuilder("tkf92", "shared", Q, pi, lam=0.04, mu=0.05, ext=0.5)
    ts = np.array([0.02, 0.4, 1.3, 7.0])
    _, tab, _ = b.build(ts)
    for i, t in enumerate(ts):
        ls, lp = emissions.log_tables_shared(scipy.linalg.expm(Q * t), pi)
        _cmp(tab[i], _tab_shared_or_domain(ls[None], lp[None]), what=f"t={t}")


def test_builder_expm_too_many_squarings_is_nan():
    """Beyond 16 squarings (1-norm of Q t above 5.37 * 2**17)
    jax.scipy.linalg.expm returns NaN; so does the device builder."""
    from tkfmi import build
    rng = np.random.default_rng(5)
    Qs, pis = _gtr(rng, 1)
    b = build.DeviceBuilder("tkf91", "shared", Qs, pis, lam=0.04, mu=0.05)
    _, tab, _ = b.build(np.array([1e7, 1.0]))
    tab = tab.cpu().numpy()
    assert np.isnan(tab[0, 0, :441].reshape(21, 21)[:20, :20]).all()
    assert np.isfinite(tab[1]).all()
