"""The score identity behind the gradients, checked on the CPU:

    d logP / d theta = sum n d log chi / d theta + sum c d log tab / d theta

with n and c (expected transition / emission counts) from the oracle DP, and
the derivatives of log chi and of the log emission tables taken by torch
autograd through the package's builders (tkfmi.mixdom, tkfmi.tables,
tkfmi.vjp.build_class_subs / build_Q_from_S_pi).  Compared with central
finite differences (h = 1e-6) of the oracle logP, for every MixDom2 and TKF92
parameter.  This is exactly the composition tkfmi.vjp performs on the GPU,
where the kernel supplies n and c."""
import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import dp2d, emissions, mixdom, tkf

A = 20
H = 1e-6


def _gtr(S_exch, pi):
    Q = S_exch * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    return Q


def _oracle_mixdom2_logp(p, x, y, want_counts=False):
    K, F = p["frag_w"].shape
    chi = mixdom.build_nested_trans(p["lam"], p["mu"], p["t"], p["dlam"], p["dmu"], p["dw"],
                                    p["frag_w"], p["ext"])
    st = mixdom.state_types(K, F)
    subs = np.stack([scipy.linalg.expm(p["t"] * _gtr(p["S"][c], p["pis"][c]))
                     for c in range(p["pis"].shape[0])])
    em = emissions.pair_hmm_emissions_per_class(st, x, y, subs, p["pis"], p["cd"], K, F)
    if not want_counts:
        return dp2d.forward(tkf.safe_log(chi), st, em)[0]
    lp, post, n, _ = dp2d.forward_backward(tkf.safe_log(chi), st, em)
    groups = [0, 0] + [((s - 2) // (5 * F)) * F + ((s - 2) % (5 * F)) % F for s in range(2, len(st))]
    ec = emissions.emission_counts(post, st, groups, x, y, K * F)
    return lp, n, ec


def _torch_score(p, n, ec):
    """sum n log chi(theta) + sum ec log tab(theta), differentiable."""
    from tkfmi import hmm, mixdom as tmix, tables, vjp
    tp = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in p.items()}
    chi = tmix.build_nested_trans(tp["lam"], tp["mu"], tp["t"], tp["dlam"], tp["dmu"], tp["dw"],
                                  tp["frag_w"], tp["ext"])
    subs = vjp.build_class_subs(tp["S"], tp["pis"], tp["t"].reshape(1))[0]
    tab = tables.per_class_tables(subs, tp["pis"], tp["cd"])
    L = (torch.tensor(n) * hmm.safe_log(chi)).sum() + (torch.tensor(ec) * tab).sum()
    L.backward()
    return {k: v.grad.numpy() for k, v in tp.items()}


def _params(rng, K=2, F=1, C=2):
    S = rng.gamma(1.0, 1.0, (C, A, A))
    S = (S + S.transpose(0, 2, 1)) / 2
    return {"lam": 0.04, "mu": 0.05, "t": 0.8, "dlam": rng.uniform(0.03, 0.05, K),
            "dmu": rng.uniform(0.05, 0.07, K), "dw": rng.dirichlet(np.ones(K)),
            "frag_w": rng.dirichlet(np.ones(F), size=K), "ext": rng.uniform(0.5, 0.9, (K, F)),
            "pis": rng.dirichlet(np.ones(A), size=C), "S": S,
            "cd": rng.dirichlet(np.ones(C), size=(K, F))}


def _fd(fn, p, key, idx):
    q1 = {k: np.array(v, dtype=np.float64, copy=True) for k, v in p.items()}
    q2 = {k: np.array(v, dtype=np.float64, copy=True) for k, v in p.items()}
    q1[key][idx] += H
    q2[key][idx] -= H
    return (fn(q1) - fn(q2)) / (2 * H)


def test_mixdom2_score_identity_all_parameters():
    rng = np.random.default_rng(31)
    p = _params(rng)
    x, y = rng.integers(0, A, 9), rng.integers(0, A, 11)
    lp, n, ec = _oracle_mixdom2_logp(p, x, y, want_counts=True)
    grads = _torch_score(p, n, ec)
    fn = lambda q: _oracle_mixdom2_logp(q, x, y)
    checks = [("lam", ()), ("mu", ()), ("t", ()), ("dlam", (1,)), ("dmu", (0,)), ("dw", (0,)),
              ("frag_w", (1, 0)), ("ext", (0, 0)), ("ext", (1, 0)), ("pis", (0, int(x[2]))),
              ("pis", (1, int(y[3]))), ("S", (0, int(x[1]), int(y[1]))), ("S", (1, 3, 7)),
              ("cd", (0, 0, 1)), ("cd", (1, 0, 0))]
    for key, idx in checks:
        want = _fd(fn, p, key, idx)
        got = float(np.asarray(grads[key])[idx])
        assert got == pytest.approx(want, rel=2e-5, abs=2e-6), (key, idx, got, want)


def test_tkf92_score_identity():
    """TKF92 with shared emissions: lambda, mu, t, ext, Q (through expm), pi."""
    from tkfmi import hmm, params, tables, vjp
    rng = np.random.default_rng(32)
    S = rng.gamma(1.0, 1.0, (A, A))
    S = (S + S.T) / 2
    p = {"lam": 0.04, "mu": 0.05, "t": 0.6, "ext": 0.55, "S": S, "pi": rng.dirichlet(np.ones(A))}
    x, y = rng.integers(0, A, 12), rng.integers(0, A, 10)
    st = tkf.TKF_STATE_TYPES

    def logp(q, counts=False):
        T = tkf.tkf92_trans(q["lam"], q["mu"], q["t"], q["ext"])
        sub = scipy.linalg.expm(q["t"] * _gtr(q["S"], q["pi"]))
        em = emissions.pair_hmm_emissions(st, x, y, sub, q["pi"])
        if not counts:
            return dp2d.forward(tkf.safe_log(T), st, em)[0]
        lp, post, n, _ = dp2d.forward_backward(tkf.safe_log(T), st, em)
        return lp, n, emissions.emission_counts(post, st, [0] * 5, x, y, 1)

    _, n, ec = logp(p, counts=True)
    tp = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in p.items()}
    T = params.tkf92_trans(tp["lam"], tp["mu"], tp["t"], tp["ext"])
    sub = vjp.transition_matrix(vjp.build_Q_from_S_pi(tp["S"], tp["pi"]), tp["t"])
    tab = tables.shared_tables(sub, tp["pi"])
    L = (torch.tensor(n) * hmm.safe_log(T)).sum() + (torch.tensor(ec) * tab).sum()
    L.backward()
    for key, idx in [("lam", ()), ("mu", ()), ("t", ()), ("ext", ()), ("pi", (int(x[0]),)),
                     ("pi", (int(y[4]),)), ("S", (int(x[3]), int(y[3]))), ("S", (2, 9))]:
        want = _fd(logp, p, key, idx)
        got = float(np.asarray(tp[key].grad.numpy())[idx])
        assert got == pytest.approx(want, rel=2e-5, abs=2e-6), (key, idx, got, want)


def test_pade13_expm_matches_scipy_and_its_gradient():
    """tkfmi.vjp.expm (degree-13 Pade, scaling and squaring chosen on the
    device) against scipy.linalg.expm on rate matrices x t with 1-norms from
    1e-3 to 2e3 (0 to 9 squarings) and the zero matrix, and its autograd
    gradient against central differences."""
    from tkfmi import vjp
    rng = np.random.default_rng(5)
    mats = [np.zeros((A, A))]
    for scale in (1e-3, 0.1, 1.0, 7.0, 60.0, 2e3):
        S = rng.gamma(1.0, 1.0, (A, A))
        S = (S + S.T) / 2
        Q = _gtr(S, rng.dirichlet(np.ones(A)))
        mats.append(Q * scale / np.abs(Q).sum(0).max())
    X = torch.tensor(np.stack(mats))
    got = vjp.expm(X).numpy()
    for m, g in zip(mats, got):
        ref = scipy.linalg.expm(m)
        np.testing.assert_allclose(g, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
    # gradient of sum(W * expm(A)) along a random direction, A of norm ~3
    Am = mats[3] * 3.0
    W = rng.standard_normal((A, A))
    D = rng.standard_normal((A, A))
    At = torch.tensor(Am, requires_grad=True)
    (vjp.expm(At) * torch.tensor(W)).sum().backward()
    fd = ((scipy.linalg.expm(Am + H * D) - scipy.linalg.expm(Am - H * D)) * W).sum() / (2 * H)
    assert abs(float((At.grad.numpy() * D).sum()) - fd) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("K,F", [(1, 1), (3, 1), (2, 2), (3, 2), (8, 2), (8, 4)])
def test_chi_by_elimination_matches_closed_form(K, F):
    """tkfmi.mixdom.build_nested_trans (reduced matrix of the exploded
    machine, one block solve) against the oracle's closed-form collapse
    (models/mixdom.py:252-423 restated) at 1e-13 relative, exact zeros
    included, batched over t; and the exploded machine's rows sum to one
    (End excepted, which has no out-transitions)."""
    from tkfmi import mixdom as tm
    rng = np.random.default_rng(100 + 10 * K + F)
    dl, dm = rng.uniform(0.02, 0.06, K), rng.uniform(0.05, 0.08, K)
    dw, fw = rng.dirichlet(np.ones(K)), rng.dirichlet(np.ones(F), size=K)
    ext = rng.uniform(0.5, 0.9, (K, F, 1)) * rng.dirichlet(np.ones(F), size=(K, F))
    ts = np.array([0.05, 0.3, 1.0, 2.0])
    got = tm.build_nested_trans(0.04, 0.05, torch.tensor(ts), dl, dm, dw, fw, ext).numpy()
    for i, t in enumerate(ts):
        want = mixdom.build_nested_trans(0.04, 0.05, t, dl, dm, dw, fw, ext)
        nz = want != 0
        assert (got[i][~nz] == 0).all()
        assert (np.abs(got[i][nz] - want[nz]) / np.abs(want[nz])).max() <= 1e-13
    T = tm.exploded_trans(0.04, 0.05, 0.7, dl, dm, dw, fw, ext).numpy()
    rows = T.sum(axis=1)
    assert rows[1] == 0.0
    np.testing.assert_allclose(np.delete(rows, 1), 1.0, rtol=1e-13)
    # MixDom1 (K,F) extension rates go on the diagonal (models/mixdom.py:284-287)
    e1 = rng.uniform(0.3, 0.8, (K, F))
    np.testing.assert_allclose(tm.build_nested_trans(0.04, 0.05, 0.5, dl, dm, dw, fw, e1).numpy(),
                               mixdom.build_nested_trans(0.04, 0.05, 0.5, dl, dm, dw, fw, e1),
                               rtol=1e-13, atol=1e-300)
