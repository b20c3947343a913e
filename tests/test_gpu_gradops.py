"""GPU: the native builders of the gradient path (tkfmi.gradops, libtkfmi's
tkfmi_tkf_trans_jac, tkfmi_chi_forward / _backward, tkfmi_expm_batched /
_vjp_batched) against their torch forms on the CPU -- which the CPU suite
pins to the oracle -- and against scipy / central differences.

* TKF91, conditioned TKF91, TKF92 matrices and d/d(lambda, mu, t, ext),
  including the lambda ~ mu limit (|1 - lambda/mu| < 1e-4) and t -> 0;
* MixDom chi and the gradient through it for several (K, F);
* exp(A) against scipy.linalg.expm (norms needing 0..10 squarings, a
  non-normal A) and its adjoint against autograd of the CPU Pade form and
  central differences.
"""
import numpy as np
import pytest
import scipy.linalg
import torch

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _tkf_inputs(rng, n):
    lam = rng.uniform(0.01, 0.2, n)
    mu = lam / rng.uniform(0.5, 0.98, n)
    mu[:4] = lam[:4] * (1 + np.array([0.0, 3e-5, -5e-5, 2e-4]))  # the lambda ~ mu switch
    t = rng.uniform(0.01, 3.0, n)
    t[4:6] = [1e-12, 0.0]                                          # t -> 0
    ext = rng.uniform(0.0, 0.95, n)
    return lam, mu, t, ext


@pytest.mark.parametrize("kind", ["tkf91", "tkf91_cond", "tkf92"])
def test_tkf_trans_and_jacobian_match_torch(kind):
    from tkfmi import params
    rng = np.random.default_rng(7)
    lam, mu, t, ext = _tkf_inputs(rng, 64)
    fn = {"tkf91": params.tkf91_trans, "tkf91_cond": params.tkf91_trans_cond,
          "tkf92": params.tkf92_trans}[kind]
    G = rng.standard_normal((64, 5, 5))
    res = {}
    for dev in ("cpu", DEV):
        xs = [torch.tensor(v, dtype=torch.float64, device=dev, requires_grad=True) for v in (lam, mu, t, ext)]
        T = fn(*xs[:3]) if kind != "tkf92" else fn(*xs)
        (T * torch.tensor(G, device=dev)).sum().backward()
        grads = [x.grad.cpu().numpy() if x.grad is not None else np.zeros(64) for x in xs]
        res[str(dev)] = (T.detach().cpu().numpy(), grads)
    (Tc, gc), (Tg, gg) = res["cpu"], res[str(DEV)]
    # Just past the switch (|1 - lambda/mu| = 2e-4, pair 3) the general
    # closed forms cancel: both sides' last-bit differences in exp grow by
    # ~1 / |1 - lambda/mu| (T) and its square (derivatives).
    ok = np.ones(64, bool)
    ok[3] = False
    assert _rel(Tg[ok], Tc[ok]) <= 1e-14
    assert _rel(Tg[3], Tc[3]) <= 1e-11
    for k in range(3 if kind != "tkf92" else 4):
        assert _rel(gg[k][ok], gc[k][ok]) <= 1e-12, k
        assert _rel(gg[k][3], gc[k][3]) <= 1e-6, k


def test_tkf_trans_broadcast_shapes():
    """Shared (scalar) rates with per-pair t, and per-domain rates with
    (B, 1) t, reduce their gradients to the inputs' shapes."""
    from tkfmi import params
    rng = np.random.default_rng(8)
    t = rng.uniform(0.1, 2.0, 6)
    dl, dm = rng.uniform(0.02, 0.05, 3), rng.uniform(0.06, 0.08, 3)
    out = {}
    for dev in ("cpu", DEV):
        tt = torch.tensor(t, device=dev, requires_grad=True)
        lam = torch.tensor(0.04, dtype=torch.float64, device=dev, requires_grad=True)
        mu = torch.tensor(0.05, dtype=torch.float64, device=dev, requires_grad=True)
        a = torch.tensor(dl, device=dev, requires_grad=True)
        b = torch.tensor(dm, device=dev, requires_grad=True)
        T0 = params.tkf91_trans(lam, mu, tt)
        T1 = params.tkf91_trans(a, b, tt[:, None])
        assert T0.shape == (6, 5, 5) and T1.shape == (6, 3, 5, 5)
        (T0.sin().sum() + T1.cos().sum()).backward()
        out[str(dev)] = [x.grad.cpu().numpy() for x in (lam, mu, tt, a, b)]
    for gc, gg in zip(out["cpu"], out[str(DEV)]):
        assert gg.shape == gc.shape
        assert _rel(gg, gc) <= 1e-12


@pytest.mark.parametrize("K,F", [(1, 1), (3, 1), (2, 2), (3, 2)])
def test_chi_and_gradient_match_torch(K, F):
    from tkfmi import mixdom
    rng = np.random.default_rng(10 + K * F)
    B = 5
    p = dict(lam=0.04, mu=0.05, t=rng.uniform(0.05, 2.0, B), dl=rng.uniform(0.03, 0.05, K),
             dm=rng.uniform(0.05, 0.07, K), dw=rng.dirichlet(np.ones(K)),
             fw=rng.dirichlet(np.ones(F), size=K), ext=rng.uniform(0.5, 0.9, (K, F)))
    res = {}
    for dev in ("cpu", DEV):
        xs = {k: torch.tensor(v, dtype=torch.float64, device=dev, requires_grad=True) for k, v in p.items()}
        chi = mixdom.build_nested_trans(xs["lam"], xs["mu"], xs["t"], xs["dl"], xs["dm"], xs["dw"],
                                        xs["fw"], xs["ext"])
        G = torch.tensor(np.random.default_rng(1).standard_normal(tuple(chi.shape)), device=dev)
        (chi * G).sum().backward()
        res[str(dev)] = (chi.detach().cpu().numpy(), {k: v.grad.cpu().numpy() for k, v in xs.items()})
    (cc, gc), (cg, gg) = res["cpu"], res[str(DEV)]
    assert cg.shape == cc.shape == (B, 2 + 5 * K * F, 2 + 5 * K * F)
    assert _rel(cg, cc) <= 1e-13
    for k in p:
        assert _rel(gg[k], gc[k]) <= 1e-11, k


def _expm_cases(rng):
    A = []
    for scale in (1e-3, 0.3, 2.0, 9.0, 60.0, 2000.0):  # degrees 3..13, 0..10 squarings
        Q = rng.gamma(1.0, 1.0, (20, 20))
        np.fill_diagonal(Q, 0.0)
        np.fill_diagonal(Q, -Q.sum(1))
        A.append(Q * scale / np.abs(Q).sum(0).max())
    N = np.triu(rng.standard_normal((20, 20)), 1) * 3.0 - np.eye(20)  # non-normal
    A.append(N)
    return np.stack(A)


def test_expm_matches_scipy():
    from tkfmi import gradops
    A = _expm_cases(np.random.default_rng(3))
    got = gradops.expm(torch.tensor(A, device=DEV)).cpu().numpy()
    for a, g in zip(A, got):
        want = scipy.linalg.expm(a)
        assert _rel(g, want) <= 1e-12


def test_expm_adjoint():
    """The adjoint against autograd of the CPU degree-13 Pade form, and
    <adjoint, E> against the complex-step derivative Im <G, exp(A + ihE)> / h
    (scipy's expm, a rational function of A at fixed scaling: no
    cancellation, so h = 1e-30 is exact to rounding)."""
    from tkfmi import gradops, vjp
    rng = np.random.default_rng(4)
    A = _expm_cases(rng)
    G = rng.standard_normal(A.shape)
    Ag = torch.tensor(A, device=DEV, requires_grad=True)
    (gradops.expm(Ag) * torch.tensor(G, device=DEV)).sum().backward()
    got = Ag.grad.cpu().numpy()
    Ac = torch.tensor(A, requires_grad=True)
    (vjp.expm(Ac) * torch.tensor(G)).sum().backward()
    assert _rel(got, Ac.grad.numpy()) <= 1e-10
    E = rng.standard_normal(A.shape)
    for a, g, e, adj in zip(A, G, E, got):
        h = 1e-30
        cs = np.sum(g * scipy.linalg.expm(a + 1j * h * e)).imag / h
        assert np.sum(adj * e) == pytest.approx(cs, rel=1e-9, abs=1e-12 * np.abs(adj).sum() * np.abs(e).max())
    # a zero cotangent gives a zero gradient
    Z = torch.tensor(A[:2], device=DEV, requires_grad=True)
    (gradops.expm(Z) * 0.0).sum().backward()
    assert float(Z.grad.abs().max()) == 0.0
