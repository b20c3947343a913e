"""Device parameter builder (tkfmi_build_params_batched) vs the torch
builders (tkfmi.params / tkfmi.mixdom / tkfmi.tables, torch.linalg.matrix_exp)
for every model and emission mode.  Log transitions agree to 1e-12; log
emission tables to 1e-10 (absolute on log values, relative on large ones):
the device forms P(t) from the eigendecomposition of the symmetrised rate
matrix, torch by Pade scaling-and-squaring, and the two FP64 routes differ by
~1e-12 relative on the smallest entries at long t."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
A = 20


def _gtr(rng, n):
    Qs, pis, Ss = [], [], []
    for _ in range(n):
        S = rng.gamma(1.0, 1.0, (A, A))
        S = (S + S.T) / 2
        pi = rng.dirichlet(np.ones(A))
        Q = S * pi[None, :]
        np.fill_diagonal(Q, 0.0)
        np.fill_diagonal(Q, -Q.sum(axis=1))
        Qs.append(Q); pis.append(pi); Ss.append(S)
    return np.stack(Qs), np.stack(pis), np.stack(Ss)


def _cmp(got, want, tol=1e-12):
    got, want = got.cpu().numpy(), want.cpu().numpy()
    err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
    assert err.max() <= tol, float(err.max())


TS = np.array([0.01, 0.05, 0.3, 0.9, 1.7, 2.0, 4.0])


@pytest.mark.parametrize("model", ["tkf91", "tkf91_cond", "tkf92"])
@pytest.mark.parametrize("floor", [False, True])
def test_tkf_builder(model, floor):
    from tkfmi import build, hmm, params, tables, vjp
    rng = np.random.default_rng(1)
    Qs, pis, _ = _gtr(rng, 1)
    b = build.DeviceBuilder(model, "shared", Qs, pis, lam=0.04, mu=0.05, ext=0.6, log_floor=floor)
    lt, tab, t = b.build(TS)
    dev = lt.device
    tt = torch.as_tensor(TS, device=dev)
    tau = {"tkf91": lambda: params.tkf91_trans(0.04, 0.05, tt),
           "tkf91_cond": lambda: params.tkf91_trans_cond(0.04, 0.05, tt),
           "tkf92": lambda: params.tkf92_trans(0.04, 0.05, tt, 0.6)}[model]()
    want = torch.log(torch.clamp(tau, min=1e-30)) if floor else hmm.safe_log(tau)
    _cmp(lt, want)
    sub = vjp.transition_matrix(torch.as_tensor(Qs[0], device=dev), tt)
    _cmp(tab, tables.shared_tables(sub, torch.as_tensor(pis[0], device=dev).expand(sub.shape[:-1])), 1e-10)


@pytest.mark.parametrize("K,F,C", [(3, 1, 1), (3, 1, 2), (2, 2, 3), (1, 3, 1)])
def test_mixdom_builder_per_class_and_per_domain(K, F, C):
    from tkfmi import build, hmm, mixdom, tables, vjp
    rng = np.random.default_rng(K * 10 + F + C)
    mix = dict(main_ins=0.04, main_del=0.05, dom_ins=rng.uniform(0.02, 0.06, K),
               dom_del=rng.uniform(0.05, 0.08, K), dom_w=rng.dirichlet(np.ones(K)),
               frag_w=rng.dirichlet(np.ones(F), size=K),
               ext=rng.uniform(0.5, 0.9, (K, F, 1)) * rng.dirichlet(np.ones(F), size=(K, F)))
    Qs, pis, Ss = _gtr(rng, C)
    cd = rng.dirichlet(np.ones(C), size=(K, F))
    cd[0, 0, 0] = 0.0
    cd[0, 0] /= cd[0, 0].sum() if cd[0, 0].sum() > 0 else 1.0
    b = build.DeviceBuilder("mixdom", "per_class", Qs, pis, mix=mix, class_dist=cd, log_floor=True)
    lt, tab, _ = b.build(TS)
    dev = lt.device
    tt = torch.as_tensor(TS, device=dev)
    chi = mixdom.build_nested_trans(0.04, 0.05, tt, mix["dom_ins"], mix["dom_del"], mix["dom_w"],
                                    mix["frag_w"], mix["ext"])
    _cmp(lt, torch.log(torch.clamp(chi, min=1e-30)))
    subs = vjp.build_class_subs(torch.as_tensor(Ss, device=dev), torch.as_tensor(pis, device=dev), tt)
    _cmp(tab, tables.per_class_tables(subs, torch.as_tensor(pis, device=dev), torch.as_tensor(cd, device=dev)),
         1e-10)
    # safe_log transitions and per-domain tables (C = K components)
    Qd, pd_, _ = _gtr(rng, K)
    b2 = build.DeviceBuilder("mixdom", "per_domain", Qd, pd_, mix=mix, log_floor=False)
    lt2, tab2, _ = b2.build(TS)
    _cmp(lt2, hmm.safe_log(chi))
    subs_d = vjp.transition_matrix(torch.as_tensor(Qd, device=dev)[None], tt[:, None])
    _cmp(tab2, tables.per_domain_tables(subs_d, torch.as_tensor(pd_, device=dev).expand(subs_d.shape[:-1])),
         1e-10)


def test_builder_rejects_nonreversible_rate_matrix():
    from tkfmi import build
    rng = np.random.default_rng(3)
    Q = rng.uniform(0.1, 1.0, (A, A))
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    with pytest.raises(ValueError):
        build.DeviceBuilder("tkf92", "shared", Q, np.full(A, 1.0 / A), lam=0.04, mu=0.05, ext=0.5)
