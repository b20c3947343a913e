"""GPU gradients (tkfmi.vjp): the autograd.Function whose backward is the
hand-written adjoint DP, composed with the torch builders.

* small pairs: every parameter against central finite differences
  (h = 1e-6) of the FP64 oracle logP (tolerance 2e-5 relative, 2e-6
  absolute -- the finite-difference error, not the kernel's);
* the 200-pair MixDom2 K=3 F=1 C=2 minibatch (lengths LogNormal median 200):
  directional derivatives of the summed logP along random directions per
  parameter group against central differences of the kernel's own logP.
"""
import numpy as np
import pytest
import scipy.linalg
import torch

pytestmark = pytest.mark.gpu
A = 20
H = 1e-6


def _gtr(S_exch, pi):
    Q = S_exch * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    return Q


def _mix2_params(rng, K, F, C):
    S = rng.gamma(1.0, 1.0, (C, A, A))
    S = (S + S.transpose(0, 2, 1)) / 2
    return {"lam": 0.04, "mu": 0.05, "dlam": rng.uniform(0.03, 0.05, K),
            "dmu": rng.uniform(0.05, 0.07, K), "dw": rng.dirichlet(np.ones(K)),
            "frag_w": rng.dirichlet(np.ones(F), size=K), "ext": rng.uniform(0.5, 0.9, (K, F)),
            "pis": rng.dirichlet(np.ones(A), size=C), "S": S,
            "cd": rng.dirichlet(np.ones(C), size=(K, F))}


def _oracle_sum_logp(p, ts, xs, ys):
    from oracle import dp2d, emissions, mixdom, tkf
    K, F = p["frag_w"].shape
    st = mixdom.state_types(K, F)
    tot = 0.0
    for t, x, y in zip(ts, xs, ys):
        chi = mixdom.build_nested_trans(p["lam"], p["mu"], t, p["dlam"], p["dmu"], p["dw"],
                                        p["frag_w"], p["ext"])
        subs = np.stack([scipy.linalg.expm(t * _gtr(p["S"][c], p["pis"][c]))
                         for c in range(p["pis"].shape[0])])
        em = emissions.pair_hmm_emissions_per_class(st, x, y, subs, p["pis"], p["cd"], K, F)
        tot += dp2d.forward(tkf.safe_log(chi), st, em)[0]
    return tot


def _gpu_sum_logp(p, ts, xs, ys, grad=False, kernel=None, checkpoint=False):
    from tkfmi import vjp
    dev = torch.device("cuda", 0)
    tp = {k: torch.tensor(v, dtype=torch.float64, device=dev, requires_grad=grad)
          for k, v in p.items()}
    tt = torch.tensor(ts, dtype=torch.float64, device=dev, requires_grad=grad)
    lp = vjp.mixdom2_log_prob_batch(tp["lam"], tp["mu"], tt, tp["dlam"], tp["dmu"], tp["dw"],
                                    tp["frag_w"], tp["ext"], tp["pis"], tp["S"], tp["cd"], xs, ys,
                                    kernel=kernel, checkpoint=checkpoint)
    tot = lp.sum()
    if not grad:
        return float(tot)
    tot.backward()
    g = {k: v.grad.cpu().numpy() for k, v in tp.items()}
    g["ts"] = tt.grad.cpu().numpy()
    return float(tot.detach()), g


def _shift(p, key, idx, d):
    q = {k: np.array(v, dtype=np.float64, copy=True) for k, v in p.items()}
    q[key][idx] += d
    return q


@pytest.mark.parametrize("kernel", ["chain", "class", "generic"])
def test_mixdom2_gradients_match_oracle_fd(kernel):
    rng = np.random.default_rng(41)
    p = _mix2_params(rng, 3, 1, 2)
    lens = [(0, 3), (7, 0), (25, 31), (40, 38), (64, 70)]
    xs = [rng.integers(0, A, a) for a, _ in lens]
    ys = [rng.integers(0, A, b) for _, b in lens]
    ts = rng.uniform(0.1, 1.5, len(lens))
    tot, g = _gpu_sum_logp(p, ts, xs, ys, grad=True, kernel=kernel)
    assert tot == pytest.approx(_oracle_sum_logp(p, ts, xs, ys), rel=1e-10)
    checks = [("lam", ()), ("mu", ()), ("dlam", (2,)), ("dmu", (0,)), ("dw", (1,)),
              ("frag_w", (0, 0)), ("ext", (1, 0)), ("pis", (0, 5)), ("pis", (1, int(xs[3][0]))),
              ("S", (0, int(xs[2][1]), int(ys[2][1]))), ("S", (1, 4, 11)), ("cd", (2, 0, 1))]
    for key, idx in checks:
        want = (_oracle_sum_logp(_shift(p, key, idx, H), ts, xs, ys)
                - _oracle_sum_logp(_shift(p, key, idx, -H), ts, xs, ys)) / (2 * H)
        got = float(np.asarray(g[key])[idx])
        print(key, idx, got, want, flush=True)
        assert got == pytest.approx(want, rel=2e-5, abs=2e-6), (key, idx)
    for b in (2, 4):  # per-pair t
        tp_, tm_ = ts.copy(), ts.copy()
        tp_[b] += H
        tm_[b] -= H
        want = (_oracle_sum_logp(p, tp_, xs, ys) - _oracle_sum_logp(p, tm_, xs, ys)) / (2 * H)
        assert g["ts"][b] == pytest.approx(want, rel=2e-5, abs=2e-6)


def test_minibatch_200_directional_derivatives():
    """The 200-pair K=3 F=1 C=2 minibatch: grad . v == central difference of
    the summed logP along v, for a random direction v per parameter group."""
    from tkfmi import synth
    rng = np.random.default_rng(3)
    w = synth.make_mixdom2(seed=3, B=200, K=3, F=1, C=2)
    p = _mix2_params(rng, 3, 1, 2)
    ts = w.ts
    tot, g = _gpu_sum_logp(p, ts, w.xs, w.ys, grad=True)
    assert np.isfinite(tot)
    for key in ("lam", "mu", "dlam", "dmu", "dw", "ext", "pis", "S", "cd"):
        v = rng.standard_normal(np.shape(p[key]))
        if key == "S":
            v = (v + np.swapaxes(v, -1, -2)) / 2
        h = 1e-6 * max(1.0, float(np.abs(np.asarray(p[key])).max()))
        qp = {k: np.array(x, dtype=np.float64, copy=True) for k, x in p.items()}
        qm = {k: np.array(x, dtype=np.float64, copy=True) for k, x in p.items()}
        qp[key] = qp[key] + h * v
        qm[key] = qm[key] - h * v
        fd = (_gpu_sum_logp(qp, ts, w.xs, w.ys) - _gpu_sum_logp(qm, ts, w.xs, w.ys)) / (2 * h)
        dd = float((np.asarray(g[key]) * v).sum())
        print(key, dd, fd, flush=True)
        assert dd == pytest.approx(fd, rel=1e-5, abs=1e-4 * max(1.0, abs(tot) * 1e-6)), key


def test_graphed_builders_equal_eager():
    """graph=True (the builders' forward and backward replayed from HIP
    graphs) on the 200-pair minibatch: log P and every gradient equal the
    eager path's over several steps with new parameter values, with eager
    work between the replays; a second call before the first's backward runs
    eagerly, and both gradients stay right."""
    from tkfmi import synth, vjp
    rng = np.random.default_rng(5)
    w = synth.make_mixdom2(seed=3, B=200, K=3, F=1, C=2)
    dev = torch.device("cuda", 0)
    keys = ("lam", "mu", "dlam", "dmu", "dw", "frag_w", "ext", "pis", "S", "cd")

    def run(p, graph, xs=w.xs, ys=w.ys):
        tp = {k: torch.tensor(p[k], dtype=torch.float64, device=dev, requires_grad=True) for k in keys}
        tt = torch.tensor(w.ts, dtype=torch.float64, device=dev, requires_grad=True)
        lp = vjp.mixdom2_log_prob_batch(tp["lam"], tp["mu"], tt, tp["dlam"], tp["dmu"], tp["dw"],
                                        tp["frag_w"], tp["ext"], tp["pis"], tp["S"], tp["cd"], xs, ys,
                                        graph=graph)
        return lp, tp, tt

    def grads(lp, tp, tt):
        lp.sum().backward()
        g = {k: v.grad.cpu().numpy() for k, v in tp.items()}
        g["ts"] = tt.grad.cpu().numpy()
        return lp.detach().cpu().numpy(), g

    def same(a, b):
        la, ga = a
        lb, gb = b
        assert np.max(np.abs(la - lb) / np.abs(lb)) <= 1e-12
        for k in ga:
            scale = max(np.abs(gb[k]).max(), 1e-300)
            assert np.max(np.abs(ga[k] - gb[k])) <= 1e-10 * scale, k

    for step in range(4):
        p = _mix2_params(rng, 3, 1, 2)
        eager = grads(*run(p, False))
        graphed = grads(*run(p, True))
        same(graphed, eager)
    # two calls before one backward: the second runs eagerly
    p = _mix2_params(rng, 3, 1, 2)
    lp1, tp1, tt1 = run(p, True)
    lp2, tp2, tt2 = run(p, True, xs=w.ys, ys=w.xs)
    (lp1.sum() + lp2.sum()).backward()
    got1 = (lp1.detach().cpu().numpy(), {**{k: v.grad.cpu().numpy() for k, v in tp1.items()},
                                        "ts": tt1.grad.cpu().numpy()})
    got2 = (lp2.detach().cpu().numpy(), {**{k: v.grad.cpu().numpy() for k, v in tp2.items()},
                                        "ts": tt2.grad.cpu().numpy()})
    same(got1, grads(*run(p, False)))
    same(got2, grads(*run(p, False, xs=w.ys, ys=w.xs)))


def test_tkf_reference_surfaces_gradients():
    """tkf91_log_prob / tkf92_log_prob (single pair, the reference's
    signatures): gradients w.r.t. lambda, mu, t, ext, Q, pi vs oracle FD."""
    from oracle import dp2d, emissions, tkf
    from tkfmi import vjp
    rng = np.random.default_rng(42)
    S = rng.gamma(1.0, 1.0, (A, A))
    S = (S + S.T) / 2
    pi = rng.dirichlet(np.ones(A))
    Q = _gtr(S, pi)
    x, y = rng.integers(0, A, 33), rng.integers(0, A, 29)
    dev = torch.device("cuda", 0)

    def oracle(lam, mu, t, ext, Qm, pim):
        T = tkf.tkf91_trans(lam, mu, t) if ext is None else tkf.tkf92_trans(lam, mu, t, ext)
        em = emissions.pair_hmm_emissions(tkf.TKF_STATE_TYPES, x, y, scipy.linalg.expm(t * Qm), pim)
        return dp2d.forward(tkf.safe_log(T), tkf.TKF_STATE_TYPES, em)[0]

    for ext in (None, 0.6):
        th = {k: torch.tensor(v, dtype=torch.float64, device=dev, requires_grad=True)
              for k, v in dict(lam=0.04, mu=0.05, t=0.7, Q=Q, pi=pi).items()}
        if ext is None:
            lp = vjp.tkf91_log_prob(th["lam"], th["mu"], th["t"], th["Q"], th["pi"], x, y)
        else:
            th["ext"] = torch.tensor(ext, dtype=torch.float64, device=dev, requires_grad=True)
            lp = vjp.tkf92_log_prob(th["lam"], th["mu"], th["t"], th["ext"], th["Q"], th["pi"], x, y)
        lp.backward()
        base = dict(lam=0.04, mu=0.05, t=0.7, ext=ext, Qm=Q, pim=pi)
        assert float(lp) == pytest.approx(oracle(**base), rel=1e-10)
        for name, key, idx in [("lam", "lam", ()), ("mu", "mu", ()), ("t", "t", ()),
                               ("ext", "ext", ()), ("Q", "Qm", (int(x[0]), int(y[0]))),
                               ("pi", "pim", (int(x[5]),))]:
            if name == "ext" and ext is None:
                continue
            bp = {k: (np.array(v, copy=True) if v is not None else None) for k, v in base.items()}
            bm = {k: (np.array(v, copy=True) if v is not None else None) for k, v in base.items()}
            bp[key][idx] += H
            bm[key][idx] -= H
            bp = {k: (float(v) if (v is not None and np.ndim(v) == 0) else v) for k, v in bp.items()}
            bm = {k: (float(v) if (v is not None and np.ndim(v) == 0) else v) for k, v in bm.items()}
            want = (oracle(**bp) - oracle(**bm)) / (2 * H)
            got = float(np.asarray(th[name].grad.cpu().numpy())[idx])
            assert got == pytest.approx(want, rel=2e-5, abs=2e-6), (ext, name)


def test_wide_gradients_with_checkpointing_equal_full_lattice():
    """MixDom2 K=4 F=2 C=2 (42 states, the wide kernel): log P and every
    parameter gradient are the same with antidiagonal checkpointing as with
    the full lattice (the backward recomputes the forward bit for bit; the
    emission-count atomics leave summation-order rounding)."""
    rng = np.random.default_rng(12)
    p = _mix2_params(rng, 4, 2, 2)
    xs = [rng.integers(0, A, n) for n in (70, 5, 130)]
    ys = [rng.integers(0, A, n) for n in (64, 9, 121)]
    ts = [0.3, 1.1, 0.7]
    lp0, g0 = _gpu_sum_logp(p, ts, xs, ys, grad=True)
    lp1, g1 = _gpu_sum_logp(p, ts, xs, ys, grad=True, checkpoint=True)
    assert lp0 == lp1
    for k in g0:
        np.testing.assert_allclose(g1[k], g0[k], rtol=1e-11, atol=1e-13 * max(1.0, float(np.abs(g0[k]).max())),
                                   err_msg=k)
