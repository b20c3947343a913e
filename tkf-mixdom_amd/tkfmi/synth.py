"""Seeded synthetic workloads for tests and benchmarks.

Parameters follow the benchmark workload definitions:

* ``c0``  TKF91, 64 pairs, ancestor length U{50..150}, lambda=0.04, mu=0.05,
          t=0.5; random reversible Q normalised to one expected substitution
          per unit time (exchangeabilities ~ Gamma(1,1), pi ~ Dirichlet(1)).
* ``c1``  MixDom2 K=3, F=1, C=1 (17 states), 1024 pairs, ancestor length
          round(LogNormal(ln 200, 0.5)) clipped to [30, 800], t ~ U(0.05, 2).
* ``c3``  MixDom2 K=3, F=1, C=2, 200 pairs, LogNormal median 200.
* ``c4``  serving mix: 50/50 TKF92 (5 states) and MixDom2 K=3 (17 states),
          lengths LogNormal median 250 clipped to [20, 2000].

Descendants are drawn by backward-guided forward sampling through the pair
HMM conditioned on the ancestor (the scheme of simulate/pair_hmm.py:39-222):
the structural backward probabilities B[a, s] (probability of consuming the
remaining ancestor suffix from state s, with insert emissions summed out) are
computed per ancestor position, and each step samples the next state with
weight T[s, s'] * (pi[x_a] B[a+1, s'] if s' consumes x else B[a, s']).
"""
from dataclasses import dataclass, field

import numpy as np
import torch

from . import mixdom, params, tables

A = 20


def random_Q_pi(rng, A_=A, normalise=True):
    pi = rng.dirichlet(np.ones(A_))
    Sx = np.zeros((A_, A_))
    iu = np.triu_indices(A_, 1)
    Sx[iu] = rng.gamma(1.0, 1.0, size=len(iu[0]))
    Sx = Sx + Sx.T
    Q = Sx * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    if normalise:
        rate = -np.sum(pi * np.diag(Q))
        Q = Q / rate
        Sx = Sx / rate
    return Q, pi, Sx


def expm_batch(Q, ts):
    """P(t) = expm(t Q) for a vector of t (torch matrix_exp, FP64, CPU)."""
    Qt = torch.as_tensor(Q, dtype=torch.float64)
    tt = torch.as_tensor(np.asarray(ts, dtype=np.float64))
    return torch.linalg.matrix_exp(tt[:, None, None] * Qt[None]).numpy()


def _structural_backward(T, types, x, pi_x_marg):
    """B[a, :] for a = L..0, each row scaled to max 1; returns (B, logscale)."""
    N = T.shape[0]
    L = len(x)
    cx = np.isin(types, (params.M, params.D))
    ity = types == params.I
    e_idx = int(np.argmax(types == params.E))
    Ai = T * ity[None, :]
    Minv = np.linalg.inv(np.eye(N) - Ai)
    TX = T * cx[None, :]
    B = np.zeros((L + 1, N))
    ls = np.zeros(L + 1)
    c = T[:, e_idx].copy()
    v = Minv @ c
    mx = v.max() if v.max() > 0 else 1.0
    B[L] = v / mx
    ls[L] = np.log(mx)
    for a in range(L - 1, -1, -1):
        c = pi_x_marg[a] * (TX @ B[a + 1])
        v = Minv @ c
        mx = v.max() if v.max() > 0 else 1.0
        B[a] = v / mx
        ls[a] = ls[a + 1] + np.log(mx)
    return B, ls


def sample_descendant(rng, x, T, types, emit_y_match, emit_y_ins, pi_x_marg, path=None):
    """Backward-guided conditional sampling of y given x.

    emit_y_match(state, a_char) -> distribution over y for a match state;
    emit_y_ins(state) -> distribution over y for an insert state;
    pi_x_marg[a] = marginal emission weight of consuming x[a].  When ``path``
    is a list, the column type (M/I/D) of every step is appended to it: the
    alignment the pair was generated from."""
    types = np.asarray(types)
    N = T.shape[0]
    L = len(x)
    B, ls = _structural_backward(T, types, x, pi_x_marg)
    cx = np.isin(types, (params.M, params.D))
    ity = types == params.I
    e_idx = int(np.argmax(types == params.E))
    y = []
    a = 0
    s = 0
    for _ in range(20 * (L + 10)):
        w = np.zeros(N)
        if a < L:
            w[cx] = T[s, cx] * pi_x_marg[a] * B[a + 1, cx] * np.exp(ls[a + 1] - ls[a])
        w[ity] = T[s, ity] * B[a, ity]
        if a == L:
            w[e_idx] = T[s, e_idx] / max(np.exp(ls[a]), 1e-300)
        tot = w.sum()
        if not tot > 0:
            break
        s2 = int(np.searchsorted(np.cumsum(w), rng.random() * tot, side="right"))
        s2 = min(s2, N - 1)
        if s2 == e_idx:
            break
        ty = types[s2]
        if path is not None:
            path.append(int(ty))
        if ty == params.M:
            p = emit_y_match(s2, x[a])
            y.append(int(np.searchsorted(np.cumsum(p), rng.random() * p.sum(), side="right")))
            a += 1
        elif ty == params.D:
            a += 1
        elif ty == params.I:
            p = emit_y_ins(s2)
            y.append(int(np.searchsorted(np.cumsum(p), rng.random() * p.sum(), side="right")))
        s = s2
    return np.minimum(np.array(y, dtype=np.int64), A - 1)


@dataclass
class Workload:
    name: str
    S: int
    state_types: list
    state_group: list
    xs: list
    ys: list
    ts: np.ndarray
    log_trans: torch.Tensor         # (B,S,S) or (S,S), CPU float64
    log_tab: torch.Tensor           # (B,G,TAB) or (G,TAB), CPU float64
    params: dict = field(default_factory=dict)

    @property
    def B(self):
        return len(self.xs)

    def cell_states(self):
        return int(sum((len(x) + 1) * (len(y) + 1) for x, y in zip(self.xs, self.ys)) * self.S)


def _lognormal_lengths(rng, n, median, sigma, lo, hi):
    return np.clip(np.round(np.exp(np.log(median) + sigma * rng.standard_normal(n))), lo, hi).astype(int)


def make_tkf(seed=0, B=64, lengths=None, lam=0.04, mu=0.05, t=0.5, ext=None, name="c0"):
    """TKF91 (ext None) or TKF92 workload with one shared t."""
    rng = np.random.default_rng(seed)
    Q, pi, _ = random_Q_pi(rng)
    Lx = rng.integers(50, 151, size=B) if lengths is None else np.asarray(lengths)
    P = expm_batch(Q, [t])[0]
    T = (params.tkf91_trans(lam, mu, t) if ext is None else params.tkf92_trans(lam, mu, t, ext)).numpy()
    types = np.array(params.TKF_STATE_TYPES)
    xs, ys = [], []
    for b in range(B):
        x = rng.choice(A, size=int(Lx[b]), p=pi)
        y = sample_descendant(rng, x, T, types, lambda s, a: P[a], lambda s: pi, pi[x])
        xs.append(x.astype(np.int64))
        ys.append(y)
    log_trans = torch.log(torch.clamp(torch.as_tensor(T), min=1e-30))
    log_tab = tables.shared_tables(torch.as_tensor(P), torch.as_tensor(pi))
    return Workload(name, 5, list(params.TKF_STATE_TYPES), [0] * 5, xs, ys,
                    np.full(B, t), log_trans, log_tab,
                    dict(lam=lam, mu=mu, t=t, ext=ext, Q=Q, pi=pi, sub=P))


def mixdom2_params(rng, K, F, C, lam=0.04, mu=0.05, normalise_Q=True):
    dom_w = rng.standard_normal(K)
    dom_w = np.exp(dom_w - dom_w.max())
    dom_w /= dom_w.sum()
    fw = rng.standard_normal((K, F))
    fw = np.exp(fw - fw.max(axis=1, keepdims=True))
    fw /= fw.sum(axis=1, keepdims=True)
    if F == 1:
        ext = rng.uniform(0.5, 0.95, size=(K, 1, 1))
    else:
        rows = rng.standard_normal((K, F, F))
        rows = np.exp(rows - rows.max(axis=2, keepdims=True))
        rows /= rows.sum(axis=2, keepdims=True)
        ext = rng.uniform(0.5, 0.95, size=(K, F, 1)) * rows
    cls = [random_Q_pi(rng, normalise=normalise_Q) for _ in range(C)]
    cd = rng.standard_normal((K, F, C))
    cd = np.exp(cd - cd.max(axis=2, keepdims=True))
    cd /= cd.sum(axis=2, keepdims=True)
    return dict(main_lam=lam, main_mu=mu, dom_lam=np.full(K, lam), dom_mu=np.full(K, mu),
                dom_w=dom_w, frag_w=fw, ext=ext,
                class_Q=np.stack([c[0] for c in cls]), class_pi=np.stack([c[1] for c in cls]),
                class_S=np.stack([c[2] for c in cls]), class_dist=cd, K=K, F=F, C=C)


def make_mixdom2(seed=1, B=1024, K=3, F=1, C=1, median=200, sigma=0.5, lo=30, hi=800,
                 t_range=(0.05, 2.0), lengths=None, name="c1"):
    rng = np.random.default_rng(seed)
    pr = mixdom2_params(rng, K, F, C)
    Lx = _lognormal_lengths(rng, B, median, sigma, lo, hi) if lengths is None else np.asarray(lengths)
    ts = rng.uniform(*t_range, size=B)
    chi = mixdom.build_nested_trans(pr["main_lam"], pr["main_mu"], torch.as_tensor(ts),
                                    pr["dom_lam"], pr["dom_mu"], pr["dom_w"], pr["frag_w"],
                                    pr["ext"]).numpy()                      # (B,N,N)
    types = np.array(mixdom.state_types(K, F))
    N = len(types)
    dom, frag = np.zeros(N, int), np.zeros(N, int)
    body = np.arange(N - 2)
    dom[2:] = body // (5 * F)
    frag[2:] = (body % (5 * F)) % F
    subs = np.stack([expm_batch(pr["class_Q"][c], ts) for c in range(C)], axis=1)  # (B,C,A,A)
    cpi = pr["class_pi"]
    cd = pr["class_dist"]
    xs, ys, alns = [], [], []
    for b in range(B):
        # ancestor from the class-mixture equilibrium of a random (d,f)
        wx = cd.reshape(-1, C).mean(axis=0) @ cpi
        x = rng.choice(A, size=int(Lx[b]), p=wx / wx.sum())

        def em_match(s, a, b=b):
            w = cd[dom[s], frag[s]] * cpi[:, a]
            return (w[:, None] * subs[b, :, a, :]).sum(axis=0)

        def em_ins(s):
            return cd[dom[s], frag[s]] @ cpi
        path = []
        y = sample_descendant(rng, x, chi[b], types, em_match, em_ins, wx[x] / wx.sum(), path=path)
        xs.append(x.astype(np.int64))
        ys.append(y)
        alns.append(np.asarray(path, dtype=np.int64))
    log_trans = torch.log(torch.clamp(torch.as_tensor(chi), min=1e-30))
    log_tab = tables.per_class_tables(torch.as_tensor(subs), torch.as_tensor(cpi), torch.as_tensor(cd))
    groups = tables.state_groups_per_class(N, F)
    pr.update(ts=ts, class_subs=subs, alignments=alns)
    return Workload(name, N, list(types), groups, xs, ys, ts, log_trans, log_tab, pr)


def alignment_columns(w):
    """Per pair (col_types, anc_full, desc_full) of the alignment each pair of
    a make_mixdom2 workload was generated from (the sampler stops at the
    first E, so the path consumes a prefix of y when y was clipped)."""
    cts, ancs, descs = [], [], []
    for x, y, ct in zip(w.xs, w.ys, w.params["alignments"]):
        a = np.zeros(len(ct), dtype=np.int64)
        d = np.zeros(len(ct), dtype=np.int64)
        ia = np.cumsum((ct == params.M) | (ct == params.D)) - 1
        idd = np.cumsum((ct == params.M) | (ct == params.I)) - 1
        mx = (ct == params.M) | (ct == params.D)
        my = (ct == params.M) | (ct == params.I)
        a[mx] = np.asarray(x)[ia[mx]]
        d[my] = np.asarray(y)[idd[my]]
        cts.append(ct)
        ancs.append(a)
        descs.append(d)
    return cts, ancs, descs
