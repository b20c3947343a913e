"""GPU 1D (alignment-constrained) E-step delta (tkfmi.estep.process_pairs_batched)
vs the oracle chain: oracle/fb1d.py forward/backward on constrained
emissions, per-column domain / class reductions written as in
train_pfam.py:2338-2546, Holmes-Rubin (oracle/hr.py) and the exploded-model
restoration (oracle/suffstats.py).  Tolerance 1e-10 relative (1e-9 through
an eigendecomposition or a BDI division)."""
import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu
A = 20


def _gtr(S, pi):
    Q = S * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    return Q


def _rand_Q(rng):
    S = rng.gamma(1.0, 1.0, (A, A))
    S = (S + S.T) / 2
    pi = rng.dirichlet(np.ones(A))
    return _gtr(S, pi), pi


def _assert_close(got, want, rtol=1e-10, what=""):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want) / (np.abs(want) + 1e-12 * scale)
    assert err.max() <= rtol, (what, float(err.max()))


def _pack(match, single):
    return np.concatenate([match.reshape(match.shape[:-2] + (441,)), single, single], axis=-1)


def _pairs(rng, n):
    out = []
    for b in range(n):
        L = 30 + 23 * b
        states = rng.integers(1, 4, L)
        n_anc = int(((states == 1) | (states == 3)).sum())
        n_desc = int(((states == 1) | (states == 2)).sum())
        out.append((None, None, states, rng.integers(0, A, n_anc), rng.integers(0, A, n_desc),
                    float(rng.uniform(0.1, 1.5))))
    return out


def _cp(K, F, rng):
    return dict(main_ins=0.04, main_del=0.05, dom_ins=rng.uniform(0.02, 0.05, K),
                dom_del=rng.uniform(0.05, 0.07, K), dom_weights=rng.dirichlet(np.ones(K)),
                frag_weights=rng.dirichlet(np.ones(F), size=K),
                ext_rates=rng.uniform(0.5, 0.9, (K, F)))


def test_process_pairs_batched_per_domain_matches_oracle():
    from oracle import emissions, fb1d, hr as ohr, mixdom, suffstats
    from tkfmi import estep
    rng = np.random.default_rng(21)
    K, F = 2, 2
    Qs, pis = zip(*[_rand_Q(rng) for _ in range(K)])
    Qs, pis = np.stack(Qs), np.stack(pis)
    cp = _cp(K, F, rng)
    st = mixdom.state_types(K, F)
    N = len(st)
    pairs = _pairs(rng, 6)
    print("[1d estep per-domain] K=2 F=2 pairs=6", flush=True)
    delta = estep.process_pairs_batched(pairs, cp, st, Qs[0], pis[0], N, None, dom_Qs=Qs,
                                        dom_pis=pis, n_dom=K, n_frag=F)
    groups = [0, 0] + [(s - 2) // (5 * F) for s in range(2, N)]
    agg = np.zeros((N, N)); ll = 0.0
    dm = np.zeros((K, A, A)); di = np.zeros((K, A)); dd = np.zeros((K, A))
    W = np.zeros((K, A)); U = np.zeros((K, A, A))
    recs = []
    for (_, _, states, ac, dc, t) in pairs:
        t32 = float(np.float32(t))
        chi = mixdom.build_nested_trans(cp["main_ins"], cp["main_del"], t32, cp["dom_ins"],
                                        cp["dom_del"], cp["dom_weights"], cp["frag_weights"],
                                        cp["ext_rates"])
        subs = np.stack([scipy.linalg.expm(Qs[k] * t) for k in range(K)])
        ls, lp_ = emissions.log_tables_per_domain(subs, pis)
        tab = _pack(lp_[:, :, None] + ls, lp_)
        ct, a, d = fb1d.columns_from_alignment(states, ac, dc)
        le = fb1d.constrained_emissions(st, groups, tab, ct, a, d)
        lp, post, n = fb1d.forward_backward_1d(np.log(np.maximum(chi, 1e-30)), le, 0, 1)
        agg += n; ll += lp
        ec = fb1d.emission_counts_1d(post, st, groups, ct, a, d, K)
        m = ec[:, :441].reshape(K, 21, 21)[:, :A, :A]
        dm += m; di += ec[:, 441:441 + A]; dd += ec[:, 462:462 + A]
        for k in range(K):
            w_, u_ = ohr.weighted_stats(Qs[k], pis[k], t32, m[k])
            W[k] += w_; U[k] += u_
        recs.append((n, t32))
    assert delta["n_pairs"] == 6 and delta["n_files"] == 1
    assert delta["total_ll"] == pytest.approx(ll, rel=1e-10)
    _assert_close(delta["agg_n_chi"], agg, what="agg_n_chi")
    _assert_close(delta["dom_match_counts"], dm, what="dom_match")
    _assert_close(delta["match_counts"], dm.sum(0), what="match")
    _assert_close(delta["dom_insert_counts"], di, what="ins")
    _assert_close(delta["dom_delete_counts"], dd, what="del")
    _assert_close(delta["dom_W"], W, rtol=1e-9, what="W")
    _assert_close(delta["dom_U"], U, rtol=1e-9, what="U")
    top = sum(suffstats.exact_suffstats(n, cp["main_ins"], cp["main_del"], t, cp["dom_ins"],
                                        cp["dom_del"], cp["dom_weights"], cp["frag_weights"],
                                        cp["ext_rates"])["top_5x5"] for n, t in recs)
    _assert_close(delta["exact_ss"]["top_5x5"], top, rtol=1e-9, what="top_5x5")


def test_process_pairs_batched_per_class_matches_oracle():
    """MixDom2 class counts: per-column responsibilities
    q = cd[d,f,c] f_c / sum_c' cd[d,f,c'] f_c' weighted by the fragment's
    posterior (train_pfam.py:2410-2546), summed over columns."""
    from oracle import emissions, fb1d, hr as ohr, mixdom
    from tkfmi import estep
    rng = np.random.default_rng(22)
    K, F, C = 2, 2, 3
    Sx = rng.gamma(1.0, 1.0, (C, A, A))
    Sx = (Sx + Sx.transpose(0, 2, 1)) / 2
    cpis = rng.dirichlet(np.ones(A), size=C)
    cd = rng.dirichlet(np.ones(C), size=(K, F))
    Qlg, pilg = _rand_Q(rng)
    cp = _cp(K, F, rng)
    st = mixdom.state_types(K, F)
    N = len(st)
    dom, frag = mixdom.state_dom_frag(K, F)
    groups = [0, 0] + [int(dom[s] * F + frag[s]) for s in range(2, N)]
    pairs = _pairs(rng, 4)
    print("[1d estep per-class] K=2 F=2 C=3 pairs=4", flush=True)
    delta = estep.process_pairs_batched(pairs, cp, st, Qlg, pilg, N, None, n_dom=K, n_frag=F,
                                        classdist_params=dict(classdist=cd, class_S_exch=Sx,
                                                              class_pis=cpis))
    Qc = np.stack([_gtr(Sx[c], cpis[c]) for c in range(C)])
    agg = np.zeros((N, N)); ll = 0.0
    cm = np.zeros((C, A, A)); ci = np.zeros((C, A)); cdl = np.zeros((C, A)); cdc = np.zeros((K, F, C))
    dmm = np.zeros((K, A, A))
    cW = np.zeros((C, A)); cU = np.zeros((C, A, A))
    stn = np.asarray(st)
    for (_, _, states, ac, dc, t) in pairs:
        t32 = float(np.float32(t))
        chi = mixdom.build_nested_trans(cp["main_ins"], cp["main_del"], t32, cp["dom_ins"],
                                        cp["dom_del"], cp["dom_weights"], cp["frag_weights"],
                                        cp["ext_rates"])
        subs = np.stack([scipy.linalg.expm(Qc[c] * t) for c in range(C)])
        mt, sg = emissions.log_mixture_tables_per_class(subs, cpis, cd)
        tab = _pack(mt, sg).reshape(K * F, -1)
        ct, a, d = fb1d.columns_from_alignment(states, ac, dc)
        le = fb1d.constrained_emissions(st, groups, tab, ct, a, d)
        lp, post, n = fb1d.forward_backward_1d(np.log(np.maximum(chi, 1e-30)), le, 0, 1)
        agg += n; ll += lp
        cm_p = np.zeros((C, A, A))
        for tcol in range(len(ct)):
            for s in range(2, N):
                if stn[s] != ct[tcol]:
                    continue
                w = post[tcol, s]
                dd_, ff = dom[s], frag[s]
                if ct[tcol] == 1:
                    fc = cpis[:, a[tcol]] * subs[:, a[tcol], d[tcol]]
                    dmm[dd_, a[tcol], d[tcol]] += w
                elif ct[tcol] == 2:
                    fc = cpis[:, d[tcol]]
                else:
                    fc = cpis[:, a[tcol]]
                q = cd[dd_, ff] * fc / (cd[dd_, ff] * fc).sum()
                cdc[dd_, ff] += w * q
                if ct[tcol] == 1:
                    cm_p[:, a[tcol], d[tcol]] += w * q
                elif ct[tcol] == 2:
                    ci[:, d[tcol]] += w * q
                else:
                    cdl[:, a[tcol]] += w * q
        cm += cm_p
        for c in range(C):
            w_, u_ = ohr.weighted_stats(Qc[c], cpis[c], t32, cm_p[c])
            cW[c] += w_; cU[c] += u_
    assert delta["total_ll"] == pytest.approx(ll, rel=1e-10)
    _assert_close(delta["agg_n_chi"], agg, what="agg_n_chi")
    _assert_close(delta["dom_match_counts"], dmm, what="dom_match")
    _assert_close(delta["class_match_counts"], cm, what="class_match")
    _assert_close(delta["class_insert_counts"], ci, what="class_ins")
    _assert_close(delta["class_delete_counts"], cdl, what="class_del")
    _assert_close(delta["classdist_counts"], cdc, what="classdist")
    _assert_close(delta["class_W"], cW, rtol=1e-9, what="class_W")
    _assert_close(delta["class_U"], cU, rtol=1e-9, what="class_U")
