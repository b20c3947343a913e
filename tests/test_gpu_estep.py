"""GPU E-step callers (tkfmi.estep) vs the oracle chain:
estep_batch_tkf92, e_step (TKF92), e_step_mixdom2 and the
process_pairs_batched_2d delta dict.  Tolerance 1e-10 relative (1e-9 for
quantities that pass through a BDI division or an eigendecomposition)."""
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
    return _gtr(S, pi), pi, S


def _assert_close(got, want, rtol=1e-10, what=""):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want) / (np.abs(want) + 1e-12 * scale)
    assert err.max() <= rtol, (what, float(err.max()))


def test_estep_batch_tkf92_matches_oracle():
    from oracle import bdi, dp2d, emissions, tkf
    from tkfmi import estep
    rng = np.random.default_rng(1)
    Q, pi, _ = _rand_Q(rng)
    lam, mu, ext = 0.04, 0.05, 0.65
    Lx = np.array([0, 7, 40, 90, 130])
    Ly = np.array([3, 0, 44, 85, 120])
    xs_pad = np.zeros((5, 140), np.int32)
    ys_pad = np.zeros((5, 140), np.int32)
    for b in range(5):
        xs_pad[b, :Lx[b]] = rng.integers(0, A, Lx[b])
        ys_pad[b, :Ly[b]] = rng.integers(0, A, Ly[b])
    ts = rng.uniform(0.1, 1.8, 5)
    out = estep.estep_batch_tkf92(xs_pad, ys_pad, ts, Lx, Ly, lam, mu, ext, Q, pi)
    o = {k: v.cpu().numpy() for k, v in out.items()}
    for b in range(5):
        x, y, t = xs_pad[b, :Lx[b]], ys_pad[b, :Ly[b]], ts[b]
        tau = tkf.tkf92_trans(lam, mu, t, ext)
        em = emissions.pair_hmm_emissions(tkf.TKF_STATE_TYPES, x, y, scipy.linalg.expm(Q * t), pi)
        lp, _, n, _ = dp2d.forward_backward(np.log(np.maximum(tau, 1e-30)), tkf.TKF_STATE_TYPES, em)
        assert o["log_p"][b] == pytest.approx(lp, rel=1e-10)
        _assert_close(o["n_chi_diag"][b], n, what="n_chi")
        res, ec, ne = bdi.tkf92_resolve(n, ext, tau)
        L, M = bdi.kappa_coefficients(res)
        eb, ed, es = bdi.bdi_stats(res, lam, mu, t, t)
        for key, want in (("ext_count", ec), ("notext_count", ne), ("L", L), ("M", M)):
            assert o[key][b] == pytest.approx(want, rel=1e-10, abs=1e-12), key
        for key, want in (("B", eb), ("D", ed), ("S", es)):
            assert o[key][b] == pytest.approx(max(want, 0.0), rel=1e-9, abs=1e-9), key
        assert o["T"][b] == t


def test_e_step_tkf92_matches_oracle():
    from oracle import dp2d, emissions, tkf
    from tkfmi import estep
    rng = np.random.default_rng(2)
    Q, pi, _ = _rand_Q(rng)
    x, y = rng.integers(0, A, 37), rng.integers(0, A, 41)
    p = dict(ins_rate=0.04, del_rate=0.05, t=0.6, ext=0.5, Q=Q, pi=pi)
    lp, n, ss = estep.e_step("tkf92", p, x, y)
    st = tkf.TKF_STATE_TYPES
    em = emissions.pair_hmm_emissions(st, x, y, scipy.linalg.expm(Q * 0.6), pi)
    lp_o, post, n_o, _ = dp2d.forward_backward(tkf.safe_log(tkf.tkf92_trans(0.04, 0.05, 0.6, 0.5)), st, em)
    assert lp == pytest.approx(lp_o, rel=1e-10)
    _assert_close(n.cpu().numpy(), n_o, what="n")
    ec = emissions.emission_counts(post, st, [0] * 5, x, y, 1)[0]
    nm = 21 * 21
    _assert_close(ss["pi_obs_counts"]["match_anc_count"], ec[:nm].reshape(21, 21)[:20, :20].sum(1))
    _assert_close(ss["pi_obs_counts"]["insert_count"], ec[nm:nm + 21][:20])
    _assert_close(ss["pi_obs_counts"]["delete_count"], ec[nm + 21:][:20])
    m_post = post[1:, 1:, 1]
    want = {(int(x[i]), int(y[j]), i, j) for i, j in zip(*np.where(m_post > 1e-8))}
    assert len(ss["match_pairs"]) == len(want)
    tot = sum(w for _, _, w in ss["match_pairs"])
    assert tot == pytest.approx(m_post[m_post > 1e-8].sum(), rel=1e-10)
    with pytest.raises(ValueError):
        estep.e_step("nope", p, x, y)


def test_e_step_mixdom2_matches_oracle():
    """Class counts from emission counts x responsibilities == the per-cell
    responsibility sums of train/likelihood.py:296-347 on oracle posteriors;
    class_W / class_U == oracle Holmes-Rubin of those counts."""
    from oracle import dp2d, emissions, hr as ohr, mixdom, tkf
    from tkfmi import estep
    rng = np.random.default_rng(3)
    K, F, C = 2, 1, 3
    S = rng.gamma(1.0, 1.0, (C, A, A))
    S = (S + S.transpose(0, 2, 1)) / 2
    cpis = rng.dirichlet(np.ones(A), size=C)
    cd = rng.dirichlet(np.ones(C), size=(K, F))
    p = dict(main_ins=0.04, main_del=0.05, t=0.9, dom_ins=[0.03, 0.05], dom_del=[0.05, 0.06],
             dom_weights=[0.3, 0.7], frag_weights=[[1.0], [1.0]], ext_rates=[[0.6], [0.8]],
             class_S_exch=S, class_pis=cpis, classdist=cd)
    x, y = rng.integers(0, A, 33), rng.integers(0, A, 28)
    lp, n, ss = estep.e_step("mixdom2", p, x, y)
    st = mixdom.state_types(K, F)
    chi = mixdom.build_nested_trans(0.04, 0.05, 0.9, p["dom_ins"], p["dom_del"], p["dom_weights"],
                                    p["frag_weights"], p["ext_rates"])
    subs = np.stack([scipy.linalg.expm(0.9 * _gtr(S[c], cpis[c])) for c in range(C)])
    em = emissions.pair_hmm_emissions_per_class(st, x, y, subs, cpis, cd, K, F)
    lp_o, post, n_o, _ = dp2d.forward_backward(tkf.safe_log(chi), st, em)
    assert lp == pytest.approx(lp_o, rel=1e-10)
    _assert_close(n.cpu().numpy(), n_o, what="n")
    # per-cell responsibilities
    dom, frag = mixdom.state_dom_frag(K, F)
    cm = np.zeros((C, A, A)); ci = np.zeros((C, A)); cdl = np.zeros((C, A)); cdc = np.zeros((K, F, C))
    for s in range(2, len(st)):
        d, f = dom[s], frag[s]
        for i in range(len(x) + 1):
            for j in range(len(y) + 1):
                w = post[i, j, s]
                if w == 0.0:
                    continue
                if st[s] == 1:
                    fc = cpis[:, x[i - 1]] * subs[:, x[i - 1], y[j - 1]]
                elif st[s] == 2:
                    fc = cpis[:, y[j - 1]]
                else:
                    fc = cpis[:, x[i - 1]]
                q = cd[d, f] * fc / (cd[d, f] * fc).sum()
                cdc[d, f] += w * q
                if st[s] == 1:
                    cm[:, x[i - 1], y[j - 1]] += w * q
                elif st[s] == 2:
                    ci[:, y[j - 1]] += w * q
                else:
                    cdl[:, x[i - 1]] += w * q
    _assert_close(ss["class_match_counts"], cm, what="match")
    _assert_close(ss["class_insert_counts"], ci, what="insert")
    _assert_close(ss["class_delete_counts"], cdl, what="delete")
    _assert_close(ss["classdist_counts"], cdc, what="classdist")
    for c in range(C):
        W, U = ohr.weighted_stats(_gtr(S[c], cpis[c]), cpis[c], 0.9, cm[c])
        _assert_close(ss["class_W"][c], W, rtol=1e-9, what="W")
        _assert_close(ss["class_U"][c], U, rtol=1e-9, what="U")


def _check_process_pairs_2d(cp, Qs, pis, K, F, pairs):
    """process_pairs_batched_2d's delta dict vs the oracle chain: chi and
    scipy expm emissions at the float32-rounded t, oracle forward/backward,
    emission counts, Holmes-Rubin and the restored top-level statistics."""
    from oracle import bdi, dp2d, emissions, hr as ohr, mixdom, suffstats
    from tkfmi import estep
    st = mixdom.state_types(K, F)
    N = len(st)
    delta = estep.process_pairs_batched_2d(pairs, cp, st, Qs[0], pis[0], N, dom_Qs=Qs, dom_pis=pis,
                                           n_dom=K, n_frag=F)
    agg = np.zeros((N, N)); ll = 0.0
    dm = np.zeros((K, A, A)); di = np.zeros((K, A)); dd = np.zeros((K, A))
    W = np.zeros((K, A)); U = np.zeros((K, A, A))
    recs = []
    for (x, y, _, _, _, t) in pairs:
        t = float(np.float32(t))
        chi = mixdom.build_nested_trans(cp["main_ins"], cp["main_del"], t, cp["dom_ins"], cp["dom_del"],
                                        cp["dom_weights"], cp["frag_weights"], cp["ext_rates"])
        subs = np.stack([scipy.linalg.expm(Qs[k] * t) for k in range(K)])
        em = emissions.pair_hmm_emissions_per_domain(st, x, y, subs, pis, K, F)
        lp, post, n, _ = dp2d.forward_backward(np.log(np.maximum(chi, 1e-30)), st, em)
        agg += n; ll += lp
        groups = [0, 0] + [(s - 2) // (5 * F) for s in range(2, N)]
        ec = emissions.emission_counts(post, st, groups, x, y, K)
        m = ec[:, :441].reshape(K, 21, 21)[:, :20, :20]
        dm += m; di += ec[:, 441:461]; dd += ec[:, 462:482]
        for k in range(K):
            w_, u_ = ohr.weighted_stats(Qs[k], pis[k], t, m[k])
            W[k] += w_; U[k] += u_
        recs.append((n, t))
    assert delta["n_pairs"] == len(pairs) and delta["n_files"] == 1
    assert delta["total_ll"] == pytest.approx(ll, rel=1e-10)
    _assert_close(delta["agg_n_chi"], agg, what="agg_n_chi")
    _assert_close(delta["dom_match_counts"], dm, what="dom_match")
    _assert_close(delta["match_counts"], dm.sum(0), what="match")
    _assert_close(delta["dom_insert_counts"], di, what="ins")
    _assert_close(delta["dom_delete_counts"], dd, what="del")
    _assert_close(delta["dom_W"], W, rtol=1e-9, what="W")
    _assert_close(delta["dom_U"], U, rtol=1e-9, what="U")
    es = delta["exact_ss"]
    sss = [suffstats.exact_suffstats(n, cp["main_ins"], cp["main_del"], t, cp["dom_ins"],
                                     cp["dom_del"], cp["dom_weights"], cp["frag_weights"],
                                     cp["ext_rates"])["top_5x5"] for n, t in recs]
    _assert_close(es["top_5x5"], sum(sss), rtol=1e-9, what="top_5x5")
    EBs = [bdi.bdi_stats(tt, cp["main_ins"], cp["main_del"], t, t * tt[:, 4].sum())[0]
           for tt, (_, t) in zip(sss, recs)]
    assert es["top_E_B"] == pytest.approx(sum(EBs), rel=1e-8)


def test_process_pairs_batched_2d_delta_matches_oracle():
    rng = np.random.default_rng(4)
    K, F = 2, 1
    Qs, pis = [], []
    for _ in range(K):
        Q, pi, _ = _rand_Q(rng)
        Qs.append(Q)
        pis.append(pi)
    Qs, pis = np.stack(Qs), np.stack(pis)
    cp = dict(main_ins=0.04, main_del=0.05, dom_ins=np.array([0.03, 0.045]),
              dom_del=np.array([0.05, 0.06]), dom_weights=np.array([0.4, 0.6]),
              frag_weights=np.array([[1.0], [1.0]]), ext_rates=np.array([[0.7], [0.85]]))
    pairs = []
    for b in range(5):
        x, y = rng.integers(0, A, 20 + 9 * b), rng.integers(0, A, 25 + 7 * b)
        pairs.append((x, y, None, None, None, float(rng.uniform(0.1, 1.5))))
    _check_process_pairs_2d(cp, Qs, pis, K, F, pairs)


@pytest.mark.parametrize("name", ["bw_d8f2_fullseed_15iter", "bw_d3f2_pfam100_20iter",
                                  "bw_d5f2_pfam100_20iter"])
def test_process_pairs_batched_2d_on_reference_parameters(name):
    """The E-step delta on the reference's own trained parameter sets
    (python/params/best, via tests/golden/params_best.npz): K = 8, 3, 5
    domains of F = 2 fragments (82, 32, 52 states).  bw_d8f2_fullseed's rate
    matrices are not exactly reversible and are used as given."""
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "params_best.npz"), allow_pickle=False)
    g = lambda k: d[f"{name}/{k}"]
    Qs, pis = g("dom_Qs"), g("dom_pis")
    K, F = g("ext_rates").shape
    cp = dict(main_ins=float(g("main_ins")), main_del=float(g("main_del")), dom_ins=g("dom_ins"),
              dom_del=g("dom_del"), dom_weights=g("dom_weights"), frag_weights=g("frag_weights"),
              ext_rates=g("ext_rates"))
    rng = np.random.default_rng(11)
    pairs = []
    for b in range(4):
        x = rng.choice(A, 18 + 11 * b, p=pis[b % K])
        y = rng.choice(A, 15 + 13 * b, p=pis[(b + 1) % K])
        pairs.append((x, y, None, None, None, float(rng.uniform(0.05, 2.0))))
    pairs.append((rng.integers(0, A, 3), np.zeros(0, np.int64), None, None, None, float(g("t_rep"))))
    _check_process_pairs_2d(cp, Qs, pis, K, F, pairs)
