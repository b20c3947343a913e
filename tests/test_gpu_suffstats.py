"""GPU restoration / sufficient statistics / BDI (libtkfmi
tkfmi_restore_counts_batched, tkfmi_bdi_stats_batched) vs the oracle
(oracle/suffstats.py, oracle/bdi.py).  Tolerance 1e-10 relative with an
absolute floor of 1e-12 x the largest entry of each statistic."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _close(a, b, rtol=RTOL, floor=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return np.all(np.abs(a - b) <= rtol * np.abs(b) + floor * scale), \
        float((np.abs(a - b) / (np.abs(b) + floor * scale)).max())


CASES = [dict(K=3, F=1), dict(K=2, F=2), dict(K=1, F=3), dict(K=4, F=2)]


def _params(K, F, seed):
    rng = np.random.default_rng(seed)
    ext = rng.uniform(0.5, 0.9, (K, F, 1)) * rng.dirichlet(np.ones(F), size=(K, F))
    return dict(main_ins=0.04, main_del=0.05, dom_ins=rng.uniform(0.02, 0.06, K),
                dom_del=rng.uniform(0.05, 0.08, K), dom_w=rng.dirichlet(np.ones(K)),
                frag_w=rng.dirichlet(np.ones(F), size=K), ext=ext)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"K{c['K']}F{c['F']}")
def test_restore_matches_oracle(case):
    from oracle import mixdom, suffstats as oss
    from tkfmi import suffstats
    K, F = case["K"], case["F"]
    p = _params(K, F, 10 * K + F)
    rng = np.random.default_rng(K + F)
    ts = np.array([0.05, 0.3, 0.9, 1.7, 2.0])
    n_chi = []
    for t in ts:
        chi = mixdom.build_nested_trans(p["main_ins"], p["main_del"], t, p["dom_ins"], p["dom_del"],
                                        p["dom_w"], p["frag_w"], p["ext"])
        n_chi.append(np.where(chi > 0, rng.uniform(0.0, 4.0, chi.shape), 0.0))
    n_chi = np.stack(n_chi)
    got = suffstats.exact_suffstats_batch(n_chi, p["main_ins"], p["main_del"], ts, p["dom_ins"],
                                          p["dom_del"], p["dom_w"], p["frag_w"], p["ext"])
    for b, t in enumerate(ts):
        want = oss.exact_suffstats(n_chi[b], t=t, **p)
        for key in ("top_5x5", "dom_kappa", "dom_1mkappa", "ext", "term", "dom_w", "frag_w"):
            ok, worst = _close(got[key][b].cpu().numpy(), want[key])
            assert ok, (b, key, worst)
        ok, worst = _close(np.stack([d[b].cpu().numpy() for d in got["dom_M_5x5"]]), want["dom_M_5x5"])
        assert ok, (b, "dom_M_5x5", worst)


def test_restore_and_bdi_on_kernel_counts():
    """fb2d expected counts -> restoration -> BDI, all on the GPU, vs the
    oracle chain (numpy restoration, mpmath BDI) on the same counts."""
    from oracle import bdi as obdi, suffstats as oss
    from tkfmi import hmm, suffstats, synth
    w = synth.make_mixdom2(seed=5, B=6, lengths=[40, 120, 3, 77, 200, 150])
    pr = w.params
    res = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group)
    n = res["n"]
    ext = pr["ext"]
    got = suffstats.exact_suffstats_batch(n, pr["main_lam"], pr["main_mu"], w.ts, pr["dom_lam"],
                                          pr["dom_mu"], pr["dom_w"], pr["frag_w"], ext, with_bdi=True)
    bdi = got["bdi"].cpu().numpy()
    nn = n.cpu().numpy()
    K = pr["K"]
    for b in range(w.B):
        print(f"pair {b}", flush=True)
        want = oss.exact_suffstats(nn[b], pr["main_lam"], pr["main_mu"], w.ts[b], pr["dom_lam"],
                                   pr["dom_mu"], pr["dom_w"], pr["frag_w"], ext)
        ok, worst = _close(got["top_5x5"][b].cpu().numpy(), want["top_5x5"])
        assert ok, worst
        levels = [(want["top_5x5"], pr["main_lam"], pr["main_mu"],
                   w.ts[b] * want["top_5x5"][:, 4].sum())]
        for d in range(K):
            Td = w.ts[b] * (want["dom_M_5x5"][d][0].sum() + want["dom_1mkappa"][d])
            levels.append((want["dom_M_5x5"][d], pr["dom_lam"][d], pr["dom_mu"][d], Td))
        for lv, (n5, lam, mu, T) in enumerate(levels):
            eb, ed, es = obdi.bdi_stats(n5, lam, mu, w.ts[b], T)
            assert bdi[b, lv, 3] == pytest.approx(T, rel=1e-12)
            # lam/mu = 0.8 here: the general branch's 1/(lam - mu) costs little
            for g, o in zip(bdi[b, lv, :3], (eb, ed, es)):
                assert g == pytest.approx(o, rel=1e-9, abs=1e-9), (b, lv)


@pytest.mark.parametrize("lam", [0.04, 0.05 * (1 - 3e-11), 0.07])
def test_bdi_batch_matches_mpmath(lam):
    from oracle import bdi as obdi, dp2d, emissions, tkf
    from tkfmi import suffstats
    mu = 0.05
    rng = np.random.default_rng(3)
    ns, ts = [], []
    for b in range(5):
        t = rng.uniform(0.1, 2.0)
        P = rng.dirichlet(np.ones(20), size=20)
        pi = rng.dirichlet(np.ones(20))
        x, y = rng.integers(0, 20, 20 + 5 * b), rng.integers(0, 20, 18 + 4 * b)
        em = emissions.pair_hmm_emissions(tkf.TKF_STATE_TYPES, x, y, P, pi)
        ns.append(dp2d.forward_backward(tkf.safe_log(tkf.tkf91_trans(0.045, mu, t)),
                                        tkf.TKF_STATE_TYPES, em)[2])
        ts.append(t)
    ns, ts = np.stack(ns), np.array(ts)
    T = ts * ns[:, :, 4].sum(axis=1)
    EB, ED, ES = (v.cpu().numpy() for v in suffstats.tkf91_stats_from_counts_batch(ns, lam, mu, ts, T))
    for b in range(5):
        eb, ed, es = obdi.bdi_stats(ns[b], lam, mu, ts[b], T[b])
        # the general branch divides by lam - mu: allow for that cancellation
        tol = 1e-9 if abs(1 - lam / mu) > 1e-3 else 1e-4
        assert EB[b] == pytest.approx(eb, rel=tol, abs=tol)
        assert ED[b] == pytest.approx(ed, rel=tol, abs=tol)
        assert ES[b] == pytest.approx(es, rel=tol, abs=tol)
        cons = ns[b][:, 2].sum() - ns[b][:, 3].sum()
        assert EB[b] - ED[b] == pytest.approx(cons, rel=1e-6, abs=1e-6)


def test_restore_full_size_identities():
    """The benchmark's 1024 MixDom2 K=3 pairs: restored statistics obey the
    exact flow identities (one unit enters and leaves the top level, domain
    visits == top-level domain transitions)."""
    from tkfmi import hmm, suffstats, synth
    w = synth.make_mixdom2(seed=1, B=1024)
    pr = w.params
    n = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group)["n"]
    ss = suffstats.exact_suffstats_batch(n, pr["main_lam"], pr["main_mu"], w.ts, pr["dom_lam"],
                                         pr["dom_mu"], pr["dom_w"], pr["frag_w"], pr["ext"],
                                         with_bdi=True)
    top = ss["top_5x5"].cpu().numpy()
    np.testing.assert_allclose(top[:, 0].sum(axis=1), 1.0, rtol=1e-10)
    np.testing.assert_allclose(top[:, :, 4].sum(axis=1), 1.0, rtol=1e-10)
    np.testing.assert_allclose(ss["dom_w"].cpu().numpy().sum(axis=1), top[:, :, 1:4].sum(axis=(1, 2)),
                               rtol=1e-9)
    assert np.all(np.isfinite(ss["bdi"].cpu().numpy()))
