"""Continuous-batching serving loop (tkfmi.serve): every request's result is
the same alone and in the mix (log-likelihood, transition counts, E-step
statistics: 1e-12 relative), a sample is checked against the oracle, and a
real-time Poisson stream completes with bounded latency."""
import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu


def _engine(**kw):
    from tkfmi import serve, synth
    rng = np.random.default_rng(0)
    Q, pi, _ = synth.random_Q_pi(rng)
    tkf = dict(lam=0.04, mu=0.05, ext=0.6, Q=Q, pi=pi)
    mix = synth.mixdom2_params(rng, 3, 1, 1)
    return serve.ServingEngine(tkf, mix, **kw), tkf, mix


def _close(a, b, rtol=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return bool(np.all(np.abs(a - b) <= rtol * np.abs(b) + 1e-14 * scale))


def test_results_same_alone_and_in_mix():
    from tkfmi import serve
    eng, _, _ = _engine(max_batch_pairs=24, slots=3)
    reqs = serve.poisson_requests(7, 80, rate=1e6, median=120, hi=600)
    mixed = eng.serve(reqs, realtime=False)
    solo_eng, _, _ = _engine(max_batch_pairs=1, slots=1)
    for r in reqs:
        alone = solo_eng.serve([serve.Request(r.rid, r.model, r.kind, r.x, r.y, r.t, 0.0)],
                               realtime=False)[r.rid]
        m = mixed[r.rid]
        assert _close(m.log_p, alone.log_p), (r.rid, m.log_p, alone.log_p)
        if r.kind == "estep":
            assert _close(m.n, alone.n), r.rid
            assert _close(m.stats, alone.stats), r.rid
        else:
            assert m.n is None


def test_served_results_match_oracle():
    from oracle import dp2d, emissions, tkf as otkf
    from tkfmi import serve
    eng, tkfp, _ = _engine()
    rng = np.random.default_rng(3)
    reqs = [serve.Request(i, "tkf92", "estep" if i % 2 else "score", rng.integers(0, 20, 30 + 9 * i),
                          rng.integers(0, 20, 25 + 11 * i), 0.3 + 0.2 * i) for i in range(4)]
    res = eng.serve(reqs, realtime=False)
    for r in reqs:
        T = otkf.tkf92_trans(tkfp["lam"], tkfp["mu"], r.t, tkfp["ext"])
        em = emissions.pair_hmm_emissions(otkf.TKF_STATE_TYPES, r.x, r.y,
                                          scipy.linalg.expm(tkfp["Q"] * r.t), tkfp["pi"])
        lp, _, n, _ = dp2d.forward_backward(np.log(np.maximum(T, 1e-30)), otkf.TKF_STATE_TYPES, em)
        assert res[r.rid].log_p == pytest.approx(lp, rel=1e-10)
        if r.kind == "estep":
            assert _close(res[r.rid].n, n, rtol=1e-10)


def test_realtime_poisson_stream_completes():
    from tkfmi import serve
    eng, _, _ = _engine()
    reqs = serve.poisson_requests(11, 400, rate=2000.0, median=200, hi=1200)
    import time
    t0 = time.perf_counter()
    res = eng.serve(reqs, realtime=True)
    wall = time.perf_counter() - t0
    assert len(res) == len(reqs)
    s = serve.summarize(reqs, res, wall, {"tkf92": 5, "mixdom2": 17})
    print(s, flush=True)
    assert np.isfinite([res[r.rid].log_p for r in reqs]).all()
    assert s["p50_ms"] <= s["p99_ms"] < 5000
