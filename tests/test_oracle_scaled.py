"""The max-shifted product form of the oracle recursion (oracle/c/fb2d_scaled.c),
in FP64 and in 80-bit long double, against the log-space numpy oracle
(oracle/dp2d.py, itself pinned by brute-force path enumeration): log P,
transition counts, emission counts and posteriors at 1e-12 relative (FP64)
and 1e-12 (long double, whose values are the more accurate), on TKF and
MixDom layouts of 5..82 states, with empty sequences and impossible
emissions.  These builds make the long-pair golden fixtures."""
import numpy as np
import pytest

from oracle import cfb2d, dp2d, emissions, mixdom, tkf


def _close(a, b, rtol):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b) / (np.abs(b) + 1e-13 * scale)
    return float(err.max()) <= rtol, float(err.max())


def _model(K, F, rng):
    if K == 0:
        st = np.asarray(tkf.TKF_STATE_TYPES)
        lt = tkf.safe_log(tkf.tkf92_trans(0.04, 0.05, 0.7, 0.4))
        G, sg = 1, np.zeros(5, np.int32)
    else:
        st = np.asarray(mixdom.state_types(K, F))
        ext = rng.uniform(0.3, 0.8, (K, F, 1)) * rng.dirichlet(np.ones(F), size=(K, F))
        chi = mixdom.build_nested_trans(0.04, 0.05, 0.9, rng.uniform(0.02, 0.05, K),
                                        rng.uniform(0.05, 0.08, K), rng.dirichlet(np.ones(K)),
                                        rng.dirichlet(np.ones(F), size=K), ext)
        lt = tkf.floor_log(chi)
        G = K
        sg = np.array([0, 0] + [(s - 2) // (5 * F) for s in range(2, len(st))], np.int32)
    tab = rng.normal(-3.0, 1.0, (G, 483))
    tab[rng.random(tab.shape) < 0.03] = -1e30
    return st, sg, lt, tab


@pytest.mark.parametrize("K,F", [(0, 1), (3, 1), (2, 2), (8, 2)])
@pytest.mark.parametrize("long_double", [False, True])
def test_scaled_oracle_matches_log_space_oracle(K, F, long_double):
    rng = np.random.default_rng(10 * K + F)
    st, sg, lt, tab = _model(K, F, rng)
    for Lx, Ly in ((0, 0), (0, 5), (6, 0), (9, 13), (31, 24)):
        x, y = rng.integers(0, 21, Lx), rng.integers(0, 21, Ly)
        em = emissions.emit_from_tables(st, sg, tab, x, y)
        lp, post, n, _ = dp2d.forward_backward(lt, st, em)
        ec = emissions.emission_counts(post, st, sg, x, y, len(tab))
        samp = np.array([(i, j) for i in range(0, Lx + 1, 3) for j in range(0, Ly + 1, 4)], np.int32)
        lp2, n2, ec2, post2 = cfb2d.fb2d_scaled(lt, st, sg, tab, x, y, samp, long_double=long_double,
                                                nthreads=4)
        assert abs(lp2 - lp) <= 1e-12 * abs(lp), (Lx, Ly, lp2, lp)
        for what, a, b in (("n", n2, n), ("ecnt", ec2, ec),
                           ("post", post2, post[samp[:, 0], samp[:, 1]])):
            ok, worst = _close(a, b, 1e-12)
            assert ok, (what, Lx, Ly, worst)
