"""TKF91 / TKF92 transition matrices (oracle restatement, numpy FP64).

Follows python/tkfmixdom/jax/core/bdi.py:40-97 (alpha, beta, gamma, kappa with
the L'Hopital switch at |1 - kappa| < 1e-4) and core/params.py:29-108
(tkf91_trans, tkf91_trans_cond, tkf92_trans).  State order S, M, I, D, E.
"""
import numpy as np

S, M, I, D, E = 0, 1, 2, 3, 4
EQUAL_RATE_THRESHOLD = 1e-4          # core/bdi.py:14
NEG_INF = -1e30                      # dp/hmm.py:108


def safe_log(x):
    """log(x) for x > 0, NEG_INF for x <= 0 (dp/hmm.py:111-122)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.full(x.shape, NEG_INF)
    pos = x > 0
    out[pos] = np.log(x[pos])
    return out


def floor_log(x, floor=1e-30):
    """log(max(x, floor)) -- the floor used by train_pfam / tkf92_svi_bw
    (python/train_pfam.py:1632-1640, train/tkf92_svi_bw.py:150)."""
    return np.log(np.maximum(np.asarray(x, dtype=np.float64), floor))


def tkf_alpha(mu, t):
    """core/bdi.py:40-42."""
    return np.exp(-mu * t)


def tkf_beta(lam, mu, t):
    """core/bdi.py:45-69: general formula, denom guard, L'Hopital limit."""
    lam = np.asarray(lam, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    s = mu * t
    eta = np.exp(-lam * t)
    alpha = np.exp(-s)
    numer = lam * (eta - alpha)
    denom = mu * eta - lam * alpha
    small = np.abs(denom) < 1e-30
    general = np.where(small, 0.0, numer / np.where(small, 1.0, denom))
    limit = s / (1.0 + s)
    kappa = lam / mu
    return np.where(np.abs(1.0 - kappa) < EQUAL_RATE_THRESHOLD, limit, general)


def tkf_gamma(lam, mu, t):
    """core/bdi.py:72-93."""
    lam = np.asarray(lam, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    s = mu * t
    alpha = tkf_alpha(mu, t)
    beta = tkf_beta(lam, mu, t)
    Phi = 1.0 - alpha
    general = np.where(np.abs(Phi) < 1e-30, 0.0,
                       1.0 - mu * beta / (lam * np.maximum(Phi, 1e-30)))
    phi = np.where(s < 1e-10, 1.0 - s / 2.0, Phi / np.maximum(s, 1e-30))
    limit = 1.0 - 1.0 / ((1.0 + s) * phi)
    kappa = lam / mu
    return np.where(np.abs(1.0 - kappa) < EQUAL_RATE_THRESHOLD, limit, general)


def tkf_kappa(lam, mu):
    """core/bdi.py:96-97."""
    return lam / mu


def tkf91_trans(lam, mu, t):
    """5x5 TKF91 pair-HMM transitions, core/params.py:29-54."""
    a = tkf_alpha(mu, t)
    b = tkf_beta(lam, mu, t)
    g = tkf_gamma(lam, mu, t)
    k = tkf_kappa(lam, mu)
    tau = np.zeros((5, 5))
    for src in (S, M, I):
        tau[src, M] = (1 - b) * k * a
        tau[src, I] = b
        tau[src, D] = (1 - b) * k * (1 - a)
        tau[src, E] = (1 - b) * (1 - k)
    tau[D, M] = (1 - g) * k * a
    tau[D, I] = g
    tau[D, D] = (1 - g) * k * (1 - a)
    tau[D, E] = (1 - g) * (1 - k)
    return tau


def tkf91_trans_cond(lam, mu, t):
    """Ancestor-length-conditioned TKF91, core/params.py:57-89."""
    a = tkf_alpha(mu, t)
    b = tkf_beta(lam, mu, t)
    g = tkf_gamma(lam, mu, t)
    tau = np.zeros((5, 5))
    for src in (S, M, I):
        tau[src, M] = (1 - b) * a
        tau[src, I] = b
        tau[src, D] = (1 - b) * (1 - a)
        tau[src, E] = 1 - b
    tau[D, M] = (1 - g) * a
    tau[D, I] = g
    tau[D, D] = (1 - g) * (1 - a)
    tau[D, E] = 1 - g
    return tau


def tkf92_trans(lam, mu, t, ext):
    """TKF92: (1-ext)*tau91 rows for M/I/D plus ext on the diagonal; S row
    unchanged (core/params.py:92-108)."""
    t91 = tkf91_trans(lam, mu, t)
    tau = np.zeros((5, 5))
    tau[S] = t91[S]
    for src in (M, I, D):
        tau[src] = (1 - ext) * t91[src]
        tau[src, src] = ext + (1 - ext) * t91[src, src]
    return tau


TKF_STATE_TYPES = np.array([S, M, I, D, E], dtype=np.int32)
