"""BDI sufficient statistics from 5x5 TKF91-semantics counts (oracle, mpmath).

Restates python/tkfmixdom/jax/core/bdi.py:
  * count groups of the six WFST factors alpha, 1-alpha, beta, 1-beta,
    gamma, 1-gamma (_bdi_general_core, bdi.py:682-723);
  * the smooth log-likelihood logP = sum_g n_g log xi_g with
    beta = lam t r / (mu t r + 1), r = expm1((mu-lam) t) / ((mu-lam) t),
    gamma = 1 - mu beta / (lam max(1 - alpha, 1e-30)) and every log argument
    floored at 1e-30 (bdi.py:101-160);
  * lam_score = lam dlogP/dlam, mu_score = mu dlogP/dmu,
    cons = sum n[:, I] - sum n[:, D];
  * general branch (|1 - kappa| >= 1e-10, bdi.py:35):
    E_S = (cons + mu_score - lam_score - lam T) / (lam - mu),
    E_B = lam_score + lam E_S + lam T, E_D = mu_score + mu E_S;
  * limit branch (bdi.py:726-778): E_S = f'(mu) with
    f(l) = cons + mu dlogP/dmu(l, mu) - l dlogP/dlam(l, mu) - l T.
The derivatives are taken numerically by mpmath at 40 significant digits of
the same closed form (the r factor evaluated exactly rather than by its
4-term Taylor series below |x| = 1e-5, a difference below 1e-20).
"""
import mpmath
import numpy as np

from .tkf import S, M, I, D, E

ES_LHOPITAL_THRESHOLD = 1e-10


def count_groups(n):
    n = np.asarray(n, dtype=np.float64)
    return (n[S, M] + n[M, M] + n[I, M] + n[D, M],
            n[S, D] + n[M, D] + n[I, D] + n[D, D],
            n[S, I] + n[M, I] + n[I, I],
            n[S, M] + n[S, D] + n[S, E] + n[M, M] + n[M, D] + n[M, E] + n[I, M] + n[I, D] + n[I, E],
            n[D, I],
            n[D, M] + n[D, D] + n[D, E])


def _logP(lam, mu, t, g):
    t = mpmath.mpf(t)
    alpha = mpmath.exp(-mu * t)
    x = (mu - lam) * t
    r = mpmath.mpf(1) if x == 0 else mpmath.expm1(x) / x
    beta = lam * t * r / (mu * t * r + 1)
    gamma = 1 - mu * beta / (lam * max(1 - alpha, mpmath.mpf(1e-30)))
    fl = lambda v: mpmath.log(max(v, mpmath.mpf(1e-30)))
    vals = (fl(alpha), fl(1 - alpha), fl(beta), fl(1 - beta), fl(gamma), fl(1 - gamma))
    return sum(mpmath.mpf(float(c)) * v for c, v in zip(g, vals))


def bdi_stats(n, lam, mu, t, T):
    """(E_B, E_D, E_S) for one 5x5 count matrix."""
    with mpmath.workdps(40):
        g = count_groups(n)
        L, Mu, TT = mpmath.mpf(float(lam)), mpmath.mpf(float(mu)), mpmath.mpf(float(T))
        f = lambda a, b: _logP(a, b, t, g)
        dl = mpmath.diff(f, (L, Mu), (1, 0))
        dm = mpmath.diff(f, (L, Mu), (0, 1))
        lam_score, mu_score = L * dl, Mu * dm
        nn = np.asarray(n, dtype=np.float64)
        cons = mpmath.mpf(float(nn[:, I].sum() - nn[:, D].sum()))
        if abs(1.0 - float(lam) / float(mu)) < ES_LHOPITAL_THRESHOLD:
            f_lm = mpmath.diff(f, (Mu, Mu), (1, 1))
            f_ll = mpmath.diff(f, (Mu, Mu), (2, 0))
            f_l = mpmath.diff(f, (Mu, Mu), (1, 0))
            E_S = Mu * f_lm - f_l - Mu * f_ll - TT
        else:
            E_S = (cons + mu_score - lam_score - L * TT) / (L - Mu)
        E_B = lam_score + L * E_S + L * TT
        E_D = mu_score + Mu * E_S
        return float(E_B), float(E_D), float(E_S)


def tkf92_resolve(n_chi, ext, tau92):
    """Split the TKF92 self-loop counts into extensions and new-fragment
    events (train/tkf92_svi_bw.py:75-123).  Returns (resolved (5,5),
    ext_count, notext_count)."""
    n = np.array(n_chi, dtype=np.float64, copy=True)
    ext_count = 0.0
    for s in (M, I, D):
        d = tau92[s, s]
        frac = min(max(ext / d, 0.0), 1.0) if d > 1e-30 else 0.0
        ext_count += frac * n[s, s]
        n[s, s] = (1.0 - frac) * n[s, s]
    notext = n[M].sum() + n[I].sum() + n[D].sum()
    return n, ext_count, notext


def kappa_coefficients(n):
    """L (log kappa) and M (log 1-kappa) coefficients of a resolved 5x5
    matrix (train/tkf92_svi_bw.py:160-167)."""
    n = np.asarray(n)
    L = n[S, M] + n[S, D] + n[M, M] + n[M, D] + n[I, M] + n[I, D] + n[D, M] + n[D, D]
    Mc = n[S, E] + n[M, E] + n[I, E] + n[D, E]
    return L, Mc
