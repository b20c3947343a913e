"""Holmes-Rubin expected dwell times and transition counts (oracle, numpy).

Restates python/tkfmixdom/jax/core/ctmc.py:
  * holmes_rubin_integrals (ctmc.py:173-232): spectral decomposition of the
    symmetrised S = D^1/2 Q D^-1/2, J^{kl}(T) = T e^{l_k T} for
    |l_k - l_l| < 1e-10 else (e^{l_k T} - e^{l_l T}) / (l_k - l_l), and the
    full tensor I[a,b,i,j] = int_0^T M_ai(s) M_jb(T-s) ds;
  * holmes_rubin_weighted_stats (ctmc.py:277-326):
    W[i] = sum_ab mc_ab I[a,b,i,i] / M_ab,
    U[i,j] = Q_ij sum_ab mc_ab I[a,b,i,j] / M_ab with a zero diagonal,
    M_ab replaced by 1 where |M_ab| < 1e-30.
The A^4 tensor is formed literally (this is the checker, not the product).
"""
import numpy as np


def integrals(Q, pi, T):
    Q = np.asarray(Q, np.float64)
    pi = np.asarray(pi, np.float64)
    sp = np.sqrt(pi)
    Sm = Q * (sp[:, None] / sp[None, :])
    Sm = (Sm + Sm.T) / 2
    lam, V = np.linalg.eigh(Sm)
    e = np.exp(lam * T)
    diff = lam[:, None] - lam[None, :]
    same = np.abs(diff) < 1e-10
    J = np.where(same, T * e[:, None], (e[:, None] - e[None, :]) / np.where(same, 1.0, diff))
    VV = V[:, None, :] * V[None, :, :]                      # (a,i,k)
    Isym = np.einsum("aik,kl,jbl->aibj", VV, J, VV)
    pf = np.sqrt((pi[None, None, :, None] * pi[None, :, None, None])
                 / (pi[:, None, None, None] * pi[None, None, None, :] + 1e-30))
    I = Isym.transpose(0, 2, 1, 3) * pf.transpose(0, 2, 1, 3)
    M = ((V * e[None, :]) @ V.T) * (sp[None, :] / sp[:, None])
    return I, M


def weighted_stats(Q, pi, T, mc):
    I, M = integrals(Q, pi, T)
    Ms = np.where(np.abs(M) < 1e-30, 1.0, M)
    ratio = np.asarray(mc, np.float64) / Ms
    W = np.einsum("ab,abii->i", ratio, I)
    U = np.asarray(Q) * np.einsum("ab,abij->ij", ratio, I)
    np.fill_diagonal(U, 0.0)
    return W, U
