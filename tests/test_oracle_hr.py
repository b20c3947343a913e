"""Pins the Holmes-Rubin oracle (oracle/hr.py):
* I[a,b,i,j] against direct quadrature of int_0^T M_ai(s) M_jb(T-s) ds
  (Gauss-Legendre, 40 nodes, scipy expm) on a 4-state chain;
* total dwell time: sum_i W_i = T sum_ab mc_ab (every path spends T);
* the time derivative: sum_ij Q_ij I[a,b,i,j] = T (Q M(T))_ab, i.e.
  sum_{i!=j} U_ij + sum_i Q_ii W_i = T sum_ab mc_ab d log M_ab / dT.
CPU only."""
import numpy as np
import pytest
import scipy.linalg

from oracle import hr


def _gtr(rng, A):
    S = rng.gamma(1.0, 1.0, (A, A))
    S = (S + S.T) / 2
    pi = rng.dirichlet(np.ones(A))
    Q = S * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    return Q, pi


def test_integrals_match_quadrature():
    rng = np.random.default_rng(1)
    Q, pi = _gtr(rng, 4)
    T = 0.8
    I, M = hr.integrals(Q, pi, T)
    np.testing.assert_allclose(M, scipy.linalg.expm(Q * T), rtol=1e-12, atol=1e-14)
    x, w = np.polynomial.legendre.leggauss(40)
    s = (x + 1) * T / 2
    w = w * T / 2
    Iq = np.zeros_like(I)
    for sk, wk in zip(s, w):
        Ml, Mr = scipy.linalg.expm(Q * sk), scipy.linalg.expm(Q * (T - sk))
        Iq += wk * np.einsum("ai,jb->abij", Ml, Mr)
    np.testing.assert_allclose(I, Iq, rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("T", [0.05, 0.7, 2.0])
def test_weighted_stats_identities(T):
    rng = np.random.default_rng(2)
    A = 20
    Q, pi = _gtr(rng, A)
    mc = rng.uniform(0.0, 2.0, (A, A))
    W, U = hr.weighted_stats(Q, pi, T, mc)
    assert W.sum() == pytest.approx(T * mc.sum(), rel=1e-11)
    assert np.all(np.diag(U) == 0) and np.all(U >= -1e-12) and np.all(W >= -1e-12)
    h = 1e-6
    dlogM = (np.log(scipy.linalg.expm(Q * (T + h))) - np.log(scipy.linalg.expm(Q * (T - h)))) / (2 * h)
    lhs = U.sum() + (np.diag(Q) * W).sum()
    assert lhs == pytest.approx(T * (mc * dlogM).sum(), rel=1e-7)
