"""2D pair-HMM forward / backward / expected counts (oracle, numpy FP64).

A log-space restatement of python/tkfmixdom/jax/dp/hmm.py:
  * forward  -- _forward_2d_core_diag (hmm.py:1186-1292): F[0,0,0] = 0; for a
    state k of type M the predecessor cell is (i-1,j-1), type I (i,j-1),
    type D (i-1,j); F[i,j,k] = logsumexp_s(F[pred,s] + T[s,k]) + e[i,j,k];
    non-emitting states are NEG_INF off (0,0); logP = logsumexp_k(F[Lx,Ly,k]
    + T[k,E]) (hmm.py:1911-1913).
  * backward -- _backward_2d_core_diag (hmm.py:1295-1437): B[Lx,Ly,:] =
    T[:,E]; B[i,j,k] = logsumexp_{k'} T[k,k'] + e[succ,k'] + B[succ,k'] with
    succ = (i+1,j+1) / (i,j+1) / (i+1,j) by the type of k'.
  * counts / posteriors -- forward_backward_2d (hmm.py:1930-1971).

The lattice is the real (Lx+1, Ly+1) one: the reference's padded lattice
only adds NEG_INF emissions outside the real region, which contribute
exp(-1e30) = 0 to every output.  Vectorised along antidiagonals; intended
for pairs up to a few hundred residues.
"""
import numpy as np

from .tkf import M, I, D, E, NEG_INF


def _lse(a, axis):
    """jax.nn.logsumexp semantics on finite NEG_INF sentinels."""
    mx = np.max(a, axis=axis, keepdims=True)
    out = mx + np.log(np.sum(np.exp(a - mx), axis=axis, keepdims=True))
    return np.squeeze(out, axis=axis)


def e_index(state_types):
    """First E-type state (hmm.py:17-23)."""
    return int(np.argmax(np.asarray(state_types) == E))


def _diag_cells(d, Lx, Ly):
    i = np.arange(max(0, d - Ly), min(d, Lx) + 1)
    return i, d - i


def forward(log_trans, state_types, emit, dtype=np.float64):
    """Returns (logP, F) with F of shape (Lx+1, Ly+1, S).  ``dtype`` may be
    np.longdouble (80-bit on x86) for long pairs, where the float64 log-space
    recursion itself loses ~1e-16 * |log F| per step."""
    T = np.asarray(log_trans, dtype=dtype)
    emit = np.asarray(emit, dtype=dtype)
    st = np.asarray(state_types)
    Lx, Ly, ns = emit.shape[0] - 1, emit.shape[1] - 1, emit.shape[2]
    masks = [(M, st == M, 1, 1), (I, st == I, 0, 1), (D, st == D, 1, 0)]
    F = np.full((Lx + 1, Ly + 1, ns), NEG_INF, dtype=dtype)
    F[0, 0, 0] = 0.0
    for d in range(1, Lx + Ly + 1):
        i, j = _diag_cells(d, Lx, Ly)
        val = np.full((len(i), ns), NEG_INF, dtype=dtype)
        for _, kmask, di, dj in masks:
            ok = (i >= di) & (j >= dj)
            if not ok.any() or not kmask.any():
                continue
            pred = F[i[ok] - di, j[ok] - dj]                       # (n, ns)
            v = _lse(pred[:, :, None] + T[None, :, :], axis=1) + emit[i[ok], j[ok]]
            sub = val[ok]
            sub[:, kmask] = v[:, kmask]
            val[ok] = sub
        F[i, j] = val
    logP = _lse(F[Lx, Ly] + T[:, e_index(st)], axis=0)
    return logP, F


def backward(log_trans, state_types, emit, dtype=np.float64):
    T = np.asarray(log_trans, dtype=dtype)
    emit = np.asarray(emit, dtype=dtype)
    st = np.asarray(state_types)
    Lx, Ly, ns = emit.shape[0] - 1, emit.shape[1] - 1, emit.shape[2]
    e = e_index(st)
    B = np.full((Lx + 1, Ly + 1, ns), NEG_INF, dtype=dtype)
    B[Lx, Ly] = T[:, e]
    masks = [(st == M, 1, 1), (st == I, 0, 1), (st == D, 1, 0)]
    for d in range(Lx + Ly - 1, -1, -1):
        i, j = _diag_cells(d, Lx, Ly)
        acc = np.full((len(i), ns), NEG_INF, dtype=dtype)
        for kmask, di, dj in masks:
            ok = (i + di <= Lx) & (j + dj <= Ly)
            if not ok.any() or not kmask.any():
                continue
            si, sj = i[ok] + di, j[ok] + dj
            succ = emit[si, sj] + B[si, sj]                         # (n, ns)
            contrib = T[None, :, :] + succ[:, None, :]
            contrib = np.where(kmask[None, None, :], contrib, NEG_INF)
            v = _lse(contrib, axis=2)
            acc[ok] = np.logaddexp(acc[ok], v)
        B[i, j] = acc
    return B


def forward_backward(log_trans, state_types, emit, dtype=np.float64):
    """Returns (logP, posteriors (Lx+1,Ly+1,S), n (S,S), logP_backward) as
    float64 arrays (computed in ``dtype``)."""
    T = np.asarray(log_trans, dtype=dtype)
    emit = np.asarray(emit, dtype=dtype)
    st = np.asarray(state_types)
    logP, F = forward(T, st, emit, dtype)
    B = backward(T, st, emit, dtype)
    Lx, Ly, ns = emit.shape[0] - 1, emit.shape[1] - 1, emit.shape[2]
    post = np.exp(F + B - logP)
    n = np.zeros((ns, ns), dtype=dtype)
    for kmask, di, dj in [(st == M, 1, 1), (st == I, 0, 1), (st == D, 1, 0)]:
        if not kmask.any():
            continue
        Fs = F[:Lx + 1 - di, :Ly + 1 - dj]
        Gs = emit[di:, dj:] + B[di:, dj:]
        ln = (Fs[:, :, :, None] + T[None, None, :, :] + Gs[:, :, None, :] - logP)
        ln = np.where(kmask[None, None, None, :], ln, NEG_INF)
        n += np.exp(ln).sum(axis=(0, 1))
    e = e_index(st)
    n[:, e] += np.exp(F[Lx, Ly] + T[:, e] - logP)
    return (float(logP), post.astype(np.float64), n.astype(np.float64), float(B[0, 0, 0]))


def forward_backward_batch(log_trans, state_types, emits):
    """Per-pair loop; log_trans may be (S,S) or (B,S,S)."""
    out = []
    for b, em in enumerate(emits):
        T = log_trans if np.ndim(log_trans) == 2 else log_trans[b]
        out.append(forward_backward(T, state_types, em))
    return out
