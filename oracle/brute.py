"""Brute-force path enumeration for tiny pairs (pins the oracle DP).

Enumerates every state path S -> k_1 ... k_n -> E whose emissions consume
exactly x (length Lx) and y (length Ly): a state of type M consumes one
character of each, I one of y, D one of x (dp/hmm.py:1120-1183 defines the
same recursion).  The path probability is
    P[S,k1] e(c1,k1) P[k1,k2] e(c2,k2) ... P[kn,E]
with c_m the lattice cell reached after m steps.  Summing gives P(x,y);
the probability-weighted transition tallies give the expected counts n.
Independent of the DP's antidiagonal bookkeeping, so it catches wrong
predecessor offsets, dropped boundary terms and transposed transitions.
"""
import numpy as np

from .tkf import M, I, D, E


def enumerate_paths(trans, state_types, emit_prob, Lx, Ly):
    """trans: (S,S) probabilities; emit_prob: (Lx+1,Ly+1,S) probabilities
    (emit_prob[i,j,k] is the emission on ENTERING cell (i,j) in state k).
    Returns (P, n (S,S), post (Lx+1,Ly+1,S))."""
    st = np.asarray(state_types)
    ns = len(st)
    e_idx = int(np.argmax(st == E))
    body = [k for k in range(ns) if st[k] in (M, I, D)]
    step = {M: (1, 1), I: (0, 1), D: (1, 0)}
    total = 0.0
    n = np.zeros((ns, ns))
    post = np.zeros((Lx + 1, Ly + 1, ns))

    def rec(i, j, k, p, hist):
        nonlocal total
        if i == Lx and j == Ly:
            pe = p * trans[k, e_idx]
            if pe != 0.0:
                total += pe
                prev = 0
                for (ci, cj, kk) in hist:
                    n[prev, kk] += pe
                    post[ci, cj, kk] += pe
                    prev = kk
                n[prev, e_idx] += pe
                if not hist:
                    post[0, 0, 0] += pe
        for k2 in body:
            di, dj = step[int(st[k2])]
            i2, j2 = i + di, j + dj
            if i2 > Lx or j2 > Ly:
                continue
            q = p * trans[k, k2] * emit_prob[i2, j2, k2]
            if q == 0.0:
                continue
            rec(i2, j2, k2, q, hist + [(i2, j2, k2)])

    rec(0, 0, 0, 1.0, [])
    if total > 0:
        n /= total
        post /= total
        post[0, 0, 0] = 1.0
    return total, n, post
