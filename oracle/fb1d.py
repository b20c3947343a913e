"""1D alignment-constrained pair-HMM forward / backward / expected counts
(oracle, numpy FP64).  Test infrastructure: only tests/, smoke() and the
bench's cpu_baseline leg use it, as the checker.

A log-space restatement of python/tkfmixdom/jax/dp/hmm.py:
  * forward_1d_scan (hmm.py:404-437): alpha_init = NEG_INF except 0 at the
    init state; alpha_0 = logsumexp_i(alpha_init[i] + T[i,:]) + e_0;
    alpha_t[j] = logsumexp_i(alpha_{t-1}[i] + T[i,j]) + e_t[j];
    logP = logsumexp_k(alpha_{L-1}[k] + T[k,final]).
  * forward_backward_1d_padded (hmm.py:471-607): the backward runs on
    T_bwd = T with T_bwd[final,final] = 0 in log space (hmm.py:532):
    beta_{L-1} = T_bwd[:,final] (backward_1d_scan hmm.py:440-468),
    beta_t[i] = logsumexp_j(T_bwd[i,j] + e_{t+1}[j] + beta_{t+1}[j]);
    posteriors = exp(alpha + beta - logP) (hmm.py:547);
    counts = sum_{t<L-1} exp(alpha_t[:,None] + T + e_{t+1} + beta_{t+1} - logP)
    (hmm.py:549-577), plus exp(alpha_{L-1} + T[:,final] - logP) in column
    final (hmm.py:579-590) and exp(T[init] + e_0 + beta_0 - logP) in row
    init (hmm.py:592-596); L = 0 gives logP = T[init,final] and
    n[init,final] = 1 (hmm.py:598-607).
  * pair_hmm_emissions_constrained (hmm.py:854-1100): each alignment column
    has a type (M=1, I=2, D=3); a state gets its type's emission when the
    types agree and NEG_INF otherwise.  Match = log pi[anc] + log P[anc,desc],
    insert = log pi[desc], delete = log pi[anc], per domain or as a
    class mixture, read here from the same per-group 21x21+21+21 tables the
    2D path uses (oracle/emissions.py).
  * alignment columns: anc_full[M|D columns] = anc_chars, desc_full[M|I
    columns] = desc_chars (python/train_pfam.py:2177-2182).

Pinned (tests/test_oracle_fb1d.py) by brute-force enumeration of state
paths, and by the identity that summing the 1D probability over every
alignment of a pair gives the 2D pair-HMM probability (and the
alignment-weighted 1D counts the 2D counts), against the separately pinned
2D oracle.
"""
import itertools

import numpy as np

from .emissions import A_PAD, TAB_SZ
from .tkf import M, I, D, NEG_INF


def _lse(a, axis):
    mx = np.max(a, axis=axis, keepdims=True)
    out = mx + np.log(np.sum(np.exp(a - mx), axis=axis, keepdims=True))
    return np.squeeze(out, axis=axis)


def forward_backward_1d(log_trans, log_emit, init_state, final_state, forward_only=False):
    """log_trans (S,S), log_emit (L,S).  Returns logP, or (logP, post (L,S),
    n (S,S))."""
    T = np.asarray(log_trans, dtype=np.float64)
    e = np.asarray(log_emit, dtype=np.float64).reshape(-1, T.shape[0])
    L, ns = e.shape
    if L == 0:
        lp = float(T[init_state, final_state])
        if forward_only:
            return lp
        n = np.zeros((ns, ns))
        n[init_state, final_state] = 1.0
        return lp, np.zeros((0, ns)), n
    alpha = np.empty((L, ns))
    alpha[0] = T[init_state] + e[0]
    for t in range(1, L):
        alpha[t] = _lse(alpha[t - 1][:, None] + T, axis=0) + e[t]
    lp = float(_lse(alpha[L - 1] + T[:, final_state], axis=0))
    if forward_only:
        return lp
    Tb = T.copy()
    Tb[final_state, final_state] = 0.0
    beta = np.empty((L, ns))
    beta[L - 1] = Tb[:, final_state]
    for t in range(L - 2, -1, -1):
        beta[t] = _lse(Tb + e[t + 1][None, :] + beta[t + 1][None, :], axis=1)
    post = np.exp(alpha + beta - lp)
    n = np.zeros((ns, ns))
    for t in range(L - 1):
        n += np.exp(alpha[t][:, None] + T + e[t + 1][None, :] + beta[t + 1][None, :] - lp)
    n[:, final_state] += np.exp(alpha[L - 1] + T[:, final_state] - lp)
    n[init_state, :] += np.exp(T[init_state] + e[0] + beta[0] - lp)
    return lp, post, n


def columns_from_alignment(states, anc_chars, desc_chars):
    """(col_types, anc_full, desc_full) from a column-type sequence and the
    compact character lists (python/train_pfam.py:2177-2182)."""
    st = np.asarray(states, dtype=np.int64)
    L = len(st)
    anc_full = np.zeros(L, dtype=np.int64)
    desc_full = np.zeros(L, dtype=np.int64)
    m_or_d = (st == M) | (st == D)
    m_or_i = (st == M) | (st == I)
    anc_full[m_or_d] = np.asarray(anc_chars, dtype=np.int64)
    desc_full[m_or_i] = np.asarray(desc_chars, dtype=np.int64)
    return st, anc_full, desc_full


def constrained_emissions(state_types, state_group, log_tab, col_types, anc_full, desc_full):
    """(L, S) log emissions: a state of the column's type gets its group's
    table entry (match (anc, desc), insert desc, delete anc; codes clipped to
    the wildcard 20), every other state NEG_INF."""
    st = np.asarray(state_types)
    grp = np.asarray(state_group, dtype=np.int64)
    tab = np.asarray(log_tab, dtype=np.float64).reshape(-1, TAB_SZ)
    ct = np.asarray(col_types, dtype=np.int64)
    a = np.minimum(np.asarray(anc_full, dtype=np.int64), A_PAD - 1)
    b = np.minimum(np.asarray(desc_full, dtype=np.int64), A_PAD - 1)
    L, ns = len(ct), len(st)
    out = np.full((L, ns), NEG_INF)
    for t in range(L):
        if ct[t] == M:
            idx = a[t] * A_PAD + b[t]
        elif ct[t] == I:
            idx = A_PAD * A_PAD + b[t]
        elif ct[t] == D:
            idx = A_PAD * A_PAD + A_PAD + a[t]
        else:
            continue
        for k in range(ns):
            if st[k] == ct[t]:
                out[t, k] = tab[grp[k], idx]
    return out


def emission_counts_1d(post, state_types, state_group, col_types, anc_full, desc_full, n_groups):
    """Expected emission counts (n_groups, TAB_SZ) in the table layout: the
    posterior of each state of the column's type, added onto the table entry
    its emission used."""
    st = np.asarray(state_types)
    grp = np.asarray(state_group, dtype=np.int64)
    ct = np.asarray(col_types, dtype=np.int64)
    a = np.minimum(np.asarray(anc_full, dtype=np.int64), A_PAD - 1)
    b = np.minimum(np.asarray(desc_full, dtype=np.int64), A_PAD - 1)
    out = np.zeros((n_groups, TAB_SZ))
    for t in range(len(ct)):
        if ct[t] == M:
            idx = a[t] * A_PAD + b[t]
        elif ct[t] == I:
            idx = A_PAD * A_PAD + b[t]
        elif ct[t] == D:
            idx = A_PAD * A_PAD + A_PAD + a[t]
        else:
            continue
        for k in range(len(st)):
            if st[k] == ct[t]:
                out[grp[k], idx] += post[t, k]
    return out


def brute_1d(trans, emit_prob, init_state, final_state):
    """Enumerates every state path init -> k_0 ... k_{L-1} -> final with
    probability trans[init,k_0] e_0(k_0) trans[k_0,k_1] ... trans[k_{L-1},final].
    trans (S,S) and emit_prob (L,S) are probabilities.  Returns (P, n, post)."""
    trans = np.asarray(trans, dtype=np.float64)
    emit_prob = np.asarray(emit_prob, dtype=np.float64).reshape(-1, trans.shape[0])
    L, ns = emit_prob.shape
    n = np.zeros((ns, ns))
    post = np.zeros((L, ns))
    total = 0.0
    for path in itertools.product(range(ns), repeat=L):
        p = 1.0
        prev = init_state
        for t, k in enumerate(path):
            p *= trans[prev, k] * emit_prob[t, k]
            prev = k
            if p == 0.0:
                break
        p *= trans[prev, final_state]
        if p == 0.0:
            continue
        total += p
        prev = init_state
        for t, k in enumerate(path):
            n[prev, k] += p
            post[t, k] += p
            prev = k
        n[prev, final_state] += p
    if total > 0:
        n /= total
        post /= total
    return total, n, post


def alignments(Lx, Ly):
    """Every column-type sequence (M=1, I=2, D=3) that consumes Lx ancestor
    and Ly descendant characters, with the characters' positions."""
    out = []

    def rec(i, j, cols):
        if i == Lx and j == Ly:
            out.append(list(cols))
            return
        if i < Lx and j < Ly:
            rec(i + 1, j + 1, cols + [(M, i, j)])
        if j < Ly:
            rec(i, j + 1, cols + [(I, i, j)])
        if i < Lx:
            rec(i + 1, j, cols + [(D, i, j)])

    rec(0, 0, [])
    return out


def alignment_columns(aln, x, y):
    """(col_types, anc_full, desc_full) of one alignment from alignments()."""
    ct = np.array([c[0] for c in aln], dtype=np.int64)
    anc = np.array([x[c[1]] if c[0] in (M, D) else 0 for c in aln], dtype=np.int64)
    desc = np.array([y[c[2]] if c[0] in (M, I) else 0 for c in aln], dtype=np.int64)
    return ct, anc, desc
