"""Log-emission tables (oracle restatement, numpy FP64).

Follows python/tkfmixdom/jax/dp/hmm.py:
  * pair_hmm_emissions (hmm.py:628-670): log(x + 1e-30) floors, wildcard
    index 20 padded with log-prob 0, grid row/col 0 use dummy character 0.
  * pair_hmm_emissions_per_domain (hmm.py:786-851): log(max(x, 1e-30)),
    state domain = (s - 2) // (5F).
  * pair_hmm_emissions_per_class (hmm.py:673-783): log-mixture over classes
    by a sequential logaddexp starting from NEG_INF, class weight log is
    NEG_INF for zero weights.
Non-emitting (S/E-type) states get 0.
"""
import numpy as np

from .tkf import M, I, D, NEG_INF

A_PAD = 21


def _pad_log_sub(log_sub):
    out = np.zeros(log_sub.shape[:-2] + (A_PAD, A_PAD))
    out[..., :log_sub.shape[-2], :log_sub.shape[-1]] = log_sub
    return out


def _pad_log_pi(log_pi):
    out = np.zeros(log_pi.shape[:-1] + (A_PAD,))
    out[..., :log_pi.shape[-1]] = log_pi
    return out


def _grid_chars(x, y):
    x_pad = np.concatenate([[0], np.asarray(x, dtype=np.int64)])
    y_pad = np.concatenate([[0], np.asarray(y, dtype=np.int64)])
    return x_pad, y_pad


def _assemble(state_types, match, ins, dele):
    """match (Lx+1,Ly+1,ns), ins (Ly+1,ns), dele (Lx+1,ns)."""
    st = np.asarray(state_types)
    isM, isI, isD = st == M, st == I, st == D
    emit = (isM[None, None, :] * match + isI[None, None, :] * ins[None, :, :]
            + isD[None, None, :] * dele[:, None, :])
    return np.where((isM | isI | isD)[None, None, :], emit, 0.0)


def log_tables_shared(sub_matrix, pi):
    """(log_sub (21,21), log_pi (21,)) as in hmm.py:643-648."""
    log_sub = _pad_log_sub(np.log(np.asarray(sub_matrix, dtype=np.float64) + 1e-30))
    log_pi = _pad_log_pi(np.log(np.asarray(pi, dtype=np.float64) + 1e-30))
    return log_sub, log_pi


def pair_hmm_emissions(state_types, x, y, sub_matrix, pi):
    log_sub, log_pi = log_tables_shared(sub_matrix, pi)
    xp, yp = _grid_chars(x, y)
    ns = len(state_types)
    match = (log_pi[xp][:, None] + log_sub[xp[:, None], yp[None, :]])[:, :, None]
    match = np.broadcast_to(match, match.shape[:2] + (ns,))
    ins = np.broadcast_to(log_pi[yp][:, None], (len(yp), ns))
    dele = np.broadcast_to(log_pi[xp][:, None], (len(xp), ns))
    return _assemble(state_types, match, ins, dele)


def log_tables_per_domain(sub_matrices, pis):
    """(log_subs (K,21,21), log_pis (K,21)) as in hmm.py:814-819."""
    log_subs = _pad_log_sub(np.log(np.maximum(np.asarray(sub_matrices, dtype=np.float64), 1e-30)))
    log_pis = _pad_log_pi(np.log(np.maximum(np.asarray(pis, dtype=np.float64), 1e-30)))
    return log_subs, log_pis


def pair_hmm_emissions_per_domain(state_types, x, y, sub_matrices, pis, K, F):
    log_subs, log_pis = log_tables_per_domain(sub_matrices, pis)
    xp, yp = _grid_chars(x, y)
    ns = len(state_types)
    dom = np.zeros(ns, dtype=np.int64)
    dom[2:] = np.arange(ns - 2) // (5 * F)
    slp = log_pis[dom]                      # (ns, 21)
    sls = log_subs[dom]                     # (ns, 21, 21)
    match = (slp[:, xp][:, :, None] + sls[:, xp][:, :, yp]).transpose(1, 2, 0)
    ins = slp[:, yp].T
    dele = slp[:, xp].T
    return _assemble(state_types, match, ins, dele)


def log_mixture_tables_per_class(class_subs, class_pis, class_dist):
    """Per-(d,f) log-mixture tables, in the reference's class order:
    match[d,f,a,b] = logaddexp-scan_c(log w[d,f,c] + log pi_c[a] + log P_c[a,b]),
    ins/del[d,f,a] = logaddexp-scan_c(log w[d,f,c] + log pi_c[a])
    (hmm.py:702-772).  Returns (K,F,21,21), (K,F,21)."""
    csubs = _pad_log_sub(np.log(np.maximum(np.asarray(class_subs, dtype=np.float64), 1e-30)))
    cpis = _pad_log_pi(np.log(np.maximum(np.asarray(class_pis, dtype=np.float64), 1e-30)))
    cd = np.asarray(class_dist, dtype=np.float64)
    log_cd = np.where(cd > 0, np.log(np.maximum(cd, 1e-300)), NEG_INF)
    K, F, C = cd.shape
    match = np.full((K, F, A_PAD, A_PAD), NEG_INF)
    single = np.full((K, F, A_PAD), NEG_INF)
    for c in range(C):
        mc = cpis[c][:, None] + csubs[c]                     # (21,21)
        match = np.logaddexp(match, mc[None, None] + log_cd[:, :, c, None, None])
        single = np.logaddexp(single, cpis[c][None, None, :] + log_cd[:, :, c, None])
    return match, single


def pair_hmm_emissions_per_class(state_types, x, y, class_subs, class_pis,
                                 class_dist, K, F):
    """Per-class emissions.  The mixture is formed per grid cell exactly as
    the reference scan (hmm.py:746-772)."""
    csubs = _pad_log_sub(np.log(np.maximum(np.asarray(class_subs, dtype=np.float64), 1e-30)))
    cpis = _pad_log_pi(np.log(np.maximum(np.asarray(class_pis, dtype=np.float64), 1e-30)))
    cd = np.asarray(class_dist, dtype=np.float64)
    log_cd = np.where(cd > 0, np.log(np.maximum(cd, 1e-300)), NEG_INF)
    C = cd.shape[2]
    xp, yp = _grid_chars(x, y)
    ns = len(state_types)
    dom = np.zeros(ns, dtype=np.int64)
    frag = np.zeros(ns, dtype=np.int64)
    body = np.arange(ns - 2)
    dom[2:] = body // (5 * F)
    frag[2:] = (body % (5 * F)) % F
    slw = log_cd[dom, frag, :]              # (ns, C)
    Lp, Lq = len(xp), len(yp)
    match = np.full((Lp, Lq, ns), NEG_INF)
    ins = np.full((Lq, ns), NEG_INF)
    dele = np.full((Lp, ns), NEG_INF)
    for c in range(C):
        mpc = cpis[c][xp][:, None] + csubs[c][xp[:, None], yp[None, :]]
        match = np.logaddexp(match, mpc[:, :, None] + slw[None, None, :, c])
        ins = np.logaddexp(ins, cpis[c][yp][:, None] + slw[None, :, c])
        dele = np.logaddexp(dele, cpis[c][xp][:, None] + slw[None, :, c])
    return _assemble(state_types, match, ins, dele)


TAB_SZ = A_PAD * A_PAD + 2 * A_PAD


def emit_from_tables(state_types, state_group, log_tab, x, y):
    """Emission lattice (Lx+1, Ly+1, S) from per-group tables laid out as
    [21x21 match | 21 insert | 21 delete] (the kernel's table format).  Grid
    row/column 0 use the dummy character 0 (hmm.py:650-651); codes are
    clipped to the wildcard index 20."""
    st = np.asarray(state_types)
    grp = np.asarray(state_group, dtype=np.int64)
    tab = np.asarray(log_tab, dtype=np.float64)
    xp, yp = _grid_chars(np.minimum(np.asarray(x, dtype=np.int64), A_PAD - 1),
                         np.minimum(np.asarray(y, dtype=np.int64), A_PAD - 1))
    mt = tab[:, :A_PAD * A_PAD].reshape(-1, A_PAD, A_PAD)
    it = tab[:, A_PAD * A_PAD:A_PAD * A_PAD + A_PAD]
    dt = tab[:, A_PAD * A_PAD + A_PAD:]
    match = mt[grp][:, xp][:, :, yp].transpose(1, 2, 0)   # (Lx+1, Ly+1, S)
    ins = it[grp][:, yp].T                                 # (Ly+1, S)
    dele = dt[grp][:, xp].T                                # (Lx+1, S)
    return _assemble(st, match, ins, dele)


def emission_counts(post, state_types, state_group, x, y, n_groups):
    """Expected emission counts (n_groups, TAB_SZ) in the table layout, from
    a posterior lattice (Lx+1, Ly+1, S).  The per-type slices follow
    observation_counts_2d (train/likelihood.py:962-1028): M over
    i in [1,Lx], j in [1,Ly] onto (x[i-1], y[j-1]); I over i in [0,Lx],
    j in [1,Ly] onto y[j-1]; D over i in [1,Lx], j in [0,Ly] onto x[i-1];
    per-group (domain) masks as in train_pfam.py:1705-1734."""
    st = np.asarray(state_types)
    grp = np.asarray(state_group, dtype=np.int64)
    post = np.asarray(post, dtype=np.float64)
    x = np.minimum(np.asarray(x, dtype=np.int64), A_PAD - 1)
    y = np.minimum(np.asarray(y, dtype=np.int64), A_PAD - 1)
    Lx, Ly = len(x), len(y)
    out = np.zeros((n_groups, TAB_SZ))
    for s in range(len(st)):
        g = grp[s]
        if st[s] == M and Lx and Ly:
            w = post[1:, 1:, s]                                         # (Lx, Ly)
            blk = np.zeros((A_PAD, A_PAD))
            np.add.at(blk, (x[:, None].repeat(Ly, 1), y[None, :].repeat(Lx, 0)), w)
            out[g, :A_PAD * A_PAD] += blk.ravel()
        elif st[s] == I and Ly:
            np.add.at(out[g], A_PAD * A_PAD + y, post[:, 1:, s].sum(axis=0))
        elif st[s] == D and Lx:
            np.add.at(out[g], A_PAD * A_PAD + A_PAD + x, post[1:, :, s].sum(axis=1))
    return out
