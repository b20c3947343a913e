"""Per-group log emission tables for the kernel (torch, FP64, on device).

The kernel looks emissions up from small per-group tables instead of the
reference's materialised (Lx+1, Ly+1, S) lattice.  Each group holds a 21x21
match block, a 21-entry insert row and a 21-entry delete row of natural-log
values, built with exactly the reference's floors and wildcard padding:

* shared      -- pair_hmm_emissions (dp/hmm.py:628-670): log(x + 1e-30),
                 M = log pi[a] + log P[a,b], I = log pi[b], D = log pi[a].
* per-domain  -- pair_hmm_emissions_per_domain (dp/hmm.py:786-851):
                 log(max(x, 1e-30)); group = domain of the state.
* per-class   -- pair_hmm_emissions_per_class (dp/hmm.py:673-783): for each
                 (domain, fragment) a log-mixture over classes accumulated by a
                 sequential logaddexp from NEG_INF, class weight log NEG_INF for
                 zero weights; group = d * F + f.

Index A (= 20 for proteins) and above is the wildcard with log-prob 0
(dp/hmm.py:646-648).
"""
import torch

A_PAD = 21
TAB_SZ = A_PAD * A_PAD + 2 * A_PAD
NEG_INF = -1e30


def _as64(x, device):
    return torch.as_tensor(x, dtype=torch.float64, device=device)


def _pad_sub(log_sub):
    A = log_sub.shape[-1]
    out = torch.zeros(log_sub.shape[:-2] + (A_PAD, A_PAD), dtype=torch.float64, device=log_sub.device)
    out[..., :A, :A] = log_sub
    return out


def _pad_pi(log_pi):
    A = log_pi.shape[-1]
    out = torch.zeros(log_pi.shape[:-1] + (A_PAD,), dtype=torch.float64, device=log_pi.device)
    out[..., :A] = log_pi
    return out


def _pack(match, ins, dele):
    """match (...,21,21), ins/dele (...,21) -> (..., TAB_SZ)."""
    return torch.cat([match.reshape(match.shape[:-2] + (A_PAD * A_PAD,)), ins, dele], dim=-1)


def shared_tables(sub_matrix, pi, device=None):
    """(..., 1, TAB_SZ) tables for pair_hmm_emissions.  sub_matrix (...,A,A),
    pi (...,A) -- a leading batch dimension gives per-pair tables."""
    sub = _as64(sub_matrix, device)
    pi = _as64(pi, sub.device)
    log_sub = _pad_sub(torch.log(sub + 1e-30))
    log_pi = _pad_pi(torch.log(pi + 1e-30))
    match = log_pi[..., :, None] + log_sub
    return _pack(match, log_pi, log_pi).unsqueeze(-2)


def per_domain_tables(sub_matrices, pis, device=None):
    """(..., K, TAB_SZ) tables for pair_hmm_emissions_per_domain."""
    subs = _as64(sub_matrices, device)
    pis = _as64(pis, subs.device)
    log_sub = _pad_sub(torch.log(torch.clamp(subs, min=1e-30)))
    log_pi = _pad_pi(torch.log(torch.clamp(pis, min=1e-30)))
    match = log_pi[..., :, None] + log_sub
    return _pack(match, log_pi, log_pi)


def per_class_tables(class_subs, class_pis, class_dist, device=None):
    """(..., K*F, TAB_SZ) tables for pair_hmm_emissions_per_class.
    class_subs (...,C,A,A), class_pis (C,A) or (...,C,A), class_dist (K,F,C)."""
    csub = _as64(class_subs, device)
    cpi = _as64(class_pis, csub.device)
    cd = _as64(class_dist, csub.device)
    log_csub = _pad_sub(torch.log(torch.clamp(csub, min=1e-30)))     # (...,C,21,21)
    log_cpi = _pad_pi(torch.log(torch.clamp(cpi, min=1e-30)))        # (...,C,21)
    log_cd = torch.where(cd > 0, torch.log(torch.clamp(cd, min=1e-300)),
                         torch.full_like(cd, NEG_INF))                # (K,F,C)
    K, F, C = cd.shape
    batch = log_csub.shape[:-3]
    match = torch.full(batch + (K, F, A_PAD, A_PAD), NEG_INF, dtype=torch.float64, device=csub.device)
    single = torch.full(batch + (K, F, A_PAD), NEG_INF, dtype=torch.float64, device=csub.device)
    for c in range(C):
        lp = log_cpi[..., c, :]
        mc = lp[..., :, None] + log_csub[..., c, :, :]               # (...,21,21)
        match = torch.logaddexp(match, mc[..., None, None, :, :] + log_cd[:, :, c, None, None])
        single = torch.logaddexp(single, lp[..., None, None, :] + log_cd[:, :, c, None])
    tab = _pack(match, single, single)                               # (...,K,F,TAB)
    return tab.reshape(batch + (K * F, TAB_SZ))


def state_groups_shared(S):
    return [0] * S


def state_groups_per_domain(S, F):
    return [0, 0] + [(s - 2) // (5 * F) for s in range(2, S)]


def state_groups_per_class(S, F):
    out = [0, 0]
    for s in range(2, S):
        b = s - 2
        d = b // (5 * F)
        f = (b % (5 * F)) % F
        out.append(d * F + f)
    return out
