"""Collapsed MixDom chi (oracle restatement, numpy FP64).

Follows python/tkfmixdom/jax/models/mixdom.py:
  * state layout: 0 = SS, 1 = EE, body index 2 + d*5F + uv*F + f with
    uv in (MM, MI, MD, II, DD) -> emission types (M, I, D, I, D)
    (mixdom.py:48-77, 232-239)
  * effective_trans_per_type: exit-type decomposition of the null fan-out,
    closure (I - T_nullnull)^-1 (mixdom.py:104-214)
  * build_nested_trans: inter-domain block, intra-domain M/I/D blocks with
    Markov fragment transitions, SS row, EE column (mixdom.py:252-423)
"""
import numpy as np

from .tkf import S, M, I, D, E, tkf91_trans, tkf_beta

_UV_EMIT_TYPE = np.array([M, I, D, I, D])
_UV_U = np.array([M, M, M, I, D])
_UV_X = np.array([M, I, D, I, D])
_IS_M_TYPE = np.array([True, True, True, False, False])
_MID = np.array([M, I, D])


def n_states(K, F):
    return 2 + 5 * K * F


def state_types(K, F):
    """mixdom.py:69-77."""
    body = np.tile(np.repeat(_UV_EMIT_TYPE, F), K)
    return np.concatenate([np.array([S, E]), body]).astype(np.int32)


def state_dom_frag(K, F):
    """(domain, fragment) of every state; SS/EE map to (0, 0)
    (dp/hmm.py:719-729)."""
    ns = n_states(K, F)
    body = np.arange(ns - 2)
    dom = np.zeros(ns, dtype=np.int64)
    frag = np.zeros(ns, dtype=np.int64)
    dom[2:] = body // (5 * F)
    frag[2:] = (body % (5 * F)) % F
    return dom, frag


def effective_trans_per_type(main_lam, main_mu, t, dom_lam, dom_mu, dom_w):
    """mixdom.py:104-214.  Returns (T_exit_k (K,5,5), T_eff (5,5))."""
    dom_lam = np.asarray(dom_lam, dtype=np.float64)
    dom_mu = np.asarray(dom_mu, dtype=np.float64)
    dom_w = np.asarray(dom_w, dtype=np.float64)
    K = dom_lam.shape[0]
    tau = tkf91_trans(main_lam, main_mu, t)
    kappas = dom_lam / dom_mu
    betas = tkf_beta(dom_lam, dom_mu, t)
    z_t = (1.0 - kappas) * (1.0 - betas)
    z_0 = 1.0 - kappas
    tau_rows = np.stack([tau[S], tau[M], tau[I], tau[D], np.zeros(5)])
    tau_MID = tau_rows[:, _MID]
    w_rep = np.repeat(dom_w, 3)
    T_Nnull = np.tile(tau_MID, (1, K)) * w_rep[None, :]
    nonnull = np.stack([1.0 - z_t, kappas, kappas], axis=1)
    null_prob = np.stack([z_t, z_0, z_0], axis=1)
    tau_E = np.stack([np.full(K, tau[M, E]), np.full(K, tau[M, E]),
                      np.full(K, tau[D, E])], axis=1)
    T_nullN = np.zeros((3 * K, 5))
    exit_col = [M, I, D]
    for e in range(3):
        rows = np.arange(K) * 3 + e
        T_nullN[rows, exit_col[e]] = nonnull[:, e]
        T_nullN[rows, E] = null_prob[:, e] * tau_E[:, e]
    tau_null_rows = np.stack([tau[M], tau[M], tau[D]])
    tau_null_MID = tau_null_rows[:, _MID]
    z_flat = null_prob.reshape(3 * K)
    idx = np.arange(3 * K)
    i_e, j_k, j_e = idx % 3, idx // 3, idx % 3
    T_nullnull = (z_flat[:, None] * dom_w[j_k][None, :]
                  * tau_null_MID[i_e[:, None], j_e[None, :]])
    closure = np.linalg.inv(np.eye(3 * K) - T_nullnull)
    closure_3d = closure.reshape(3 * K, K, 3)
    T_nullN_3d = T_nullN.reshape(K, 3, 5)
    closure_exit = np.einsum('ike,kev->kiv', closure_3d, T_nullN_3d)
    T_exit_k = np.einsum('ui,kiv->kuv', T_Nnull, closure_exit)
    T_NN = np.zeros((5, 5))
    T_NN[:, E] = tau_rows[:, E]
    return T_exit_k, T_NN + T_exit_k.sum(axis=0)


def _block_diag(*blocks):
    n = sum(b.shape[0] for b in blocks)
    out = np.zeros((n, n))
    o = 0
    for b in blocks:
        k = b.shape[0]
        out[o:o + k, o:o + k] = b
        o += k
    return out


def build_nested_trans(main_lam, main_mu, t, dom_lam, dom_mu, dom_w,
                       frag_w, ext_rates):
    """Collapsed MixDom chi, mixdom.py:252-423.  ext_rates may be (K,F)
    (MixDom1, made diagonal as at mixdom.py:284-287) or (K,F,F)."""
    dom_lam = np.asarray(dom_lam, dtype=np.float64)
    dom_mu = np.asarray(dom_mu, dtype=np.float64)
    dom_w = np.asarray(dom_w, dtype=np.float64)
    frag_w = np.asarray(frag_w, dtype=np.float64)
    ext = np.asarray(ext_rates, dtype=np.float64)
    if ext.ndim == 2:
        ext = np.stack([np.diag(e) for e in ext])
    K = dom_lam.shape[0]
    F = frag_w.shape[1]
    n_body = 5 * K * F
    N = 2 + n_body
    T_exit_k, T = effective_trans_per_type(main_lam, main_mu, t,
                                           dom_lam, dom_mu, dom_w)
    tau_all = np.stack([tkf91_trans(dom_lam[d], dom_mu[d], t) for d in range(K)])
    kappas = dom_lam / dom_mu
    beta_d = np.maximum(1.0 - tau_all[:, S, E], 1e-30)
    entry_M = tau_all[:, S, :][:, _UV_X] / beta_d[:, None]
    entry_factor = np.where(_IS_M_TYPE[None, :], entry_M, np.ones((K, 5)))
    exit_inner_M = tau_all[:, :, E][:, _UV_X]
    exit_inner_ID = (1.0 - kappas)[:, None] * np.ones(5)[None, :]
    exit_inner = np.where(_IS_M_TYPE[None, :], exit_inner_M, exit_inner_ID)
    notext = 1.0 - ext.sum(axis=-1)
    exit_full = notext[:, None, :] * exit_inner[:, :, None]       # (K,5,F)
    exit_flat = exit_full.reshape(n_body)
    dest_flat = (frag_w[:, None, :] * entry_factor[:, :, None]).reshape(n_body)
    T_k_compound = T_exit_k[:, _UV_U][:, :, _UV_U]
    block = 5 * F
    d_idx = np.arange(n_body) // block
    suv = np.tile(np.repeat(np.arange(5), F), K)
    T_expanded = T_k_compound[d_idx[None, :], suv[:, None], suv[None, :]]
    chi_inter = exit_flat[:, None] * T_expanded * dest_flat[None, :]
    tau_MID_all = tau_all[:, _MID][:, :, _MID]
    non_ext_M = np.einsum('df,dxy,dg->dxfyg', notext, tau_MID_all,
                          frag_w).reshape(K, 3 * F, 3 * F)
    ext_M = np.stack([_block_diag(e, e, e) for e in ext])
    intra_M = non_ext_M + ext_M
    non_ext_I = notext[:, :, None] * kappas[:, None, None] * frag_w[:, None, :]
    intra_I = non_ext_I + ext
    intra_D = non_ext_I + ext
    block_all = np.stack([_block_diag(intra_M[d], intra_I[d], intra_D[d])
                          for d in range(K)])
    local = np.arange(n_body) % block
    same = d_idx[:, None] == d_idx[None, :]
    chi_intra = block_all[d_idx[:, None], local[:, None], local[None, :]] * same
    chi_body = chi_inter + chi_intra
    ss_T_k = T_exit_k[:, S, :][:, _UV_U]
    ss_body = (ss_T_k[:, :, None] * frag_w[:, None, :]
               * entry_factor[:, :, None]).reshape(n_body)
    ee_body = (exit_full * T[_UV_U, E][None, :, None]).reshape(n_body)
    chi = np.zeros((N, N))
    chi[0, 1] = T[S, E]
    chi[0, 2:] = ss_body
    chi[2:, 1] = ee_body
    chi[2:, 2:] = chi_body
    return chi
