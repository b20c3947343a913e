"""Collapsed MixDom chi in torch (FP64, differentiable, batched over t).

Reference: models/mixdom.py -- state layout (:48-77), exit-type decomposed
effective transitions with the (I - T_nullnull)^-1 closure (:104-214) and the
nested chi (:252-423).  ``t`` may be a scalar or a (B,) tensor; the result is
(N, N) or (B, N, N), N = 2 + 5 K F.  Index 0 = SS, 1 = EE, body state
2 + d*5F + uv*F + f with uv in (MM, MI, MD, II, DD).
"""
import torch

from .params import S, M, I, D, E, tkf91_trans, tkf_beta

_UV_EMIT_TYPE = [M, I, D, I, D]
_UV_U = [M, M, M, I, D]
_UV_X = [M, I, D, I, D]
_IS_M_TYPE = [True, True, True, False, False]
_MID = [M, I, D]


def n_states(K, F):
    return 2 + 5 * K * F


def state_types(K, F):
    body = []
    for _d in range(K):
        for uv in range(5):
            body += [_UV_EMIT_TYPE[uv]] * F
    return [S, E] + body


def _f64(x, dev):
    return torch.as_tensor(x, dtype=torch.float64, device=dev) if not isinstance(x, torch.Tensor) \
        else x.to(device=dev, dtype=torch.float64)


def effective_trans_per_type(main_lam, main_mu, t, dom_lam, dom_mu, dom_w):
    """Returns T_exit_k (...,K,5,5) and T_eff (...,5,5) (mixdom.py:104-214)."""
    dev = t.device
    dom_lam, dom_mu, dom_w = (_f64(x, dev) for x in (dom_lam, dom_mu, dom_w))
    K = dom_lam.shape[0]
    bt = t.shape
    tau = tkf91_trans(_f64(main_lam, dev), _f64(main_mu, dev), t)     # (...,5,5)
    kappas = dom_lam / dom_mu                                          # (K,)
    betas = tkf_beta(dom_lam, dom_mu, t[..., None])                    # (...,K)
    z_t = (1.0 - kappas) * (1.0 - betas)                               # (...,K)
    z_0 = (1.0 - kappas).expand_as(z_t)
    tau_rows = torch.stack([tau[..., S, :], tau[..., M, :], tau[..., I, :], tau[..., D, :],
                            torch.zeros_like(tau[..., S, :])], dim=-2)  # (...,5,5)
    tau_MID = tau_rows[..., _MID]                                      # (...,5,3)
    w_rep = dom_w.repeat_interleave(3)                                 # (3K,)
    T_Nnull = tau_MID.repeat(*([1] * len(bt)), 1, K) * w_rep           # (...,5,3K)
    nonnull = torch.stack([1.0 - z_t, kappas.expand_as(z_t), kappas.expand_as(z_t)], dim=-1)  # (...,K,3)
    null_prob = torch.stack([z_t, z_0, z_0], dim=-1)                   # (...,K,3)
    tau_E = torch.stack([tau[..., M, E], tau[..., M, E], tau[..., D, E]], dim=-1)  # (...,3)
    T_nullN = torch.zeros(bt + (K, 3, 5), dtype=torch.float64, device=dev)
    for e_, col in enumerate(_MID):
        T_nullN[..., e_, col] = nonnull[..., e_]
        T_nullN[..., e_, E] = null_prob[..., e_] * tau_E[..., None, e_]
    T_nullN = T_nullN.reshape(bt + (3 * K, 5))
    tau_null_rows = torch.stack([tau[..., M, :], tau[..., M, :], tau[..., D, :]], dim=-2)
    tau_null_MID = tau_null_rows[..., _MID]                            # (...,3,3)
    z_flat = null_prob.reshape(bt + (3 * K,))
    idx = torch.arange(3 * K, device=dev)
    i_e, j_k, j_e = idx % 3, idx // 3, idx % 3
    T_nullnull = (z_flat[..., :, None] * dom_w[j_k][None, :]
                  * tau_null_MID[..., i_e[:, None], j_e[None, :]])
    eye = torch.eye(3 * K, dtype=torch.float64, device=dev)
    closure = torch.linalg.inv_ex(eye - T_nullnull)[0]  # no singularity check: it would wait for the GPU
    closure_3d = closure.reshape(bt + (3 * K, K, 3))
    T_nullN_3d = T_nullN.reshape(bt + (K, 3, 5))
    closure_exit = torch.einsum('...ike,...kev->...kiv', closure_3d, T_nullN_3d)
    T_exit_k = torch.einsum('...ui,...kiv->...kuv', T_Nnull, closure_exit)
    T_NN = torch.zeros(bt + (5, 5), dtype=torch.float64, device=dev)
    T_NN[..., :, E] = tau_rows[..., :, E]
    return T_exit_k, T_NN + T_exit_k.sum(dim=-3)


def _block_diag3(a, b, c):
    """Block-diagonal of three (...,n,n) blocks (n may differ)."""
    n = a.shape[-1] + b.shape[-1] + c.shape[-1]
    out = torch.zeros(a.shape[:-2] + (n, n), dtype=a.dtype, device=a.device)
    o = 0
    for blk in (a, b, c):
        q = blk.shape[-1]
        out[..., o:o + q, o:o + q] = blk
        o += q
    return out


def build_nested_trans(main_lam, main_mu, t, dom_lam, dom_mu, dom_w, frag_w, ext_rates):
    """Collapsed MixDom chi (mixdom.py:252-423).  ext_rates (K,F) is made
    diagonal (MixDom1 format, mixdom.py:284-287) or given as (K,F,F)."""
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t, dtype=torch.float64)
    t = t.to(torch.float64)
    dev = t.device
    dom_lam, dom_mu, dom_w, frag_w, ext = (_f64(x, dev) for x in
                                           (dom_lam, dom_mu, dom_w, frag_w, ext_rates))
    if ext.dim() == 2:
        ext = torch.diag_embed(ext)
    K, F = frag_w.shape
    bt = t.shape
    n_body = 5 * K * F
    N = 2 + n_body
    T_exit_k, T = effective_trans_per_type(main_lam, main_mu, t, dom_lam, dom_mu, dom_w)
    tau_all = tkf91_trans(dom_lam, dom_mu, t[..., None])               # (...,K,5,5)
    kappas = dom_lam / dom_mu
    beta_d = torch.clamp(1.0 - tau_all[..., S, E], min=1e-30)         # (...,K)
    isM = torch.tensor(_IS_M_TYPE, device=dev)
    entry_M = tau_all[..., S, :][..., _UV_X] / beta_d[..., None]      # (...,K,5)
    entry_factor = torch.where(isM, entry_M, torch.ones_like(entry_M))
    exit_inner_M = tau_all[..., :, E][..., _UV_X]                      # (...,K,5)
    exit_inner_ID = ((1.0 - kappas)[:, None]).expand(K, 5).expand_as(exit_inner_M)
    exit_inner = torch.where(isM, exit_inner_M, exit_inner_ID)
    notext = 1.0 - ext.sum(dim=-1)                                     # (K,F)
    exit_full = notext[:, None, :] * exit_inner[..., :, :, None]       # (...,K,5,F)
    exit_flat = exit_full.reshape(bt + (n_body,))
    dest_flat = (frag_w[:, None, :] * entry_factor[..., :, :, None]).reshape(bt + (n_body,))
    T_k_compound = T_exit_k[..., _UV_U, :][..., _UV_U]                 # (...,K,5,5)
    block = 5 * F
    d_idx = torch.arange(n_body, device=dev) // block
    suv = torch.arange(5, device=dev).repeat_interleave(F).repeat(K)
    T_expanded = T_k_compound[..., d_idx[None, :], suv[:, None], suv[None, :]]
    chi_inter = exit_flat[..., :, None] * T_expanded * dest_flat[..., None, :]
    tau_MID_all = tau_all[..., _MID, :][..., _MID]                     # (...,K,3,3)
    non_ext_M = torch.einsum('df,...dxy,dg->...dxfyg', notext, tau_MID_all,
                             frag_w).reshape(bt + (K, 3 * F, 3 * F))
    zF = torch.zeros((K, F, F), dtype=torch.float64, device=dev)
    ext_M = _block_diag3(ext, ext, ext)                                # (K,3F,3F)
    intra_M = non_ext_M + ext_M
    non_ext_I = notext[:, :, None] * kappas[:, None, None] * frag_w[:, None, :]
    intra_I = (non_ext_I + ext).expand(bt + (K, F, F))
    block_all = _block_diag3(intra_M, intra_I, intra_I)               # (...,K,5F,5F)
    del zF
    local = torch.arange(n_body, device=dev) % block
    same = (d_idx[:, None] == d_idx[None, :]).to(torch.float64)
    chi_intra = block_all[..., d_idx[:, None], local[:, None], local[None, :]] * same
    chi_body = chi_inter + chi_intra
    ss_T_k = T_exit_k[..., S, :][..., _UV_U]                           # (...,K,5)
    ss_body = (ss_T_k[..., :, :, None] * frag_w[:, None, :]
               * entry_factor[..., :, :, None]).reshape(bt + (n_body,))
    uvu_E = T[..., _UV_U, E]                                           # (...,5)
    ee_body = (exit_full * uvu_E[..., None, :, None]).reshape(bt + (n_body,))
    chi = torch.zeros(bt + (N, N), dtype=torch.float64, device=dev)
    chi[..., 0, 1] = T[..., S, E]
    chi[..., 0, 2:] = ss_body
    chi[..., 2:, 1] = ee_body
    chi[..., 2:, 2:] = chi_body
    return chi
