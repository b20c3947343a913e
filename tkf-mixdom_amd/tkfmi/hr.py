"""Holmes-Rubin weighted statistics on the GPU (reference: core/ctmc.py:173-326,
python/train_pfam.py:1991-2046).

The symmetric eigendecomposition of each class's rate matrix is computed
once per call on the host (it does not depend on t); the per-(pair, class)
work -- M(t), the spectral J(t) factors and the weighted sums -- is one
libtkfmi kernel launch (tkfmi_hr_stats_batched).
"""
import numpy as np
import torch

from . import _abi, hmm


def _np64(x):
    return np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x, dtype=np.float64)


def eig_data(Qs, pis):
    """Per class: eigenvectors and eigenvalues of (S + S^T)/2 with
    S = D^1/2 Q D^-1/2 (ctmc.py:181-185)."""
    Qs, pis = _np64(Qs), _np64(pis)
    sp = np.sqrt(pis)
    S = Qs * (sp[:, :, None] / sp[:, None, :])
    S = (S + np.swapaxes(S, 1, 2)) / 2
    lam, V = np.linalg.eigh(S)
    return np.ascontiguousarray(V), np.ascontiguousarray(lam)


def per_pair_hr_stats(Qs, pis, t_array, mc_stack):
    """W (B,C,A) and U (B,C,A,A) for per-pair t and match counts
    mc_stack (B,C,A,A), classes (or domains) with rate matrices Qs (C,A,A)
    and equilibria pis (C,A) (train_pfam.py:1991-2046)."""
    L = _abi.lib()
    dev = hmm.device()
    Qs_n, pis_n = _np64(Qs), _np64(pis)
    C, A = pis_n.shape
    V, lam = eig_data(Qs_n, pis_n)
    t = torch.as_tensor(_np64(t_array).reshape(-1), device=dev)
    B = int(t.shape[0])
    mc = torch.as_tensor(mc_stack, dtype=torch.float64).to(dev).contiguous()
    if tuple(mc.shape) != (B, C, A, A):
        raise ValueError(f"mc_stack must be ({B}, {C}, {A}, {A})")
    dV, dl = torch.as_tensor(V, device=dev), torch.as_tensor(lam, device=dev)
    dpi, dQ = torch.as_tensor(pis_n, device=dev), torch.as_tensor(Qs_n, device=dev)
    W = torch.empty((B, C, A), dtype=torch.float64, device=dev)
    U = torch.empty((B, C, A, A), dtype=torch.float64, device=dev)
    if B:
        _abi.check(L.tkfmi_hr_stats_batched(B, C, A, _abi._ptr(dV), _abi._ptr(dl), _abi._ptr(dpi),
                                            _abi._ptr(dQ), _abi._ptr(t), _abi._ptr(mc),
                                            _abi._ptr(W), _abi._ptr(U),
                                            _abi.current_stream_ptr(dev)))
    return W, U


def holmes_rubin_weighted_stats(Q, pi, t, mc):
    """Single (Q, pi, t, mc): reference surface of ctmc.py:277-326."""
    W, U = per_pair_hr_stats(_np64(Q)[None], _np64(pi)[None], [float(t)],
                             torch.as_tensor(mc, dtype=torch.float64)[None, None])
    return W[0, 0], U[0, 0]
