"""Null-state count restoration, sufficient statistics and BDI statistics on
the GPU (reference: models/exact_suffstats_batch.py, models/exact_suffstats.py,
core/bdi.py).

Same call surfaces as the reference:

* ``exact_suffstats_batch`` (exact_suffstats_batch.py:485-541)
* ``exact_suffstats_per_pair_batch`` (exact_suffstats.py:266-386)
* ``tkf91_stats_from_counts_batch`` / ``bdi_stats_from_counts_batch``
  (core/bdi.py:800-888)

All of them run one libtkfmi kernel launch over the batch (FP64, one
workgroup per pair); inputs may be numpy or torch, outputs are FP64 torch
tensors on the device (the per-pair-batch function returns summed host
values like the reference).
"""
import ctypes

import numpy as np
import torch

from . import _abi, hmm

S, M, I, D, E = 0, 1, 2, 3, 4


def _dev64(x, dev):
    return torch.as_tensor(x, dtype=torch.float64).to(dev).contiguous()


class _Model:
    """Host copies of the shared structural parameters + the C struct."""

    def __init__(self, main_ins, main_del, dom_ins, dom_del, dom_weights, frag_weights, ext_rates):
        f64 = lambda x: np.ascontiguousarray(
            np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x, dtype=np.float64))
        self.dom_ins, self.dom_del = f64(dom_ins).ravel(), f64(dom_del).ravel()
        self.dom_w = f64(dom_weights).ravel()
        self.frag_w = f64(frag_weights)
        K, F = self.frag_w.shape
        ext = f64(ext_rates)
        if ext.ndim == 2:  # MixDom1 (K,F) rates -> diagonal (K,F,F) (models/mixdom.py:284-287)
            ext = np.stack([np.diag(e) for e in ext])
        self.ext = np.ascontiguousarray(ext)
        if self.ext.shape != (K, F, F) or self.dom_ins.shape != (K,):
            raise ValueError("inconsistent MixDom parameter shapes")
        self.K, self.F = K, F
        self.c = _abi.MixdomModel(K, F, float(main_ins), float(main_del),
                                  _abi._np_ptr(self.dom_ins), _abi._np_ptr(self.dom_del),
                                  _abi._np_ptr(self.dom_w), _abi._np_ptr(self.frag_w),
                                  _abi._np_ptr(self.ext))


def _split_stats(st, K, F):
    B = st.shape[0]
    o = 0

    def take(n, shape):
        nonlocal o
        v = st[:, o:o + n].reshape((B,) + shape)
        o += n
        return v
    top = take(25, (5, 5))
    domM = take(25 * K, (K, 5, 5))
    out = {"top_5x5": top, "dom_M_5x5": [domM[:, k] for k in range(K)],
           "dom_kappa": take(K, (K,)), "dom_1mkappa": take(K, (K,)),
           "ext": take(K * F * F, (K, F, F)), "term": take(K * F, (K, F)),
           "dom_w": take(K, (K,)), "frag_w": take(K * F, (K, F))}
    return out


def exact_suffstats_batch(n_chi_stack, main_ins, main_del, t_array, dom_ins, dom_del,
                          dom_weights, frag_weights, ext_rates, *, with_bdi=False):
    """Per-pair restored sufficient statistics (exact_suffstats_batch.py:485-541).

    n_chi_stack: (B, N, N) collapsed expected counts (N = 2 + 5KF), e.g. the
    ``n`` output of the forward/backward kernel; t_array: (B,).  Returns the
    reference's dict with leading B axes (dom_M_5x5 a list of K (B,5,5));
    with ``with_bdi`` also 'bdi' (B, 1+K, 4): E[B], E[D], E[S], T per level
    (top, then each domain)."""
    L = _abi.lib()
    m = _Model(main_ins, main_del, dom_ins, dom_del, dom_weights, frag_weights, ext_rates)
    dev = hmm.device()
    n = _dev64(n_chi_stack, dev)
    t = _dev64(t_array, dev).reshape(-1)
    B = int(t.shape[0])
    Nc = 2 + 5 * m.K * m.F
    if tuple(n.shape) != (B, Nc, Nc):
        raise ValueError(f"n_chi_stack must be ({B}, {Nc}, {Nc})")
    nstat = int(L.tkfmi_suffstats_size(m.K, m.F))
    stats = torch.zeros((B, nstat), dtype=torch.float64, device=dev)
    bdi = torch.zeros((B, m.K + 1, 4), dtype=torch.float64, device=dev) if with_bdi else None
    wsb = int(L.tkfmi_restore_workspace_bytes(ctypes.byref(m.c), B))
    if wsb == 0:  # invalid model: the call reports why
        L.tkfmi_restore_counts_batched(ctypes.byref(m.c), B, None, None, None, None, None, 0, None)
        raise _abi.TkfmiError(L.tkfmi_last_error().decode() or "invalid MixDom model")
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    _abi.check(L.tkfmi_restore_counts_batched(ctypes.byref(m.c), B, _abi._ptr(t), _abi._ptr(n),
                                              _abi._ptr(stats), _abi._ptr(bdi), _abi._ptr(ws),
                                              ctypes.c_size_t(wsb), _abi.current_stream_ptr(dev)))
    out = _split_stats(stats, m.K, m.F)
    if with_bdi:
        out["bdi"] = bdi
    return out


def exact_suffstats_per_pair_batch(per_pair_records, main_ins, main_del, dom_ins, dom_del,
                                   dom_weights, frag_weights, ext_rates, accumulate_bdi=True):
    """Summed restored statistics over (n_chi_p, t_p) records, plus the
    summed post-divide BDI statistics (exact_suffstats.py:266-386)."""
    fw = np.asarray(frag_weights.detach().cpu() if isinstance(frag_weights, torch.Tensor)
                    else frag_weights)
    K, F = fw.shape
    recs = list(per_pair_records)
    if not recs:
        acc = {"top_5x5": np.zeros((5, 5)), "dom_M_5x5": [np.zeros((5, 5)) for _ in range(K)],
               "dom_kappa": np.zeros(K), "dom_1mkappa": np.zeros(K), "ext": np.zeros((K, F, F)),
               "term": np.zeros((K, F)), "dom_w": np.zeros(K), "frag_w": np.zeros((K, F))}
        if accumulate_bdi:
            acc.update(top_E_B=0.0, top_E_D=0.0, top_E_S=0.0, top_T_obs=0.0,
                       dom_E_B=np.zeros(K), dom_E_D=np.zeros(K), dom_E_S=np.zeros(K),
                       dom_T_obs=np.zeros(K))
        return acc
    dev = hmm.device()
    n_chi = torch.stack([_dev64(r[0], dev) for r in recs])
    t = torch.tensor([float(r[1]) for r in recs], dtype=torch.float64, device=dev)
    ss = exact_suffstats_batch(n_chi, main_ins, main_del, t, dom_ins, dom_del, dom_weights,
                               frag_weights, ext_rates, with_bdi=accumulate_bdi)
    acc = {k: v.sum(0).cpu().numpy() for k, v in ss.items() if k not in ("dom_M_5x5", "bdi")}
    acc["dom_M_5x5"] = [v.sum(0).cpu().numpy() for v in ss["dom_M_5x5"]]
    if accumulate_bdi:
        b = ss["bdi"].sum(0).cpu().numpy()  # (1+K, 4)
        acc.update(top_E_B=float(b[0, 0]), top_E_D=float(b[0, 1]), top_E_S=float(b[0, 2]),
                   top_T_obs=float(b[0, 3]), dom_E_B=b[1:, 0].copy(), dom_E_D=b[1:, 1].copy(),
                   dom_E_S=b[1:, 2].copy(), dom_T_obs=b[1:, 3].copy())
    return acc


def tkf91_stats_from_counts_batch(n_trans_batch, ins_rate, del_rate, t_batch, T_batch=None):
    """(E_B, E_D, E_S) per pair of (B,5,5) TKF91-semantics counts at shared
    (lambda, mu) and per-pair t, T (core/bdi.py:800-888)."""
    L = _abi.lib()
    dev = hmm.device()
    n = _dev64(n_trans_batch, dev)
    t = _dev64(t_batch, dev).reshape(-1)
    T = t if T_batch is None else _dev64(T_batch, dev).reshape(-1)
    B = int(t.shape[0])
    if tuple(n.shape) != (B, 5, 5):
        raise ValueError("n_trans_batch must be (B,5,5)")
    out = torch.empty((B, 3), dtype=torch.float64, device=dev)
    _abi.check(L.tkfmi_bdi_stats_batched(B, float(ins_rate), float(del_rate), _abi._ptr(n),
                                         _abi._ptr(t), _abi._ptr(T), _abi._ptr(out),
                                         _abi.current_stream_ptr(dev)))
    return out[:, 0], out[:, 1], out[:, 2]


bdi_stats_from_counts_batch = tkf91_stats_from_counts_batch
