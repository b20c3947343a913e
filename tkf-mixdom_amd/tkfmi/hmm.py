"""Drop-in 2D pair-HMM dynamic programming (reference: dp/hmm.py).

Same call surfaces as the reference's 2D functions:

* ``forward_backward_2d`` (dp/hmm.py:1834-1973), with table emissions or a
  dense ``log_emit_table`` and ``real_Lx`` / ``real_Ly``
* ``_pad_to_bin`` (dp/hmm.py:66-91), ``safe_log`` (dp/hmm.py:111-122)

plus the batched entry ``fb2d_batch`` that replaces the vmapped, bin-bucketed
callers with one launch over a ragged batch.  All arrays are FP64 torch
tensors on the current HIP device; numpy inputs are accepted and moved.
Everything runs through libtkfmi.so -- there is no CPU fallback.
"""
import os

import numpy as np
import torch

from . import _abi, tables

S_T, M_T, I_T, D_T, E_T = 0, 1, 2, 3, 4
NEG_INF = -1e30

BIN_BASE = 4
BIN_MAX = 4096
BIN_MULTIPLIER = 1.2


def _make_geom_bins(base=BIN_BASE, max_bin=BIN_MAX, multiplier=BIN_MULTIPLIER):
    """Geometric bin sizes (dp/hmm.py:46-59)."""
    bins = [int(base)]
    val = float(base)
    while val < max_bin:
        val *= multiplier
        b = int(round(val))
        if b > bins[-1]:
            bins.append(b)
    if bins[-1] < max_bin:
        bins.append(int(max_bin))
    return bins


_GEOM_BINS = [0, 1, 2, 3] + _make_geom_bins()


def _pad_to_bin(n):
    """Padded length used by the reference for output shapes (dp/hmm.py:66-91).
    The kernels themselves never pad; this only fixes the shape of returned
    posterior lattices so that callers see the reference's shapes."""
    n = max(int(n), 1)
    cap = os.environ.get("TKFMIXDOM_MAX_PAD")
    if cap is not None:
        cap = int(cap)
        if n > cap:
            raise ValueError(f"_pad_to_bin: n={n} exceeds TKFMIXDOM_MAX_PAD={cap}")
        return max(min(((n + 31) // 32) * 32, cap), 1)
    for b in _GEOM_BINS:
        if b >= n:
            return b
    return int(2 ** np.ceil(np.log2(max(n, 1))))


def safe_log(x):
    """log(x) where x > 0, NEG_INF elsewhere (dp/hmm.py:111-122)."""
    x = torch.as_tensor(x, dtype=torch.float64)
    return torch.where(x > 0, torch.log(torch.clamp(x, min=1e-300)), torch.full_like(x, NEG_INF))


def device():
    if not torch.cuda.is_available():
        raise _abi.TkfmiError("tkfmi needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _dev64(x, dev):
    return torch.as_tensor(x, dtype=torch.float64).to(dev)


def _types_list(state_types):
    if isinstance(state_types, torch.Tensor):
        state_types = state_types.detach().cpu().numpy()
    return [int(v) for v in np.asarray(state_types).ravel()]


def fb2d_batch(log_trans, state_types, xs, ys, log_tab, state_group=None, *,
               forward_only=False, want_posteriors=False, want_counts=True,
               want_emit_counts=False, generic=False, kernel=None, checkpoint=False,
               log_emit=None):
    """One launch over a ragged batch of pairs.

    log_trans: (S,S) or (B,S,S) log transitions; log_tab: (G,TAB) or
    (B,G,TAB) log emission tables from ``tkfmi.tables``; xs, ys: sequences
    of integer code arrays (real lengths, no padding).
    Returns dict with 'log_p' (B,), 'n' (B,S,S) (unless forward_only or
    want_counts=False), 'posteriors' (list of (Lx+1,Ly+1,S) views) and
    'emit_counts' (B,G,TAB): expected uses of each emission-table entry,
    i.e. d log_p / d log_tab.
    ``generic=True`` (or kernel="generic") forces the any-layout kernel; the
    TKF and MixDom K*F<=3 layouts otherwise run on the cell-per-lane path,
    chained workgroups per pair by default or one workgroup per pair with
    kernel="class".  ``checkpoint=True`` keeps no full lattice and recomputes
    it in the backward (the 17..162-state kernels: per band from the forward's
    boundary rows, or per K-antidiagonal segment with kernel="class" and for
    batches of few bands; same results, less lattice memory).
    ``log_emit``: a list of dense (Lx+1, Ly+1, S) log-emission tables, one
    per pair (the reference's log_emit_table), used instead of ``log_tab``
    (then None); xs / ys give only the lengths.  Runs on the generic kernel
    (<= 15 emitting states) or the 16..160-state kernel.
    """
    dev = device()
    lt = _dev64(log_trans, dev)
    st = _types_list(state_types)
    S = len(st)
    if state_group is None:
        state_group = [0] * S
    pairs = _abi.PackedPairs(list(xs), list(ys), dev)
    dense = None
    if log_emit is not None:
        if len(log_emit) != pairs.B:
            raise ValueError("log_emit needs one table per pair")
        blocks = []
        for b, e in enumerate(log_emit):
            e = _dev64(e, dev)
            shp = (int(pairs.x_len[b]) + 1, int(pairs.y_len[b]) + 1, S)
            if tuple(e.shape) != shp:
                raise ValueError(f"log_emit[{b}] must be {shp}, got {tuple(e.shape)}")
            blocks.append(e.reshape(-1))
        dense = torch.cat(blocks) if blocks else torch.zeros(0, dtype=torch.float64, device=dev)
    tab = _dev64(log_tab, dev) if log_tab is not None else None
    out = _abi.fb2d_batched(lt, st, state_group, tab, pairs, forward_only=forward_only,
                            want_posteriors=want_posteriors, want_counts=want_counts,
                            want_emit_counts=want_emit_counts, generic=generic,
                            kernel=kernel, checkpoint=checkpoint, log_emit=dense)
    if "posteriors" in out:
        out["posteriors"] = _abi.split_posteriors(out["posteriors"], pairs, S)
    return out


def _real_len(seq, real):
    n = int(seq.shape[0]) if hasattr(seq, "shape") else len(seq)
    return n if real is None else int(real)


def forward_backward_2d(log_trans, state_types, x_seq, y_seq, sub_matrix, pi,
                        log_emit_table=None, real_Lx=None, real_Ly=None,
                        forward_only=False):
    """Reference surface of dp/hmm.py:1834-1973.

    Returns log_prob, or (log_prob, posteriors (Lx_pad+1, Ly_pad+1, S),
    expected_trans (S,S)) with (Lx_pad, Ly_pad) the _pad_to_bin sizes of
    the array shapes; posteriors are zero outside the real region.
    Emissions are the shared model (pair_hmm_emissions) of sub_matrix / pi,
    or the caller's dense ``log_emit_table`` ((Lx_arr+1, Ly_arr+1, S)),
    whose region beyond the real lengths is NEG_INF as in the reference
    (dp/hmm.py:1878-1889).  ``real_Lx`` / ``real_Ly`` are the real lengths
    of sequences already padded to their arrays' shapes.  The DP runs at
    the real lengths only: cells beyond them are unreachable.
    """
    Lx_arr = int(np.asarray(x_seq.shape[0] if hasattr(x_seq, "shape") else len(x_seq)))
    Ly_arr = int(np.asarray(y_seq.shape[0] if hasattr(y_seq, "shape") else len(y_seq)))
    Lx = _real_len(x_seq, real_Lx)
    Ly = _real_len(y_seq, real_Ly)
    if not (0 <= Lx <= Lx_arr and 0 <= Ly <= Ly_arr):
        raise ValueError("real lengths must lie within the sequence arrays")
    x = np.asarray(x_seq.detach().cpu() if isinstance(x_seq, torch.Tensor) else x_seq)[:Lx]
    y = np.asarray(y_seq.detach().cpu() if isinstance(y_seq, torch.Tensor) else y_seq)[:Ly]
    dev = device()
    S = len(_types_list(state_types))
    if log_emit_table is not None:
        le = _dev64(log_emit_table, dev)
        if tuple(le.shape) != (Lx_arr + 1, Ly_arr + 1, S):
            raise ValueError(f"log_emit_table must be {(Lx_arr + 1, Ly_arr + 1, S)}, "
                             f"got {tuple(le.shape)}")
        res = fb2d_batch(log_trans, state_types, [x], [y], None, forward_only=forward_only,
                         want_posteriors=not forward_only,
                         log_emit=[le[:Lx + 1, :Ly + 1].contiguous()])
    else:
        tab = tables.shared_tables(_dev64(sub_matrix, dev), _dev64(pi, dev))
        res = fb2d_batch(log_trans, state_types, [x], [y], tab, forward_only=forward_only,
                         want_posteriors=not forward_only)
    _abi.check_status(res)
    logp = res["log_p"][0]
    if forward_only:
        return logp
    Lxp, Lyp = _pad_to_bin(Lx_arr), _pad_to_bin(Ly_arr)
    post = torch.zeros((Lxp + 1, Lyp + 1, S), dtype=torch.float64, device=dev)
    post[:Lx + 1, :Ly + 1] = res["posteriors"][0]
    return logp, post, res["n"][0]
