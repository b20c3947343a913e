"""1D alignment-constrained forward/backward: the reference's Python surface
over libtkfmi's ``tkfmi_fb1d_batched``.

* :func:`forward_backward_1d_padded` -- dp/hmm.py:471-607, same arguments and
  return values (log P; or log P, posteriors (padded_L, S) with zero rows
  past ``seq_length``, expected transitions (S, S)).  ``scan_mode`` is
  accepted and ignored: both reference modes compute the same quantities.
* :func:`fb1d_batch` -- the vmapped caller (train_pfam.py:1563-1610) for a
  ragged batch of dense log-emission tables, without _pad_to_bin padding.
* :func:`fb1d_constrained_batch` -- the fused path: emissions looked up on
  the device from column types, characters and the per-group tables of
  tkfmi.tables, as pair_hmm_emissions_constrained builds them.
* :func:`pair_hmm_emissions_constrained` -- dp/hmm.py:945-978 (per-domain
  :854-942, per-class :981-1100) as a dense (L, S) table.
"""
import numpy as np
import torch

from . import _abi, hmm, tables

S_STATE, E_STATE = 0, 4  # dp/hmm.py defaults: init_state=S, final_state=E (core/params.py:22)
M, I, D = 1, 2, 3
NEG_INF = -1e30


def _np(x):
    return np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x)


def fb1d_batch(log_trans, log_emits, lengths=None, *, init_state=S_STATE, final_state=E_STATE,
               forward_only=False, want_posteriors=True, want_counts=True):
    """Batched dense-emission 1D forward/backward.

    log_trans: (S,S) or (B,S,S); log_emits: list of (L_b, S) tables, or a
    (B, L_pad, S) tensor with ``lengths`` (rows past a pair's length are
    ignored).  Returns a dict with 'log_p' (B,), and unless forward_only
    'n' (B,S,S) and 'posteriors' (list of (L_b, S))."""
    dev = hmm.device()
    lt = hmm._dev64(log_trans, dev)
    S = int(lt.shape[-1])
    if isinstance(log_emits, (list, tuple)):
        tabs = [hmm._dev64(e, dev).reshape(-1, S) for e in log_emits]
        lens = np.array([t.shape[0] for t in tabs], dtype=np.int32)
    else:
        le = hmm._dev64(log_emits, dev)
        lens = (np.full(le.shape[0], le.shape[1], dtype=np.int32) if lengths is None
                else np.asarray(_np(lengths), dtype=np.int32).reshape(-1))
        if lens.size and (lens.min() < 0 or lens.max() > le.shape[1]):
            raise ValueError("lengths must lie in [0, L_pad]")
        tabs = [le[b, :int(lens[b])] for b in range(le.shape[0])]
    flat = (torch.cat(tabs, 0) if tabs and int(lens.sum()) > 0
            else torch.zeros((0, S), dtype=torch.float64, device=dev))
    if flat.shape[0] == 0:
        flat = torch.zeros((1, S), dtype=torch.float64, device=dev)
    r = _abi.Runner1D(lt, init_state=init_state, final_state=final_state, lengths=lens,
                      log_emit=flat, forward_only=forward_only,
                      want_posteriors=want_posteriors and not forward_only,
                      want_counts=want_counts)
    r.run(zero_outputs=False)
    out = {"log_p": r.log_p}
    if r.n is not None:
        out["n"] = r.n
    if r.post is not None:
        out["posteriors"] = r.posteriors()
    return out


def fb1d_constrained_batch(log_trans, state_types, state_group, log_tab, col_types, ancs, descs, *,
                           init_state=0, final_state=1, forward_only=False, want_posteriors=False,
                           want_counts=True, want_emit_counts=True, runner=False):
    """Batched 1D forward/backward with emissions looked up on the device.

    col_types / ancs / descs: per pair, one entry per alignment column (type
    1=M, 2=I, 3=D; the ancestor code at M/D columns, the descendant code at
    M/I columns, anything elsewhere).  log_tab (G, TAB_SZ) or (B, G, TAB_SZ).
    Returns a dict with 'log_p', 'n', 'emit_counts' (B, G, TAB_SZ) and
    'posteriors' as requested (or the Runner1D itself when ``runner``)."""
    dev = hmm.device()
    lt = hmm._dev64(log_trans, dev)
    cols = _abi.PackedColumns(col_types, ancs, descs, dev)
    r = _abi.Runner1D(lt, init_state=init_state, final_state=final_state, columns=cols,
                      state_types=state_types, state_group=state_group,
                      log_tab=hmm._dev64(log_tab, dev), forward_only=forward_only,
                      want_posteriors=want_posteriors, want_counts=want_counts,
                      want_emit_counts=want_emit_counts)
    if runner:
        return r
    r.run(zero_outputs=False)
    out = {"log_p": r.log_p}
    if r.n is not None:
        out["n"] = r.n
    if r.ecnt is not None:
        out["emit_counts"] = r.ecnt
    if r.post is not None:
        out["posteriors"] = r.posteriors()
    return out


def forward_backward_1d_padded(log_trans, log_emissions, seq_length, init_state=S_STATE,
                               final_state=E_STATE, forward_only=False, scan_mode="associative"):
    """dp/hmm.py:471-607 for one (padded) sequence of log emissions."""
    if scan_mode not in ("associative", "sequential"):
        raise ValueError(f"unknown scan_mode {scan_mode!r}")
    le = hmm._dev64(log_emissions, hmm.device())
    Lp, S = int(le.shape[0]), int(le.shape[1])
    L = int(_np(seq_length))
    if L < 0 or L > Lp:
        raise ValueError("seq_length must lie in [0, padded_L]")
    res = fb1d_batch(log_trans, le[None], [L], init_state=init_state, final_state=final_state,
                     forward_only=forward_only)
    lp = res["log_p"][0]
    if forward_only:
        return lp
    post = torch.zeros((Lp, S), dtype=torch.float64, device=le.device)
    post[:L] = res["posteriors"][0]
    return lp, post, res["n"][0]


def _state_dom_frag(ns, n_frag):
    body = np.arange(ns - 2)
    dom = np.zeros(ns, dtype=np.int64)
    frag = np.zeros(ns, dtype=np.int64)
    dom[2:] = body // (5 * n_frag)
    frag[2:] = (body % (5 * n_frag)) % n_frag
    return dom, frag


def pair_hmm_emissions_constrained(state_types, state_seq, anc_chars, desc_chars, sub_matrices, pis,
                                   n_dom, n_frag, classdist=None, class_sub_matrices=None,
                                   class_pis=None):
    """(L, S) log emissions with NEG_INF at type-mismatched (column, state)
    entries.  anc_chars / desc_chars are filled-length (one per column).
    Per-class (MixDom2) when classdist and the per-class arguments are all
    given, else per-domain (dp/hmm.py:972-978)."""
    dev = hmm.device()
    st = np.asarray(_np(state_types), dtype=np.int64)
    ns = len(st)
    dom, frag = _state_dom_frag(ns, n_frag)
    if classdist is not None and class_sub_matrices is not None and class_pis is not None:
        tab = tables.per_class_tables(class_sub_matrices, class_pis, classdist, device=dev)
        grp = dom * n_frag + frag
    else:
        tab = tables.per_domain_tables(sub_matrices, pis, device=dev)
        grp = dom
    ct = np.asarray(_np(state_seq), dtype=np.int64)
    a = np.clip(np.asarray(_np(anc_chars), dtype=np.int64), 0, tables.A_PAD - 1)
    b = np.clip(np.asarray(_np(desc_chars), dtype=np.int64), 0, tables.A_PAD - 1)
    idx = np.where(ct == M, a * tables.A_PAD + b,
                   np.where(ct == I, tables.A_PAD ** 2 + b, tables.A_PAD ** 2 + tables.A_PAD + a))
    grp_t = torch.as_tensor(grp, device=dev)
    vals = tab[grp_t[None, :], torch.as_tensor(idx, device=dev)[:, None]]   # (L, S)
    ok = torch.as_tensor(st[None, :] == ct[:, None], device=dev)
    return torch.where(ok, vals, torch.full_like(vals, NEG_INF))
