"""E-step callers of the 2D forward/backward path (reference surfaces):

* ``estep_batch_tkf92`` / ``estep_batch_tkf92_forward_only``
  (train/tkf92_svi_bw.py:263-360)
* ``e_step`` / ``e_step_mixdom2`` (train/likelihood.py:49-369)
* ``process_pairs_batched_2d`` -- the ``delta`` dict of
  ``_process_pairs_batched_2d`` (python/train_pfam.py:2672-2903)
* ``process_pairs_batched`` -- the ``delta`` dict of the alignment-constrained
  (1D) ``_process_pairs_batched`` (python/train_pfam.py:2049-2603)

Each runs the batched forward/backward kernel once over its pairs (no
geometric-bin bucketing: every pair at its real lengths), then the
statistics kernels (TKF92 split + BDI, null-state restoration, Holmes-Rubin)
on the device.  Per-class / per-domain emission statistics come from the
kernel's expected emission counts: the class responsibility of a cell
depends only on its (group, characters), so summing posterior x
responsibility over cells equals responsibility x expected count per table
entry.
"""
import ctypes

import numpy as np
import torch

from . import _abi, build as dbuild, hmm, hr, mixdom, params, suffstats, tables, vjp

AA = 20
_NM = tables.A_PAD * tables.A_PAD


def _np(x):
    return np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x)


def _dev64(x, dev):
    return torch.as_tensor(x, dtype=torch.float64).to(dev)


def _unpad(seqs_pad, lens):
    seqs_pad = _np(seqs_pad)
    return [np.asarray(seqs_pad[b][:int(n)]) for b, n in enumerate(_np(lens))]


def _floor_log(x):
    return torch.log(torch.clamp(x, min=1e-30))


def _split_ecnt(ec, A=AA):
    """(..., TAB) emission counts -> match (..., A, A), insert (..., A),
    delete (..., A) over the first A codes (the reference's one_hot(., A)
    drops the wildcard)."""
    m = ec[..., :_NM].reshape(ec.shape[:-1] + (tables.A_PAD, tables.A_PAD))[..., :A, :A]
    i = ec[..., _NM:_NM + tables.A_PAD][..., :A]
    d = ec[..., _NM + tables.A_PAD:][..., :A]
    return m, i, d


# ------------------------------------------------------------------- TKF92
def _tkf92_fb(xs, ys, ts, lam, mu, ext, Q, pi, forward_only):
    dev = hmm.device()
    b = dbuild.DeviceBuilder("tkf92", "shared", _np(Q), _np(pi), lam=float(lam), mu=float(mu),
                             ext=float(ext), log_floor=True, device=dev)
    lt, tab, ts = b.build(_np(ts))
    pairs = _abi.PackedPairs(xs, ys, dev)
    r = _abi.BatchRunner(lt, params.TKF_STATE_TYPES, [0] * 5, tab, pairs,
                         forward_only=forward_only)
    r.run(zero_outputs=False)
    r.check_status()
    return r, ts


def estep_batch_tkf92(xs_pad, ys_pad, ts, Lx_real, Ly_real, lam, mu, ext, Q, pi):
    """Batched TKF92 E-step (train/tkf92_svi_bw.py:263-334): per pair log_p,
    BDI statistics B, D, S (clamped at 0), L, M, T = t, ext_count,
    notext_count and the chi counts n_chi_diag (B,5,5).  The transition
    matrix is log(max(tau92, 1e-30)) (:149-150); t stays FP64 (:299)."""
    xs, ys = _unpad(xs_pad, Lx_real), _unpad(ys_pad, Ly_real)
    r, t = _tkf92_fb(xs, ys, ts, lam, mu, ext, Q, pi, False)
    B = len(xs)
    out = torch.empty((B, _abi.TKF92_STATS), dtype=torch.float64, device=t.device)
    _abi.check(_abi.lib().tkfmi_tkf92_estep_stats(B, float(lam), float(mu), float(ext),
                                                   _abi._ptr(t), _abi._ptr(r.n), _abi._ptr(out),
                                                   _abi.current_stream_ptr(t.device)))
    return {"log_p": r.log_p, "B": out[:, 29], "D": out[:, 30], "S": out[:, 31],
            "L": out[:, 27], "M": out[:, 28], "T": out[:, 32], "ext_count": out[:, 25],
            "notext_count": out[:, 26], "n_chi_diag": r.n, "n_trans_resolved": out[:, :25].reshape(B, 5, 5)}


def estep_batch_tkf92_forward_only(xs_pad, ys_pad, ts, Lx_real, Ly_real, lam, mu, ext, Q, pi):
    """Per-pair log_p only (train/tkf92_svi_bw.py:337-360)."""
    xs, ys = _unpad(xs_pad, Lx_real), _unpad(ys_pad, Ly_real)
    r, _ = _tkf92_fb(xs, ys, ts, lam, mu, ext, Q, pi, True)
    return r.log_p


# ------------------------------------------------------------------ e_step
def e_step(model, p, x, y):
    """E-step for one pair (train/likelihood.py:49-137).

    Returns (log_prob, n_trans (N,N), ss) with ss = {'match_pairs': list of
    (anc, desc, weight) for match posteriors > 1e-8, 'pi_obs_counts':
    {'match_anc_count', 'insert_count', 'delete_count'} (A,)}."""
    if model == "mixdom2":
        return e_step_mixdom2(p, x, y)
    dev = hmm.device()
    t = _dev64(p["t"], dev)
    if model in ("tkf91", "tkf91_cond"):
        tau = params.tkf91_trans(p["ins_rate"], p["del_rate"], t)
        st = list(params.TKF_STATE_TYPES)
    elif model == "tkf92":
        tau = params.tkf92_trans(p["ins_rate"], p["del_rate"], t, p["ext"])
        st = list(params.TKF_STATE_TYPES)
    elif model == "mixdom":
        tau = mixdom.build_nested_trans(p["main_ins"], p["main_del"], t, p["dom_ins"], p["dom_del"],
                                        p["dom_weights"], p["frag_weights"], p["ext_rates"])
        K, F = np.asarray(p["frag_weights"]).shape
        st = mixdom.state_types(K, F)
    else:
        raise ValueError(f"Unknown model: {model}")
    A = int(np.asarray(p["pi"]).shape[0])
    sub = vjp.transition_matrix(_dev64(p["Q"], dev), t)
    tab = tables.shared_tables(sub, _dev64(p["pi"], dev))
    res = hmm.fb2d_batch(hmm.safe_log(tau), st, [x], [y], tab, want_posteriors=True,
                         want_emit_counts=True)
    _abi.check_status(res)
    post = res["posteriors"][0]
    stt = torch.as_tensor(st, device=dev)
    m_post = (post[1:, 1:, :] * (stt == params.M)).sum(-1).cpu().numpy()   # (Lx, Ly)
    ii, jj = np.where(m_post > 1e-8)
    xa, ya = np.asarray(x, dtype=np.int64), np.asarray(y, dtype=np.int64)
    match_pairs = list(zip(xa[ii].tolist(), ya[jj].tolist(), m_post[ii, jj].astype(np.float64).tolist()))
    mt, it, dt = _split_ecnt(res["emit_counts"][0].sum(0), A)
    pi_obs = {"match_anc_count": mt.sum(1).cpu().numpy(), "insert_count": it.cpu().numpy(),
              "delete_count": dt.cpu().numpy()}
    return float(res["log_p"][0]), res["n"][0], {"match_pairs": match_pairs, "pi_obs_counts": pi_obs}


def _class_counts(ecnt, class_subs, class_pis, cd, K, F):
    """Per-class statistics from per-(d,f) emission counts (B, KF, TAB):
    responsibilities q[c | d, f, type, a, b] = cd[d,f,c] f_c / sum_c' (...)
    with f_c = pi_c[a] P_c[a,b] (M), pi_c[b] (I), pi_c[a] (D)
    (train/likelihood.py:296-319).  class_subs (B, C, A, A), class_pis
    (C, A), cd (K, F, C).  Returns the (B, ...) class counts."""
    C, A = class_pis.shape
    Ap = tables.A_PAD
    ec = ecnt.reshape(ecnt.shape[0], K, F, -1)
    em = ec[..., :_NM].reshape(ec.shape[:3] + (Ap, Ap))
    ei = ec[..., _NM:_NM + Ap]
    ed = ec[..., _NM + Ap:]
    codes = torch.clamp(torch.arange(Ap, device=ecnt.device), max=A - 1)   # gather clamps
    pic = class_pis[:, codes]                                          # (C, 21)
    fM = pic[None, :, :, None] * class_subs[:, :, codes][:, :, :, codes]  # (B, C, 21, 21)
    jM = cd[None, :, :, :, None, None] * fM[:, None, None]             # (B, K, F, C, 21, 21)
    qM = jM / (jM.sum(3, keepdim=True) + 1e-300)
    jS = cd[:, :, :, None] * pic[None, None]                           # (K, F, C, 21)
    qS = jS / (jS.sum(2, keepdim=True) + 1e-300)
    wM = qM * em[:, :, :, None]                                        # (B, K, F, C, 21, 21)
    wI = qS[None] * ei[:, :, :, None]                                  # (B, K, F, C, 21)
    wD = qS[None] * ed[:, :, :, None]
    valid = torch.arange(Ap, device=ecnt.device) < A
    cm = wM[..., valid, :][..., valid].sum((1, 2))                     # (B, C, A, A)
    ci = wI[..., valid].sum((1, 2))
    cdl = wD[..., valid].sum((1, 2))
    cdc = wM.sum((-1, -2)) + wI.sum(-1) + wD.sum(-1)                   # (B, K, F, C)
    return cm, ci, cdl, cdc


def e_step_mixdom2(p, x, y):
    """E-step for MixDom2 (train/likelihood.py:143-369): (log_prob, n_trans,
    {class_W, class_U, class_match_counts, class_insert_counts,
    class_delete_counts, classdist_counts})."""
    dev = hmm.device()
    t = _dev64(p["t"], dev).reshape(1)
    K, F = np.asarray(p["frag_weights"]).shape
    chi = mixdom.build_nested_trans(p["main_ins"], p["main_del"], t, p["dom_ins"], p["dom_del"],
                                    p["dom_weights"], p["frag_weights"], p["ext_rates"])
    st = mixdom.state_types(K, F)
    S_exch, cpis = _dev64(p["class_S_exch"], dev), _dev64(p["class_pis"], dev)
    cd = _dev64(p["classdist"], dev)
    subs = vjp.build_class_subs(S_exch, cpis, t)                      # (1, C, A, A)
    tab = tables.per_class_tables(subs, cpis, cd)                     # (1, KF, TAB)
    groups = tables.state_groups_per_class(len(st), F)
    res = hmm.fb2d_batch(hmm.safe_log(chi), st, [x], [y], tab, groups, want_emit_counts=True)
    _abi.check_status(res)
    cm, ci, cdl, cdc = _class_counts(res["emit_counts"], subs, cpis, cd, K, F)
    Qc = vjp.build_Q_from_S_pi(S_exch, cpis)
    W, U = hr.per_pair_hr_stats(Qc, cpis, t, cm)
    ss = {"class_W": W[0].cpu().numpy(), "class_U": U[0].cpu().numpy(),
          "class_match_counts": cm[0].cpu().numpy(), "class_insert_counts": ci[0].cpu().numpy(),
          "class_delete_counts": cdl[0].cpu().numpy(), "classdist_counts": cdc[0].cpu().numpy()}
    return float(res["log_p"][0]), res["n"][0], ss


# -------------------------------------------------- batched 2D E-step delta
def process_pairs_batched_2d(batch_pairs, chi_params, st, Q_lg, pi_lg, N, dom_Qs=None,
                             dom_pis=None, n_dom=1, n_frag=1, params_=None):
    """Sufficient-statistics delta of a minibatch (train_pfam.py:2672-2903).

    batch_pairs: (x, y, states, anc_chars, desc_chars, t_est) tuples (only
    x, y, t_est used), or a tkfmi.pairstore.PairStore, whose device-decoded
    residue codes are used as they are (no host packing).  Per-pair chi at t_est (log floor 1e-30,
    train_pfam.py:1632-1640), per-domain substitution matrices expm(Q_d t)
    with per-domain emissions (LG Q/pi tiled when dom_Qs is None), t_est
    rounded to float32 first as the reference does (:2828).  Pairs with a
    non-finite log-likelihood are skipped (:2859-2860).  Returns the delta
    dict or None for an empty batch."""
    if not len(batch_pairs):
        return None
    dev = hmm.device()
    from .pairstore import PairStore
    if isinstance(batch_pairs, PairStore):
        store = batch_pairs
        t32 = store.t.astype(np.float32).astype(np.float64)
    else:
        store = None
        xs = [np.asarray(it[0], dtype=np.int64) for it in batch_pairs]
        ys = [np.asarray(it[1], dtype=np.int64) for it in batch_pairs]
        t32 = np.array([it[5] for it in batch_pairs], dtype=np.float32).astype(np.float64)
    cp = chi_params
    if dom_Qs is not None and dom_pis is not None:
        Qs_n, pis_n = _np(dom_Qs).astype(np.float64), _np(dom_pis).astype(np.float64)
    else:
        Qs_n = np.tile(_np(Q_lg).astype(np.float64)[None], (n_dom, 1, 1))
        pis_n = np.tile(_np(pi_lg).astype(np.float64)[None], (n_dom, 1))
    mix = dict(main_ins=cp["main_ins"], main_del=cp["main_del"], dom_ins=cp["dom_ins"],
               dom_del=cp["dom_del"], dom_w=cp["dom_weights"], frag_w=cp["frag_weights"],
               ext=cp["ext_rates"])
    b = dbuild.DeviceBuilder("mixdom", "per_domain", Qs_n, pis_n, mix=mix, log_floor=True, device=dev)
    lt, tab, t = b.build(t32)                                         # chi and (B, K, TAB) tables
    Qs, pis = _dev64(Qs_n, dev), _dev64(pis_n, dev)
    pairs = store if store is not None else _abi.PackedPairs(xs, ys, dev)
    r = _abi.BatchRunner(lt, b.state_types, b.state_group, tab, pairs, want_emit_counts=True)
    r.run(zero_outputs=False)
    r.check_status()  # a timed-out wait is an error, not an impossible pair
    ok = torch.isfinite(r.log_p)
    dm, di, dd = _split_ecnt(r.ecnt)                                  # (B, K, 20, 20) ...
    okf = ok.to(torch.float64)
    delta = {
        "agg_n_chi": (r.n * okf[:, None, None]).sum(0).cpu().numpy(),
        "total_ll": float(torch.where(ok, r.log_p, torch.zeros_like(r.log_p)).sum()),
        "n_pairs": int(ok.sum()),
        "n_files": 1,
        "match_counts": (dm * okf[:, None, None, None]).sum((0, 1)).cpu().numpy(),
        "dom_match_counts": (dm * okf[:, None, None, None]).sum(0).cpu().numpy(),
        "dom_insert_counts": (di * okf[:, None, None]).sum(0).cpu().numpy(),
        "dom_delete_counts": (dd * okf[:, None, None]).sum(0).cpu().numpy(),
    }
    if delta["n_pairs"] == 0:
        return None
    sel = torch.nonzero(ok).flatten()
    W, U = hr.per_pair_hr_stats(Qs, pis, t[sel], dm[sel].contiguous())
    delta["dom_W"] = W.sum(0).cpu().numpy()
    delta["dom_U"] = U.sum(0).cpu().numpy()
    ss = suffstats.exact_suffstats_batch(r.n[sel], cp["main_ins"], cp["main_del"], t[sel],
                                         cp["dom_ins"], cp["dom_del"], cp["dom_weights"],
                                         cp["frag_weights"], cp["ext_rates"], with_bdi=True)
    acc = {k: v.sum(0).cpu().numpy() for k, v in ss.items() if k not in ("dom_M_5x5", "bdi")}
    acc["dom_M_5x5"] = [v.sum(0).cpu().numpy() for v in ss["dom_M_5x5"]]
    b = ss["bdi"].sum(0).cpu().numpy()
    acc.update(top_E_B=float(b[0, 0]), top_E_D=float(b[0, 1]), top_E_S=float(b[0, 2]),
               top_T_obs=float(b[0, 3]), dom_E_B=b[1:, 0].copy(), dom_E_D=b[1:, 1].copy(),
               dom_E_S=b[1:, 2].copy(), dom_T_obs=b[1:, 3].copy())
    delta["exact_ss"] = acc
    return delta


# ---------------------------------------- batched 1D (MSA-guided) E-step delta
def process_pairs_batched(batch_pairs, chi_params, st, Q_lg, pi_lg, N, alignment_to_states=None,
                          dom_Qs=None, dom_pis=None, n_dom=1, n_frag=1, classdist_params=None,
                          fb_scan_mode="associative"):
    """Sufficient-statistics delta of a minibatch of aligned pairs
    (train_pfam.py:2049-2603), on the 1D alignment-constrained DP.

    batch_pairs: (x, y, states, anc_chars, desc_chars, t_est) tuples; the
    column types ``states`` and the compact character lists define each
    pair's alignment (x, y are unused, as in the reference).  As the
    reference does: chi is built at t_est rounded to float32 (:2322) and the
    substitution matrices at the FP64 t_est (:2112); Holmes-Rubin and the
    per-pair count restoration use the float32 t (:2364, :2343); emissions
    are per-domain expm(Q_d t) (LG tiled when dom_Qs is None), or per-class
    mixtures when ``classdist_params`` holds 'classdist', 'class_S_exch',
    'class_pis' (:2137-2157).  Every pair is accumulated (no non-finite
    filter on this path).  The reference rounds the log emissions to float32
    before its DP (:2243); here they stay FP64.  Its per-column 1e-30
    cut-offs in the class responsibilities (:2455-2462) drop contributions
    below 1e-30 that are kept here.  ``alignment_to_states`` and
    ``fb_scan_mode`` are accepted for signature parity (unused there too /
    same result in both modes).  Returns the delta dict or None."""
    if not batch_pairs:
        return None
    dev = hmm.device()
    t64 = np.array([it[5] for it in batch_pairs], dtype=np.float64)
    t32 = t64.astype(np.float32).astype(np.float64)
    cp = chi_params
    K, F = int(n_dom), int(n_frag)
    mix = dict(main_ins=cp["main_ins"], main_del=cp["main_del"], dom_ins=cp["dom_ins"],
               dom_del=cp["dom_del"], dom_w=cp["dom_weights"], frag_w=cp["frag_weights"],
               ext=cp["ext_rates"])
    use_cd = classdist_params is not None and "classdist" in classdist_params
    if dom_Qs is not None and dom_pis is not None:
        Qs_n, pis_n = _np(dom_Qs).astype(np.float64), _np(dom_pis).astype(np.float64)
    else:
        Qs_n = np.tile(_np(Q_lg).astype(np.float64)[None], (K, 1, 1))
        pis_n = np.tile(_np(pi_lg).astype(np.float64)[None], (K, 1))
    if use_cd:
        S_exch = _np(classdist_params["class_S_exch"]).astype(np.float64)
        cpis_n = _np(classdist_params["class_pis"]).astype(np.float64)
        cd_n = _np(classdist_params["classdist"]).astype(np.float64)
        Qc_n = S_exch * cpis_n[:, None, :]
        for c in range(Qc_n.shape[0]):
            np.fill_diagonal(Qc_n[c], 0.0)
            Qc_n[c][np.diag_indices(Qc_n.shape[1])] = -Qc_n[c].sum(axis=1)
        b = dbuild.DeviceBuilder("mixdom", "per_class", Qc_n, cpis_n, mix=mix, class_dist=cd_n,
                                 log_floor=True, device=dev)
    else:
        b = dbuild.DeviceBuilder("mixdom", "per_domain", Qs_n, pis_n, mix=mix, log_floor=True,
                                 device=dev)
    if b.S != N:
        raise ValueError(f"N = {N} but the model has {b.S} states")
    lt, _, t_dev32 = b.build(t32)                    # chi at the float32 t
    _, tab, _ = b.build(t64)                         # emission tables at the FP64 t
    cts, ancs, descs = [], [], []
    for it in batch_pairs:
        ct = np.asarray(it[2], dtype=np.int64).ravel()
        a = np.zeros(len(ct), dtype=np.int64)
        d = np.zeros(len(ct), dtype=np.int64)
        a[(ct == 1) | (ct == 3)] = np.asarray(it[3], dtype=np.int64).ravel()
        d[(ct == 1) | (ct == 2)] = np.asarray(it[4], dtype=np.int64).ravel()
        cts.append(ct)
        ancs.append(a)
        descs.append(d)
    from . import hmm1d
    res = hmm1d.fb1d_constrained_batch(lt, b.state_types, b.state_group, tab, cts, ancs, descs,
                                       init_state=0, final_state=1, want_emit_counts=True)
    n, lp, ec = res["n"], res["log_p"], res["emit_counts"]
    Bn = len(batch_pairs)
    ec_dom = ec.reshape(Bn, K, -1, tables.TAB_SZ).sum(2) if use_cd else ec   # (B, K, TAB)
    dm, di, dd = _split_ecnt(ec_dom)
    delta = {
        "agg_n_chi": n.sum(0).cpu().numpy(),
        "total_ll": float(lp.sum()),
        "n_pairs": Bn,
        "n_files": 1,
        "match_counts": dm.sum((0, 1)).cpu().numpy(),
        "dom_match_counts": dm.sum(0).cpu().numpy(),
        "dom_insert_counts": di.sum(0).cpu().numpy(),
        "dom_delete_counts": dd.sum(0).cpu().numpy(),
    }
    Qs, pis = _dev64(Qs_n, dev), _dev64(pis_n, dev)
    W, U = hr.per_pair_hr_stats(Qs, pis, t_dev32, dm.contiguous())
    delta["dom_W"] = W.sum(0).cpu().numpy()
    delta["dom_U"] = U.sum(0).cpu().numpy()
    if use_cd:
        cpis = _dev64(cpis_n, dev)
        cdv = _dev64(cd_n, dev)
        subs = vjp.build_class_subs(_dev64(S_exch, dev), cpis, _dev64(t64, dev))  # (B, C, A, A)
        cm, ci, cdl, cdc = _class_counts(ec, subs, cpis, cdv, K, F)
        Wc, Uc = hr.per_pair_hr_stats(_dev64(Qc_n, dev), cpis, t_dev32, cm.contiguous())
        delta.update(class_match_counts=cm.sum(0).cpu().numpy(),
                     class_insert_counts=ci.sum(0).cpu().numpy(),
                     class_delete_counts=cdl.sum(0).cpu().numpy(),
                     classdist_counts=cdc.sum(0).cpu().numpy(),
                     class_W=Wc.sum(0).cpu().numpy(), class_U=Uc.sum(0).cpu().numpy())
    ss = suffstats.exact_suffstats_batch(n, cp["main_ins"], cp["main_del"], t_dev32,
                                         cp["dom_ins"], cp["dom_del"], cp["dom_weights"],
                                         cp["frag_weights"], cp["ext_rates"], with_bdi=True)
    acc = {k: v.sum(0).cpu().numpy() for k, v in ss.items() if k not in ("dom_M_5x5", "bdi")}
    acc["dom_M_5x5"] = [v.sum(0).cpu().numpy() for v in ss["dom_M_5x5"]]
    bb = ss["bdi"].sum(0).cpu().numpy()
    acc.update(top_E_B=float(bb[0, 0]), top_E_D=float(bb[0, 1]), top_E_S=float(bb[0, 2]),
               top_T_obs=float(bb[0, 3]), dom_E_B=bb[1:, 0].copy(), dom_E_D=bb[1:, 1].copy(),
               dom_E_S=bb[1:, 2].copy(), dom_T_obs=bb[1:, 3].copy())
    delta["exact_ss"] = acc
    return delta
