"""Null-state count restoration and sufficient-statistic extraction (oracle,
numpy FP64).

Restates python/tkfmixdom/jax/models/exact_suffstats_batch.py:
  * the fully exploded MixDom machine (exact_suffstats_batch.py:61-253):
    Start, End, {Mat,Ins,Del}Dom, per domain {Mat,Ins,Del}DomType and
    {Mat,Ins,Del,I,D}Frag, per (domain, fragment) the five FragType, Emit and
    FragEnd states, then {Mat,Ins,Del}DomEnd -- N = 8 + 8K + 15KF states;
  * elimination of the null groups in chain order FragType, FragEnd, Frag,
    DomType, Dom, DomEnd (:468-482), each by T_red = T_KK + T_KZ C T_ZK with
    C = (I - T_ZZ)^-1 (:272-282);
  * ghost-usage count restoration (:285-339) back through the six steps;
  * extraction of top_5x5, dom_M_5x5, dom_kappa/1mkappa, ext, term, dom_w,
    frag_w (:347-458).
The kept states after the six eliminations are Start, End and the Emit
states in the collapsed chi order, so the reduced matrix is chi.
"""
import numpy as np

from .tkf import S, M, I, D, E, tkf91_trans

_FRAG_KINDS = ("Mat", "Ins", "Del", "I", "D")
_EMIT_KINDS = ("MatEmit", "InsEmit", "DelEmit", "IEmit", "DEmit")


def state_names(K, F):
    names = ["Start", "End", "MatDom", "InsDom", "DelDom"]
    names += [f"{p}DomType[{k}]" for k in range(K) for p in ("Mat", "Ins", "Del")]
    names += [f"{p}Frag[{k}]" for k in range(K) for p in _FRAG_KINDS]
    names += [f"{p}FragType[{k},{f}]" for k in range(K) for f in range(F) for p in _FRAG_KINDS]
    names += [f"{e}[{k},{f}]" for k in range(K) for e in _EMIT_KINDS for f in range(F)]
    names += [f"{p}FragEnd[{k},{f}]" for k in range(K) for f in range(F) for p in _FRAG_KINDS]
    names += ["MatDomEnd", "InsDomEnd", "DelDomEnd"]
    return names


def build_exploded(main_ins, main_del, t, dom_ins, dom_del, dom_w, frag_w, ext):
    """(N, N) exploded transition matrix at one t (exact_suffstats_batch.py:61-253)."""
    dom_ins, dom_del = np.asarray(dom_ins, float), np.asarray(dom_del, float)
    dom_w, frag_w = np.asarray(dom_w, float), np.asarray(frag_w, float)
    ext = np.asarray(ext, float)
    K, F = frag_w.shape
    if ext.ndim == 2:
        ext = np.stack([np.diag(e) for e in ext])
    names = state_names(K, F)
    ix = {n: i for i, n in enumerate(names)}
    T = np.zeros((len(names), len(names)))
    tau0 = tkf91_trans(main_ins, main_del, t)
    tauk = [tkf91_trans(dom_ins[k], dom_del[k], t) for k in range(K)]
    kap = dom_ins / dom_del
    for src, row in (("Start", S), ("MatDomEnd", S), ("InsDomEnd", S), ("DelDomEnd", D)):
        for dst, col in (("MatDom", M), ("InsDom", I), ("DelDom", D), ("End", E)):
            T[ix[src], ix[dst]] = tau0[row, col]
    for k in range(K):
        for p in ("Mat", "Ins", "Del"):
            T[ix[f"{p}Dom"], ix[f"{p}DomType[{k}]"]] = dom_w[k]
        for dst, col in ((f"MatFrag[{k}]", M), (f"InsFrag[{k}]", I), (f"DelFrag[{k}]", D),
                         ("MatDomEnd", E)):
            T[ix[f"MatDomType[{k}]"], ix[dst]] = tauk[k][S, col]
        for p, fr, end in (("Ins", "I", "InsDomEnd"), ("Del", "D", "DelDomEnd")):
            T[ix[f"{p}DomType[{k}]"], ix[f"{fr}Frag[{k}]"]] = kap[k]
            T[ix[f"{p}DomType[{k}]"], ix[end]] = 1.0 - kap[k]
        for f in range(F):
            for e_i, p in enumerate(_FRAG_KINDS):
                T[ix[f"{p}Frag[{k}]"], ix[f"{p}FragType[{k},{f}]"]] = frag_w[k, f]
                T[ix[f"{p}FragType[{k},{f}]"], ix[f"{_EMIT_KINDS[e_i]}[{k},{f}]"]] = 1.0
                T[ix[f"{_EMIT_KINDS[e_i]}[{k},{f}]"], ix[f"{p}FragEnd[{k},{f}]"]] = 1.0
            notext = 1.0 - ext[k, f].sum()
            for g in range(F):
                for p in _FRAG_KINDS:
                    T[ix[f"{p}FragEnd[{k},{f}]"], ix[f"{p}FragType[{k},{g}]"]] = ext[k, f, g]
            for p, row in (("Mat", M), ("Ins", I), ("Del", D)):
                for dst, col in ((f"MatFrag[{k}]", M), (f"InsFrag[{k}]", I),
                                 (f"DelFrag[{k}]", D), ("MatDomEnd", E)):
                    T[ix[f"{p}FragEnd[{k},{f}]"], ix[dst]] = notext * tauk[k][row, col]
            for fr, end in (("I", "InsDomEnd"), ("D", "DelDomEnd")):
                T[ix[f"{fr}FragEnd[{k},{f}]"], ix[f"{fr}Frag[{k}]"]] = notext * kap[k]
                T[ix[f"{fr}FragEnd[{k},{f}]"], ix[end]] = notext * (1.0 - kap[k])
    return T, names


def elim_groups(K, F):
    """exact_suffstats_batch.py:468-482."""
    return [
        [f"{p}FragType[{k},{f}]" for k in range(K) for f in range(F) for p in _FRAG_KINDS],
        [f"{p}FragEnd[{k},{f}]" for k in range(K) for f in range(F) for p in _FRAG_KINDS],
        [f"{p}Frag[{k}]" for k in range(K) for p in _FRAG_KINDS],
        [f"{p}DomType[{k}]" for k in range(K) for p in ("Mat", "Ins", "Del")],
        ["MatDom", "InsDom", "DelDom"],
        ["MatDomEnd", "InsDomEnd", "DelDomEnd"],
    ]


def eliminate_chain(T, names, K, F):
    """Returns the reduced matrix, its state names and the steps
    [(keep, elim, C, T_pre)] (exact_suffstats_batch.py:521-533)."""
    steps = []
    nm = list(names)
    for grp in elim_groups(K, F):
        pos = {n: i for i, n in enumerate(nm)}
        el = sorted(pos[n] for n in grp if n in pos)
        ke = [i for i in range(len(nm)) if i not in el]
        C = np.linalg.inv(np.eye(len(el)) - T[np.ix_(el, el)])
        Tr = T[np.ix_(ke, ke)] + T[np.ix_(ke, el)] @ C @ T[np.ix_(el, ke)]
        steps.append((ke, el, C, T))
        T = Tr
        nm = [nm[i] for i in ke]
    return T, nm, steps


def restore(n_red, T, keep, elim, C):
    """Ghost-usage restoration of one step (exact_suffstats_batch.py:285-339)."""
    TKK, TKZ = T[np.ix_(keep, keep)], T[np.ix_(keep, elim)]
    TZK, TZZ = T[np.ix_(elim, keep)], T[np.ix_(elim, elim)]
    Tr = TKK + TKZ @ C @ TZK
    safe = np.abs(Tr) > 1e-30
    scale = np.where(safe, n_red / np.where(safe, Tr, 1.0), 0.0)
    Hm = C @ TZK
    CtS = (TKZ @ C).T @ scale
    n = np.zeros_like(T)
    n[np.ix_(keep, keep)] = scale * TKK
    n[np.ix_(keep, elim)] = TKZ * (scale @ Hm.T)
    n[np.ix_(elim, keep)] = TZK * CtS
    n[np.ix_(elim, elim)] = TZZ * (CtS @ Hm.T)
    return n


def extract(n_full, names, K, F):
    """exact_suffstats_batch.py:347-458."""
    ix = {n: i for i, n in enumerate(names)}
    top = np.zeros((5, 5))
    dmap = {"MatDom": M, "InsDom": I, "DelDom": D}
    emap = {"MatDomEnd": M, "InsDomEnd": I, "DelDomEnd": D}
    for v, V in dmap.items():
        top[S, V] += n_full[ix["Start"], ix[v]]
    top[S, E] += n_full[ix["Start"], ix["End"]]
    for u, U in emap.items():
        for v, V in dmap.items():
            top[U, V] += n_full[ix[u], ix[v]]
        top[U, E] += n_full[ix[u], ix["End"]]
    domM = np.zeros((K, 5, 5))
    kap = np.zeros(K)
    kap1 = np.zeros(K)
    ext = np.zeros((K, F, F))
    term = np.zeros((K, F))
    dw = np.zeros(K)
    fw = np.zeros((K, F))
    for k in range(K):
        dests = ((f"MatFrag[{k}]", M), (f"InsFrag[{k}]", I), (f"DelFrag[{k}]", D), ("MatDomEnd", E))
        for dst, col in dests:
            domM[k, S, col] = n_full[ix[f"MatDomType[{k}]"], ix[dst]]
        for f in range(F):
            for p, X in (("Mat", M), ("Ins", I), ("Del", D)):
                for dst, col in dests:
                    domM[k, X, col] += n_full[ix[f"{p}FragEnd[{k},{f}]"], ix[dst]]
        for p, fr, end in (("Ins", "I", "InsDomEnd"), ("Del", "D", "DelDomEnd")):
            kap[k] += n_full[ix[f"{p}DomType[{k}]"], ix[f"{fr}Frag[{k}]"]]
            kap1[k] += n_full[ix[f"{p}DomType[{k}]"], ix[end]]
            for f in range(F):
                kap[k] += n_full[ix[f"{fr}FragEnd[{k},{f}]"], ix[f"{fr}Frag[{k}]"]]
                kap1[k] += n_full[ix[f"{fr}FragEnd[{k},{f}]"], ix[end]]
        for f in range(F):
            for p in _FRAG_KINDS:
                fe = ix[f"{p}FragEnd[{k},{f}]"]
                tot = 0.0
                for g in range(F):
                    v = n_full[fe, ix[f"{p}FragType[{k},{g}]"]]
                    ext[k, f, g] += v
                    tot += v
                term[k, f] += n_full[fe, :].sum() - tot
                fw[k, f] += n_full[ix[f"{p}Frag[{k}]"], ix[f"{p}FragType[{k},{f}]"]]
        for p in ("Mat", "Ins", "Del"):
            dw[k] += n_full[:, ix[f"{p}DomType[{k}]"]].sum()
    return {"top_5x5": top, "dom_M_5x5": domM, "dom_kappa": kap, "dom_1mkappa": kap1,
            "ext": ext, "term": term, "dom_w": dw, "frag_w": fw}


def exact_suffstats(n_chi, main_ins, main_del, t, dom_ins, dom_del, dom_w, frag_w, ext):
    """One pair: restore n_chi (collapsed counts) to the exploded machine
    and extract the per-group statistics.  Also returns the restored counts."""
    K, F = np.asarray(frag_w).shape
    T, names = build_exploded(main_ins, main_del, t, dom_ins, dom_del, dom_w, frag_w, ext)
    _, _, steps = eliminate_chain(T, names, K, F)
    n = np.asarray(n_chi, dtype=np.float64)
    for keep, elim, C, Tpre in reversed(steps):
        n = restore(n, Tpre, keep, elim, C)
    out = extract(n, names, K, F)
    out["n_full"] = n
    return out
