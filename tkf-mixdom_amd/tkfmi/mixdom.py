"""Collapsed MixDom chi in torch (FP64, differentiable, batched over t).

chi is obtained as the reduced matrix of the exploded MixDom machine -- the
route libtkfmi's device builder takes -- instead of the reference's
closed-form collapse (models/mixdom.py:104-214, 252-423); both give the same
matrix, and the oracle checks that identity.

The exploded machine (the reference's fully exploded model,
models/exact_suffstats_batch.py:61-253) has N0 = 8 + 8K + 15KF states: Start,
End, the three top-level domain-type states, per domain three domain-type
states and five fragment-type states, per (domain, fragment) five
fragment-entry, five emitting and five fragment-end states, and three
domain-end states.  Every transition is the product of at most two "atoms":
an entry of the top-level TKF91 matrix or of a domain's TKF91 matrix at t,
a domain weight, kappa or 1 - kappa of a domain, a fragment weight, a
fragment extension probability, or the non-extension probability
1 - sum_g ext[d,f,g].  Keeping Start, End and the 5KF emitting states and
eliminating the null states in one block,

    chi = T_KK + T_KZ (I - T_ZZ)^-1 T_ZK,

gives the collapsed matrix in the reference's index order: 0 = Start (SS),
1 = End (EE), body state 2 + d*5F + uv*F + f with uv in (MM, MI, MD, II, DD)
(models/mixdom.py:48-77).  The solve is one batched LU per call and never
waits for the GPU; autograd differentiates through it.
"""
import functools

import numpy as np
import torch

from .params import S, M, I, D, E, tkf91_trans

# emission type of body slot uv = MM, MI, MD, II, DD (models/mixdom.py:66)
_UV_TYPE = (M, I, D, I, D)


def n_states(K, F):
    return 2 + 5 * K * F


def state_types(K, F):
    body = []
    for _d in range(K):
        for uv in range(5):
            body += [_UV_TYPE[uv]] * F
    return [S, E] + body


class _Plan:
    """Index structure of the exploded machine of a (K, F) model: the flat
    position of every transition and the two atoms whose product is its
    value, plus the kept / eliminated state lists."""

    def __init__(self, K, F):
        self.K, self.F = K, F
        self._on_dev = {}
        N0 = 8 + 8 * K + 15 * K * F
        self.N0 = N0
        # atom offsets: tau_main (25) | tau_dom (K*25) | dom_w (K) | kappa (K) |
        # 1 - kappa (K) | frag_w (K*F) | ext (K*F*F) | notext (K*F) | one
        self.o_tau0 = 0
        self.o_taud = 25
        self.o_w = self.o_taud + 25 * K
        self.o_kap = self.o_w + K
        self.o_kap1 = self.o_kap + K
        self.o_fw = self.o_kap1 + K
        self.o_ext = self.o_fw + K * F
        self.o_notext = self.o_ext + K * F * F
        self.o_one = self.o_notext + K * F
        self.n_atoms = self.o_one + 1

        start, end = 0, 1
        dom = lambda p: 2 + p                                           # noqa: E731
        domtype = lambda k, p: 5 + 3 * k + p                            # noqa: E731
        frag = lambda k, q: 5 + 3 * K + 5 * k + q                       # noqa: E731
        fragtype = lambda k, f, q: 5 + 8 * K + (k * F + f) * 5 + q      # noqa: E731
        emit = lambda k, q, f: 5 + 8 * K + 5 * K * F + k * 5 * F + q * F + f  # noqa: E731
        fragend = lambda k, f, q: 5 + 8 * K + 10 * K * F + (k * F + f) * 5 + q  # noqa: E731
        domend = lambda p: 5 + 8 * K + 15 * K * F + p                   # noqa: E731

        one = self.o_one
        tau0 = lambda r, c: self.o_tau0 + 5 * r + c                     # noqa: E731
        taud = lambda k, r, c: self.o_taud + 25 * k + 5 * r + c         # noqa: E731
        rows, cols, a1, a2 = [], [], [], []

        def put(r, c, x=one, y=one):
            rows.append(r); cols.append(c); a1.append(x); a2.append(y)

        # top level: Start and the domain ends re-enter the top TKF91 chain
        # (a domain end is the top-level chain's visit to its emitting state)
        for src, row in ((start, S), (domend(0), M), (domend(1), I), (domend(2), D)):
            for p, col in enumerate((M, I, D)):
                put(src, dom(p), tau0(row, col))
            put(src, end, tau0(row, E))
        for k in range(K):
            for p in range(3):
                put(dom(p), domtype(k, p), self.o_w + k)
            # a Mat-type domain runs the domain's own TKF91 chain; Ins/Del-type
            # domains emit only their own kind, continuing with probability kappa
            for q, col in enumerate((M, I, D)):
                put(domtype(k, 0), frag(k, q), taud(k, S, col))
            put(domtype(k, 0), domend(0), taud(k, S, E))
            put(domtype(k, 1), frag(k, 3), self.o_kap + k)
            put(domtype(k, 1), domend(1), self.o_kap1 + k)
            put(domtype(k, 2), frag(k, 4), self.o_kap + k)
            put(domtype(k, 2), domend(2), self.o_kap1 + k)
            for f in range(F):
                nx = self.o_notext + k * F + f
                for q in range(5):
                    put(frag(k, q), fragtype(k, f, q), self.o_fw + k * F + f)
                    put(fragtype(k, f, q), emit(k, q, f))
                    put(emit(k, q, f), fragend(k, f, q))
                    for g in range(F):
                        put(fragend(k, f, q), fragtype(k, g, q), self.o_ext + (k * F + f) * F + g)
                for q, row in enumerate((M, I, D)):
                    for p, col in enumerate((M, I, D)):
                        put(fragend(k, f, q), frag(k, p), nx, taud(k, row, col))
                    put(fragend(k, f, q), domend(0), nx, taud(k, row, E))
                put(fragend(k, f, 3), frag(k, 3), nx, self.o_kap + k)
                put(fragend(k, f, 3), domend(1), nx, self.o_kap1 + k)
                put(fragend(k, f, 4), frag(k, 4), nx, self.o_kap + k)
                put(fragend(k, f, 4), domend(2), nx, self.o_kap1 + k)
        flat = np.asarray(rows) * N0 + np.asarray(cols)
        if len(np.unique(flat)) != len(flat):
            raise AssertionError("exploded machine has a repeated transition")
        self.flat, self.a1, self.a2 = flat, np.asarray(a1), np.asarray(a2)
        first_emit = emit(0, 0, 0)
        self.keep = np.array([start, end] + list(range(first_emit, first_emit + 5 * K * F)))
        mask = np.ones(N0, bool)
        mask[self.keep] = False
        self.elim = np.nonzero(mask)[0]
        # the native elimination's view (libtkfmi tkfmi_chi_forward): int32
        # transition lists and each state's index among the kept / eliminated
        self.kmap = np.full(N0, -1, np.int32)
        self.kmap[self.keep] = np.arange(len(self.keep), dtype=np.int32)
        self.zmap = np.full(N0, -1, np.int32)
        self.zmap[self.elim] = np.arange(len(self.elim), dtype=np.int32)
        self.flat32, self.a1_32, self.a2_32 = (x.astype(np.int32) for x in (flat, self.a1, self.a2))


    def on(self, name, dev):
        """Index array ``name`` as a tensor on ``dev``, uploaded once: a
        call then issues no host-to-device copy (and can be captured in a
        graph)."""
        key = (name, str(dev))
        v = self._on_dev.get(key)
        if v is None:
            v = torch.as_tensor(getattr(self, name), device=dev)
            self._on_dev[key] = v
        return v


@functools.lru_cache(maxsize=None)
def _plan(K, F):
    return _Plan(K, F)


def _f64(x, dev):
    return x.to(device=dev, dtype=torch.float64) if isinstance(x, torch.Tensor) \
        else torch.as_tensor(x, dtype=torch.float64, device=dev)


def _atoms(main_lam, main_mu, t, dom_lam, dom_mu, dom_w, frag_w, ext_rates):
    """The exploded machine's atoms (..., n_atoms) at t and its plan."""
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t, dtype=torch.float64)
    t = t.to(torch.float64)
    dev = t.device
    dom_lam, dom_mu, dom_w, frag_w, ext = (_f64(x, dev) for x in
                                           (dom_lam, dom_mu, dom_w, frag_w, ext_rates))
    if ext.dim() == 2:
        ext = torch.diag_embed(ext)
    K, F = frag_w.shape
    pl = _plan(K, F)
    bt = t.shape
    tau0 = tkf91_trans(_f64(main_lam, dev), _f64(main_mu, dev), t)          # (...,5,5)
    taud = tkf91_trans(dom_lam, dom_mu, t[..., None])                        # (...,K,5,5)
    kap = dom_lam / dom_mu
    notext = 1.0 - ext.sum(dim=-1)
    fixed = torch.cat([dom_w, kap, 1.0 - kap, frag_w.reshape(-1), ext.reshape(-1),
                       notext.reshape(-1), torch.ones(1, dtype=torch.float64, device=dev)])
    atoms = torch.cat([tau0.reshape(bt + (25,)), taud.reshape(bt + (25 * K,)),
                       fixed.expand(bt + fixed.shape)], dim=-1)
    return atoms, pl


def _scatter(atoms, pl):
    dev = atoms.device
    bt = atoms.shape[:-1]
    vals = atoms[..., pl.on("a1", dev)] * atoms[..., pl.on("a2", dev)]
    T = torch.zeros(bt + (pl.N0 * pl.N0,), dtype=torch.float64, device=dev)
    T = T.scatter(-1, pl.on("flat", dev).expand(bt + (len(pl.flat),)), vals)
    return T.reshape(bt + (pl.N0, pl.N0))


def exploded_trans(main_lam, main_mu, t, dom_lam, dom_mu, dom_w, frag_w, ext_rates):
    """The exploded transition matrix (..., N0, N0) at t (scalar or (B,));
    ext_rates (K,F) is put on the diagonal of (K,F,F) (models/mixdom.py:284-287)."""
    return _scatter(*_atoms(main_lam, main_mu, t, dom_lam, dom_mu, dom_w, frag_w, ext_rates))


def build_nested_trans(main_lam, main_mu, t, dom_lam, dom_mu, dom_w, frag_w, ext_rates):
    """Collapsed MixDom chi (..., N, N), N = 2 + 5KF, the reference's
    build_nested_trans (models/mixdom.py:252-423) obtained by eliminating the
    exploded machine's null states.  ext_rates (K,F) is made diagonal
    (MixDom1 format, models/mixdom.py:284-287) or given as (K,F,F).  On the
    GPU the elimination and its adjoint are libtkfmi's (tkfmi.gradops)."""
    atoms, pl = _atoms(main_lam, main_mu, t, dom_lam, dom_mu, dom_w, frag_w, ext_rates)
    if atoms.is_cuda:
        from . import gradops
        return gradops.chi(atoms, pl)
    T = _scatter(atoms, pl)
    dev = T.device
    kp = pl.on("keep", dev)
    ez = pl.on("elim", dev)
    T_kk = T[..., kp[:, None], kp[None, :]]
    T_kz = T[..., kp[:, None], ez[None, :]]
    T_zk = T[..., ez[:, None], kp[None, :]]
    T_zz = T[..., ez[:, None], ez[None, :]]
    eye = torch.eye(len(pl.elim), dtype=torch.float64, device=dev)
    # I - T_ZZ is a nonsingular M-matrix (the null states reach End), so no
    # error check is needed; solve_ex does not wait for the GPU
    W = torch.linalg.solve_ex(eye - T_zz, T_zk)[0]
    return T_kk + T_kz @ W
