"""Per-pair transition matrices and emission tables built on the device
(libtkfmi ``tkfmi_build_params_batched``): one launch per batch replaces the
per-pair torch builders when no gradient is needed (E-steps, serving).

A :class:`DeviceBuilder` holds everything that does not depend on t: the
model's structural parameters, the substitution rate matrices (any Q: the
kernel exponentiates Q t by Pade scaling and squaring, as the reference's
transition_matrix does, core/ctmc.py:31-60, so reversibility is not
assumed) and the class distribution.  ``build(ts)`` then returns
(log_trans (B,S,S), log_tab (B,G,TAB)) for per-pair t.
"""
import ctypes

import numpy as np
import torch

from . import _abi, hmm, mixdom, params, tables
from .suffstats import _Model as _MixModel

MODELS = {"tkf91": 0, "tkf91_cond": 1, "tkf92": 2, "mixdom": 3}
EMIT = {"shared": 0, "per_domain": 1, "per_class": 2}


def _np64(x):
    return np.ascontiguousarray(np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x,
                                           dtype=np.float64))


class DeviceBuilder:
    """model: 'tkf91' | 'tkf91_cond' | 'tkf92' | 'mixdom';
    emission: 'shared' | 'per_domain' | 'per_class';
    log_floor: False -> safe_log (-1e30 for zeros), True -> log(max(x, 1e-30)).

    TKF models take lam, mu (and ext); MixDom takes ``mix`` = dict(main_ins,
    main_del, dom_ins, dom_del, dom_w, frag_w, ext).  Substitution: Qs (C,A,A)
    and pis (C,A) (C = 1 shared, K per-domain, classes per-class); per-class
    also class_dist (K,F,C)."""

    def __init__(self, model, emission, Qs, pis, *, lam=0.0, mu=1.0, ext=0.0, mix=None,
                 class_dist=None, log_floor=True, device=None):
        self.L = _abi.lib()
        self.dev = device or hmm.device()
        Qs, pis = _np64(Qs), _np64(pis)
        if Qs.ndim == 2:
            Qs, pis = Qs[None], pis[None]
        C, A = pis.shape
        if Qs.shape != (C, A, A):
            raise ValueError(f"Qs must be ({C}, {A}, {A}), got {Qs.shape}")
        d = lambda a: torch.as_tensor(a, dtype=torch.float64, device=self.dev).contiguous()
        self.Q, self.pi = d(Qs), d(pis)
        self.cd = d(_np64(class_dist)) if class_dist is not None else None
        self.mix = None
        if model == "mixdom":
            self.mix = _MixModel(mix["main_ins"], mix["main_del"], mix["dom_ins"], mix["dom_del"],
                                 mix["dom_w"], mix["frag_w"], mix["ext"])
            K, F = self.mix.K, self.mix.F
            self.state_types = mixdom.state_types(K, F)
        else:
            K, F = 1, 1
            self.state_types = list(params.TKF_STATE_TYPES)
        self.S = len(self.state_types)
        if emission == "per_class":
            self.G = K * F
            self.state_group = tables.state_groups_per_class(self.S, F)
        elif emission == "per_domain":
            self.G = K
            self.state_group = [0, 0] + [(s - 2) // (5 * F) for s in range(2, self.S)]
        else:
            self.G = 1
            self.state_group = [0] * self.S
        self.c = _abi.Builder(MODELS[model], EMIT[emission], int(bool(log_floor)), C, A, 0,
                              float(lam), float(mu), float(ext),
                              ctypes.pointer(self.mix.c) if self.mix else None,
                              _abi._ptr(self.Q), _abi._ptr(self.pi),
                              _abi._ptr(self.cd))

    def build(self, ts, stream=None):
        t = torch.as_tensor(_np64(ts).reshape(-1), device=self.dev)
        B = int(t.shape[0])
        lt = torch.empty((B, self.S, self.S), dtype=torch.float64, device=self.dev)
        tab = torch.empty((B, self.G, tables.TAB_SZ), dtype=torch.float64, device=self.dev)
        wsb = int(self.L.tkfmi_build_workspace_bytes(ctypes.byref(self.c), B))
        if wsb == 0:
            self.L.tkfmi_build_params_batched(ctypes.byref(self.c), B, None, None, None, None, 0, None)
            raise _abi.TkfmiError(self.L.tkfmi_last_error().decode() or "invalid builder")
        ws = torch.empty(wsb, dtype=torch.uint8, device=self.dev)
        sp = ctypes.c_void_p((stream or torch.cuda.current_stream(self.dev)).cuda_stream)
        _abi.check(self.L.tkfmi_build_params_batched(ctypes.byref(self.c), B, _abi._ptr(t),
                                                     _abi._ptr(lt), _abi._ptr(tab), _abi._ptr(ws),
                                                     ctypes.c_size_t(wsb), sp))
        self._keep = (ws, t)  # in use until the launch completes (stream-ordered reuse)
        return lt, tab, t
