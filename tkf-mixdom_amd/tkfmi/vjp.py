"""Differentiable pair log-likelihoods (reference: train/vjp.py).

The gradient of a pair-HMM log-likelihood with respect to any parameter
theta is the score identity the reference's custom VJPs use
(train/vjp.py:156-191, 282-326, 428-455, 770-817):

    d logP / d theta = sum_{k,k'} n[k,k'] d log T[k,k'] / d theta
                     + sum_{g,e}  c[g,e] d log tab[g,e] / d theta

where n are the expected transition counts and c the expected emission
counts, both produced by the backward DP kernel.  :class:`FB2DLogProb` is a
``torch.autograd.Function`` whose backward returns exactly (g n, g c) for the
log transition matrices and log emission tables it was given; torch autograd
then carries them through the differentiable builders (tkfmi.params,
tkfmi.mixdom, tkfmi.tables, matrix_exp) to lambda, mu, t, ext, the domain /
fragment weights, exchangeabilities, pi and the class distribution.  The DP
itself is never differentiated by autograd -- the hand-written backward
sweep is its adjoint.

Every function takes a scalar ``t`` and one (x, y) pair like the reference,
and has a ``*_batch`` form over a ragged batch with per-pair ``ts`` that runs
as one kernel call.
"""
import torch
from torch.autograd.function import once_differentiable

from . import _abi, hmm, mixdom, params, tables


class _Spec:
    """Non-tensor inputs of one batched call."""

    def __init__(self, state_types, state_group, xs, ys, device, generic=False, kernel=None,
                 checkpoint=False):
        self.state_types = hmm._types_list(state_types)
        self.state_group = list(state_group) if state_group is not None else [0] * len(self.state_types)
        self.pairs = _abi.PackedPairs(list(xs), list(ys), device)
        self.generic = generic
        self.kernel = kernel
        self.checkpoint = checkpoint


class FB2DLogProb(torch.autograd.Function):
    """log P of each pair of a batch; backward = expected counts.

    log_trans: (S,S) shared or (B,S,S) per pair; log_tab: (G,TAB) shared or
    (B,G,TAB) per pair.  Returns (B,) log-likelihoods."""

    @staticmethod
    def forward(ctx, log_trans, log_tab, spec):
        r = _abi.BatchRunner(log_trans.detach(), spec.state_types, spec.state_group,
                             log_tab.detach(), spec.pairs, want_emit_counts=True,
                             generic=spec.generic, kernel=spec.kernel, checkpoint=spec.checkpoint)
        r.run(zero_outputs=False)
        ctx.save_for_backward(r.n, r.ecnt)
        ctx.shared = (log_trans.dim() == 2, log_tab.dim() == 2)
        return r.log_p

    @staticmethod
    @once_differentiable
    def backward(ctx, g):
        n, ec = ctx.saved_tensors
        g = g.to(torch.float64)[:, None, None]
        gt, ge = n * g, ec * g
        if ctx.shared[0]:
            gt = gt.sum(0)
        if ctx.shared[1]:
            ge = ge.sum(0)
        return gt, ge, None


def log_prob_batch(log_trans, state_types, xs, ys, log_tab, state_group=None, generic=False,
                   kernel=None, checkpoint=False):
    """Differentiable (B,) log-likelihoods of a ragged batch (one call).
    ``generic`` / ``kernel`` select the kernel and ``checkpoint`` turns on
    lattice checkpointing (the reference's rematerialisation in its
    gradient path) as in tkfmi.hmm.fb2d_batch."""
    dev = log_trans.device if isinstance(log_trans, torch.Tensor) else hmm.device()
    spec = _Spec(state_types, state_group, xs, ys, dev, generic, kernel, checkpoint)
    return FB2DLogProb.apply(log_trans.to(torch.float64), log_tab.to(torch.float64), spec)


# ------------------------------------------------------------ substitution
# Pade(13) numerator coefficients b_0..b_13 and the 1-norm bound theta_13 of
# the scaling and squaring method (Higham, "The scaling and squaring method
# for the matrix exponential revisited", SIAM J. Matrix Anal. Appl. 26, 2005,
# Table 2.3), the method of scipy.linalg.expm and jax.scipy.linalg.expm.
_PADE13 = (64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
           129060195264000.0, 10559470521600.0, 670442572800.0, 33522128640.0, 1323241920.0,
           40840800.0, 960960.0, 16380.0, 182.0, 1.0)
_THETA13 = 5.371920351148152
_MAX_SQUARINGS = 16  # 1-norms up to theta_13 * 2**16 (~3.5e5)


def expm(A):
    """exp(A) for a batch of square FP64 matrices.  On the GPU (matrices up
    to 20 x 20) libtkfmi's batched kernel, by the rule of
    jax.scipy.linalg.expm, with the adjoint of exp's Frechet derivative as
    its backward (tkfmi.gradops).  Otherwise degree-13 Pade with scaling and
    squaring in torch: the scaling exponent is chosen per matrix on the
    device and the squarings are masked, so the call never waits for the GPU
    (torch.linalg.matrix_exp copies norms to the host to pick its degree);
    autograd differentiates through it.  Matrices with 1-norm above
    theta_13 * 2**16 get too few squarings there, NaN on the GPU route."""
    from . import gradops
    if gradops.expm_ok(A):
        return gradops.expm(A)
    b = _PADE13
    eye = torch.eye(A.shape[-1], dtype=A.dtype, device=A.device)
    norm = A.abs().sum(-2).amax(-1)
    s = torch.clamp(torch.ceil(torch.log2(norm / _THETA13)), min=0.0, max=float(_MAX_SQUARINGS))
    s = torch.nan_to_num(s, nan=0.0).detach()
    A = A * torch.exp2(-s)[..., None, None]
    A2 = A @ A
    A4 = A2 @ A2
    A6 = A4 @ A2
    U = A @ (A6 @ (b[13] * A6 + b[11] * A4 + b[9] * A2) + b[7] * A6 + b[5] * A4 + b[3] * A2 + b[1] * eye)
    V = A6 @ (b[12] * A6 + b[10] * A4 + b[8] * A2) + b[6] * A6 + b[4] * A4 + b[2] * A2 + b[0] * eye
    X = torch.linalg.solve_ex(V - U, V + U)[0]
    for k in range(_MAX_SQUARINGS):
        X = torch.where((s > k)[..., None, None], X @ X, X)
    return X


def transition_matrix(Q, t):
    """P(t) = expm(Q t) (core/ctmc.py:31-60, Pade with its canonical
    tangent); ``t`` may carry leading batch dimensions."""
    Q = torch.as_tensor(Q, dtype=torch.float64)
    t = torch.as_tensor(t, dtype=torch.float64, device=Q.device)
    return expm(t[..., None, None] * Q)


def build_Q_from_S_pi(S_exch, pi):
    """GTR rate matrix Q[a,b] = S[a,b] pi[b] off the diagonal, rows summing to
    zero, no rate normalisation (core/ctmc.py:63-95; train/vjp.py:463-475)."""
    S_exch = torch.as_tensor(S_exch, dtype=torch.float64)
    pi = torch.as_tensor(pi, dtype=torch.float64, device=S_exch.device)
    A = pi.shape[-1]
    off = 1.0 - torch.eye(A, dtype=torch.float64, device=S_exch.device)
    Q_off = S_exch * pi[..., None, :] * off
    return Q_off - torch.diag_embed(Q_off.sum(dim=-1))


def _dev():
    return hmm.device()


# --------------------------------------------------------- graphed builders
class _GraphedBuilder:
    """A builder's forward and backward captured as two HIP graphs for one
    set of argument shapes (torch.cuda.make_graphed_callables).  The graphs
    replay into fixed buffers, so a call's outputs and saved values stay
    valid only until the next call: a call made while the previous call's
    backward has not run yet is computed eagerly instead."""

    def __init__(self, fn, args):
        sample = tuple(a.detach().clone().requires_grad_(a.requires_grad) for a in args)
        self.g = torch.cuda.make_graphed_callables(fn, sample)
        self.fn = fn
        self.pending = False

    def _done(self, grad):
        self.pending = False
        return grad

    def __call__(self, args):
        if self.pending:
            return self.fn(*args)
        out = self.g(*args)
        self.pending = True
        out[0].register_hook(self._done)
        return out


_GRAPHED = {}


def _built(fn, args, graph):
    """fn(*args) -> (log transitions, emission tables), or, with graph=True
    and a gradient to take, the same through a _GraphedBuilder for these
    argument shapes: a training step over a fixed minibatch shape then
    replays two graphs instead of launching the builders' ~2,600 small
    kernels one by one."""
    if not (graph and torch.is_grad_enabled() and any(a.requires_grad for a in args)):
        return fn(*args)
    key = (fn, tuple((tuple(a.shape), a.requires_grad, str(a.device)) for a in args))
    gb = _GRAPHED.get(key)
    if gb is None:
        gb = _GRAPHED[key] = _GraphedBuilder(fn, args)
    return gb(args)


def _as(x, dev):
    return x.to(device=dev, dtype=torch.float64) if isinstance(x, torch.Tensor) \
        else torch.as_tensor(x, dtype=torch.float64, device=dev)


# ----------------------------------------------------------------- TKF91/92
def _shared_tab(Q, pi, ts):
    sub = transition_matrix(Q, ts)                                 # (B,A,A)
    return tables.shared_tables(sub, pi.expand(sub.shape[:-1]))    # (B,1,TAB)


def _tkf91_build(ins_rate, del_rate, ts, Q, pi):
    return hmm.safe_log(params.tkf91_trans(ins_rate, del_rate, ts)), _shared_tab(Q, pi, ts)


def _tkf91c_build(ins_rate, del_rate, ts, Q, pi):
    return hmm.safe_log(params.tkf91_trans_cond(ins_rate, del_rate, ts)), _shared_tab(Q, pi, ts)


def tkf91_log_prob_batch(ins_rate, del_rate, ts, Q, pi, xs, ys, *, cond=False, generic=False, kernel=None,
                         graph=False):
    """(B,) log P(x_b, y_b | lambda, mu, t_b, Q, pi) under TKF91 with the
    shared emissions of pair_hmm_emissions (train/vjp.py:110-194); ``cond``
    uses the ancestor-length-conditioned matrix (train/vjp.py:198-249).
    ``graph`` replays the parameter builders from HIP graphs (see _built)."""
    dev = _dev()
    args = tuple(_as(x, dev) for x in (ins_rate, del_rate, ts, Q, pi))
    log_trans, tab = _built(_tkf91c_build if cond else _tkf91_build, args, graph)
    return log_prob_batch(log_trans, params.TKF_STATE_TYPES, xs, ys, tab, generic=generic, kernel=kernel)


def tkf91_log_prob(ins_rate, del_rate, t, Q, pi, x_seq, y_seq):
    """Reference surface of train/vjp.py:110-127 (scalar log P)."""
    return tkf91_log_prob_batch(ins_rate, del_rate, _as(t, _dev()).reshape(1), Q, pi,
                                [x_seq], [y_seq])[0]


def tkf91_log_prob_cond(ins_rate, del_rate, t, Q, pi, x_seq, y_seq):
    """Reference surface of train/vjp.py:198-205."""
    return tkf91_log_prob_batch(ins_rate, del_rate, _as(t, _dev()).reshape(1), Q, pi,
                                [x_seq], [y_seq], cond=True)[0]


def _tkf92_build(ins_rate, del_rate, ts, ext, Q, pi):
    return hmm.safe_log(params.tkf92_trans(ins_rate, del_rate, ts, ext)), _shared_tab(Q, pi, ts)


def tkf92_log_prob_batch(ins_rate, del_rate, ts, ext, Q, pi, xs, ys, *, generic=False, kernel=None,
                         graph=False):
    """(B,) TKF92 log-likelihoods (train/vjp.py:253-329)."""
    dev = _dev()
    args = tuple(_as(x, dev) for x in (ins_rate, del_rate, ts, ext, Q, pi))
    log_trans, tab = _built(_tkf92_build, args, graph)
    return log_prob_batch(log_trans, params.TKF_STATE_TYPES, xs, ys, tab, generic=generic, kernel=kernel)


def tkf92_log_prob(ins_rate, del_rate, t, ext, Q, pi, x_seq, y_seq):
    """Reference surface of train/vjp.py:253-261."""
    dev = _dev()
    return tkf92_log_prob_batch(ins_rate, del_rate, _as(t, dev).reshape(1), ext, Q, pi,
                                [x_seq], [y_seq])[0]


# ------------------------------------------------------------------ MixDom
def _mixdom_build(main_ins_rate, main_del_rate, ts, dom_ins_rates, dom_del_rates, dom_weights,
                  frag_weights, ext_rates, Q, pi):
    chi = mixdom.build_nested_trans(main_ins_rate, main_del_rate, ts, dom_ins_rates, dom_del_rates,
                                    dom_weights, frag_weights, ext_rates)
    return hmm.safe_log(chi), _shared_tab(Q, pi, ts)


def mixdom_log_prob_batch(main_ins_rate, main_del_rate, ts, dom_ins_rates, dom_del_rates,
                          dom_weights, frag_weights, ext_rates, Q, pi, xs, ys, *, generic=False, kernel=None,
                          graph=False):
    """(B,) MixDom log-likelihoods with shared emissions (train/vjp.py:351-458)."""
    dev = _dev()
    args = tuple(_as(x, dev) for x in (main_ins_rate, main_del_rate, ts, dom_ins_rates, dom_del_rates,
                                       dom_weights, frag_weights, ext_rates, Q, pi))
    K, F = args[6].shape
    log_trans, tab = _built(_mixdom_build, args, graph)
    return log_prob_batch(log_trans, mixdom.state_types(K, F), xs, ys, tab, generic=generic, kernel=kernel)


def mixdom_log_prob(main_ins_rate, main_del_rate, t, dom_ins_rates, dom_del_rates, dom_weights,
                    frag_weights, ext_rates, Q, pi, x_seq, y_seq):
    """Reference surface of train/vjp.py:351-386."""
    dev = _dev()
    return mixdom_log_prob_batch(main_ins_rate, main_del_rate, _as(t, dev).reshape(1),
                                 dom_ins_rates, dom_del_rates, dom_weights, frag_weights,
                                 ext_rates, Q, pi, [x_seq], [y_seq])[0]


def build_class_subs(class_S_exch, class_pis, ts):
    """Per-class P_c(t_b) = expm(t_b Q_c), Q_c = GTR(S_c, pi_c), not rate
    normalised (train/vjp.py:463-510).  Returns (B, C, A, A)."""
    Qc = build_Q_from_S_pi(class_S_exch, class_pis)               # (C,A,A)
    return transition_matrix(Qc[None], ts[:, None])               # (B,C,A,A)


def _mixdom2_build(main_ins_rate, main_del_rate, ts, dom_ins_rates, dom_del_rates, dom_weights,
                   frag_weights, ext_rates, class_pis, class_S_exch, class_dist):
    chi = mixdom.build_nested_trans(main_ins_rate, main_del_rate, ts, dom_ins_rates, dom_del_rates,
                                    dom_weights, frag_weights, ext_rates)
    subs = build_class_subs(class_S_exch, class_pis, ts)
    tab = tables.per_class_tables(subs, class_pis, class_dist)     # (B,KF,TAB)
    return hmm.safe_log(chi), tab


def mixdom2_log_prob_batch(main_ins_rate, main_del_rate, ts, dom_ins_rates, dom_del_rates,
                           dom_weights, frag_weights, ext_rates, class_pis, class_S_exch,
                           class_dist, xs, ys, *, generic=False, kernel=None, checkpoint=False,
                           graph=False):
    """(B,) MixDom2 log-likelihoods: per-(domain, fragment) class-mixture
    emissions (train/vjp.py:628-820; dp/hmm.py:673-783).  ``graph`` replays
    the parameter builders from HIP graphs (see _built)."""
    dev = _dev()
    args = tuple(_as(x, dev) for x in (main_ins_rate, main_del_rate, ts, dom_ins_rates, dom_del_rates,
                                       dom_weights, frag_weights, ext_rates, class_pis, class_S_exch,
                                       class_dist))
    K, F = args[6].shape
    st = mixdom.state_types(K, F)
    log_trans, tab = _built(_mixdom2_build, args, graph)
    groups = tables.state_groups_per_class(len(st), F)
    return log_prob_batch(log_trans, st, xs, ys, tab, groups, generic=generic, kernel=kernel,
                          checkpoint=checkpoint)


def mixdom2_log_prob(main_ins_rate, main_del_rate, t, dom_ins_rates, dom_del_rates, dom_weights,
                     frag_weights, ext_rates, class_pis, class_S_exch, class_dist, x_seq, y_seq):
    """Reference surface of train/vjp.py:628-679."""
    dev = _dev()
    return mixdom2_log_prob_batch(main_ins_rate, main_del_rate, _as(t, dev).reshape(1),
                                  dom_ins_rates, dom_del_rates, dom_weights, frag_weights,
                                  ext_rates, class_pis, class_S_exch, class_dist,
                                  [x_seq], [y_seq])[0]
