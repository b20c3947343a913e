"""Native builders of the gradient path with their adjoints, as
torch.autograd.Functions over CUDA FP64 tensors (libtkfmi's
tkfmi_tkf_trans_jac, tkfmi_chi_forward / tkfmi_chi_backward,
tkfmi_expm_batched / tkfmi_expm_vjp_batched).

tkfmi.params (TKF transition matrices), tkfmi.mixdom (the collapsed MixDom
chi) and tkfmi.vjp.expm call these for CUDA tensors, so a value + gradient
step through tkfmi.vjp runs a few kernels per builder where the torch form
launches hundreds of tiny ones; for CPU tensors they keep their torch form.
The derivatives are the same functions the reference obtains by autodiff of
its JAX builders (train/vjp.py): the TKF closed forms differentiated in
forward-mode dual arithmetic, the elimination by its linear-solve adjoint,
exp by the adjoint of its Frechet derivative.
"""
import torch
from torch.autograd.function import once_differentiable

from . import _abi

_MODES = {"tkf91": 0, "tkf91_cond": 1, "tkf92": 2}


def _stream(dev):
    return _abi.current_stream_ptr(dev)


class _TkfTrans(torch.autograd.Function):
    """T (..., 5, 5) of TKF91 / conditioned TKF91 / TKF92 at broadcast
    (lam, mu, t, ext); the forward also keeps dT / d(lam, mu, t, ext)."""

    @staticmethod
    def forward(ctx, lam, mu, t, ext, mode):
        shapes = (lam.shape, mu.shape, t.shape, ext.shape if ext is not None else None)
        shape = torch.broadcast_shapes(*(s for s in shapes if s is not None))
        dev = t.device
        flat = [x.detach().to(torch.float64).expand(shape).contiguous() if x is not None else None
                for x in (lam, mu, t, ext)]
        n = 1
        for d in shape:
            n *= d
        T = torch.empty(tuple(shape) + (5, 5), dtype=torch.float64, device=dev)
        J = torch.empty(tuple(shape) + (5, 5, 4), dtype=torch.float64, device=dev)
        L = _abi.lib()
        _abi.check(L.tkfmi_tkf_trans_jac(_abi._ptr(flat[0]), _abi._ptr(flat[1]), _abi._ptr(flat[2]),
                                         _abi._ptr(flat[3]), n, mode, _abi._ptr(T), _abi._ptr(J),
                                         _stream(dev)))
        ctx.save_for_backward(J)
        ctx.shapes = shapes
        return T

    @staticmethod
    @once_differentiable
    def backward(ctx, G):
        (J,) = ctx.saved_tensors
        g = (G.to(torch.float64)[..., None] * J).sum(dim=(-3, -2))  # (..., 4)
        out = []
        for k, shp in enumerate(ctx.shapes):
            out.append(g[..., k].sum_to_size(shp) if shp is not None and ctx.needs_input_grad[k] else None)
        return out[0], out[1], out[2], out[3], None


def tkf_trans(lam, mu, t, ext=None, kind="tkf91"):
    """TKF transition matrix (..., 5, 5) on the GPU (tkfmi.params semantics)."""
    return _TkfTrans.apply(lam, mu, t, ext, _MODES[kind])


class _Chi(torch.autograd.Function):
    """chi (..., nk, nk) from the exploded machine's atoms (..., n_atoms)."""

    @staticmethod
    def forward(ctx, atoms, pl):
        a = atoms.detach().to(torch.float64).contiguous()
        bt = a.shape[:-1]
        B = a.numel() // pl.n_atoms
        nk, nz = len(pl.keep), len(pl.elim)
        dev = a.device
        chi = torch.empty(tuple(bt) + (nk, nk), dtype=torch.float64, device=dev)
        W = torch.empty(tuple(bt) + (nz, nk), dtype=torch.float64, device=dev)
        L = _abi.lib()
        wsb = L.tkfmi_chi_workspace_bytes(B, nk, nz)
        ws = torch.empty(max(wsb // 8, 1), dtype=torch.float64, device=dev)
        _abi.check(L.tkfmi_chi_forward(_abi._ptr(a), B, pl.n_atoms, *_plan_args(pl, dev), _abi._ptr(chi),
                                       _abi._ptr(W), _abi._ptr(ws), wsb, _stream(dev)))
        ctx.save_for_backward(a, W)
        ctx.pl = pl
        return chi

    @staticmethod
    @once_differentiable
    def backward(ctx, G):
        a, W = ctx.saved_tensors
        pl = ctx.pl
        dev = a.device
        B = a.numel() // pl.n_atoms
        nk, nz = len(pl.keep), len(pl.elim)
        Gc = G.to(torch.float64).contiguous()
        ga = torch.empty_like(a)
        L = _abi.lib()
        wsb = L.tkfmi_chi_workspace_bytes(B, nk, nz)
        ws = torch.empty(max(wsb // 8, 1), dtype=torch.float64, device=dev)
        _abi.check(L.tkfmi_chi_backward(_abi._ptr(a), B, pl.n_atoms, *_plan_args(pl, dev), _abi._ptr(Gc),
                                        _abi._ptr(W), _abi._ptr(ga), _abi._ptr(ws), wsb, _stream(dev)))
        return ga, None


def _plan_args(pl, dev):
    return (_abi._ptr(pl.on("a1_32", dev)), _abi._ptr(pl.on("a2_32", dev)), _abi._ptr(pl.on("flat32", dev)),
            len(pl.flat), _abi._ptr(pl.on("kmap", dev)), _abi._ptr(pl.on("zmap", dev)), pl.N0,
            len(pl.keep), len(pl.elim))


def chi(atoms, pl):
    """Collapsed MixDom chi from the atoms of tkfmi.mixdom's exploded machine."""
    return _Chi.apply(atoms, pl)


class _Expm(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A):
        Ac = A.detach().to(torch.float64).contiguous()
        n = Ac.shape[-1]
        out = torch.empty_like(Ac)
        _abi.check(_abi.lib().tkfmi_expm_batched(_abi._ptr(Ac), Ac.numel() // (n * n), n, _abi._ptr(out),
                                                 _stream(Ac.device)))
        ctx.save_for_backward(Ac)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, G):
        (Ac,) = ctx.saved_tensors
        n = Ac.shape[-1]
        Gc = G.to(torch.float64).contiguous()
        out = torch.empty_like(Ac)
        _abi.check(_abi.lib().tkfmi_expm_vjp_batched(_abi._ptr(Ac), _abi._ptr(Gc), Ac.numel() // (n * n), n,
                                                     _abi._ptr(out), _stream(Ac.device)))
        return out


def expm(A):
    """exp of each n x n matrix of A (..., n, n) on the GPU, n <= 20."""
    return _Expm.apply(A)


def expm_ok(A):
    """Whether expm() takes A: a CUDA tensor of matrices of at most 20 x 20
    (tkfmi_expm_vjp_batched's limit, 2 n <= TKFMI_EXPM_MAX_N)."""
    return isinstance(A, torch.Tensor) and A.is_cuda and A.dim() >= 2 and A.shape[-1] == A.shape[-2] \
        and 2 * A.shape[-1] <= _abi.EXPM_MAX_N and A.numel() > 0
