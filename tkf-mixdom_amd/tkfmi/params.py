"""TKF91 / TKF92 transition matrices in torch (FP64, differentiable,
broadcast over leading batch dimensions of t / ext).

Reference: core/bdi.py:40-97 (alpha, beta, gamma, kappa; L'Hopital branch
when |1 - kappa| < 1e-4) and core/params.py:29-108 (tkf91_trans,
tkf91_trans_cond, tkf92_trans).  State order S, M, I, D, E.
"""
import torch

S, M, I, D, E = 0, 1, 2, 3, 4
EQUAL_RATE_THRESHOLD = 1e-4
TKF_STATE_TYPES = (S, M, I, D, E)


def _t(x, like=None):
    if isinstance(x, torch.Tensor):
        return x.to(torch.float64)
    dev = like.device if isinstance(like, torch.Tensor) else None
    return torch.as_tensor(x, dtype=torch.float64, device=dev)


def _common(*xs):
    """All arguments as FP64 tensors on the device of the first tensor among
    them (scalars broadcast to it)."""
    dev = next((x.device for x in xs if isinstance(x, torch.Tensor)), None)
    return tuple(x.to(device=dev, dtype=torch.float64) if isinstance(x, torch.Tensor)
                 else torch.as_tensor(x, dtype=torch.float64, device=dev) for x in xs)


def tkf_alpha(mu, t):
    return torch.exp(-_t(mu) * _t(t))


def tkf_beta(lam, mu, t):
    lam, mu, t = _t(lam), _t(mu, lam), _t(t, lam)
    s = mu * t
    eta = torch.exp(-lam * t)
    alpha = torch.exp(-s)
    numer = lam * (eta - alpha)
    denom = mu * eta - lam * alpha
    small = denom.abs() < 1e-30
    general = torch.where(small, torch.zeros_like(numer),
                          numer / torch.where(small, torch.ones_like(denom), denom))
    limit = s / (1.0 + s)
    kappa = lam / mu
    return torch.where((1.0 - kappa).abs() < EQUAL_RATE_THRESHOLD, limit, general)


def tkf_gamma(lam, mu, t):
    lam, mu, t = _t(lam), _t(mu, lam), _t(t, lam)
    s = mu * t
    alpha = torch.exp(-s)
    beta = tkf_beta(lam, mu, t)
    Phi = 1.0 - alpha
    general = torch.where(Phi.abs() < 1e-30, torch.zeros_like(Phi),
                          1.0 - mu * beta / (lam * torch.clamp(Phi, min=1e-30)))
    phi = torch.where(s < 1e-10, 1.0 - s / 2.0, Phi / torch.clamp(s, min=1e-30))
    limit = 1.0 - 1.0 / ((1.0 + s) * phi)
    kappa = lam / mu
    return torch.where((1.0 - kappa).abs() < EQUAL_RATE_THRESHOLD, limit, general)


def tkf_kappa(lam, mu):
    return _t(lam) / _t(mu, lam)


def _assemble91(a, b, g, k, cond=False):
    """Rows S, M, I use beta, row D uses gamma (core/params.py:29-54)."""
    shape = torch.broadcast_shapes(a.shape, b.shape, g.shape, k.shape)
    a, b, g, k = (x.expand(shape) for x in (a, b, g, k))
    z = torch.zeros(shape, dtype=torch.float64, device=a.device)
    kk = torch.ones_like(k) if cond else k
    e_b = (1 - b) if cond else (1 - b) * (1 - k)
    e_g = (1 - g) if cond else (1 - g) * (1 - k)
    row_b = torch.stack([z, (1 - b) * kk * a, b, (1 - b) * kk * (1 - a), e_b], dim=-1)
    row_d = torch.stack([z, (1 - g) * kk * a, g, (1 - g) * kk * (1 - a), e_g], dim=-1)
    row_e = torch.zeros(shape + (5,), dtype=torch.float64, device=a.device)
    return torch.stack([row_b, row_b, row_b, row_d, row_e], dim=-2)


def _native(*xs):
    """On the GPU the closed forms and their derivatives are libtkfmi's
    (tkfmi.gradops): one kernel instead of a few dozen elementwise ones."""
    return xs[0].is_cuda


def tkf91_trans(lam, mu, t):
    lam, mu, t = _common(lam, mu, t)
    if _native(lam, mu, t):
        from . import gradops
        return gradops.tkf_trans(lam, mu, t, kind="tkf91")
    return _assemble91(tkf_alpha(mu, t), tkf_beta(lam, mu, t), tkf_gamma(lam, mu, t),
                       tkf_kappa(lam, mu))


def tkf91_trans_cond(lam, mu, t):
    """core/params.py:57-89 (kappa and 1-kappa factors removed)."""
    lam, mu, t = _common(lam, mu, t)
    if _native(lam, mu, t):
        from . import gradops
        return gradops.tkf_trans(lam, mu, t, kind="tkf91_cond")
    return _assemble91(tkf_alpha(mu, t), tkf_beta(lam, mu, t), tkf_gamma(lam, mu, t),
                       tkf_kappa(lam, mu), cond=True)


def tkf92_trans(lam, mu, t, ext):
    """core/params.py:92-108: S row unchanged; M/I/D rows scaled by (1-ext)
    with ext added on the diagonal."""
    lam, mu, t, ext = _common(lam, mu, t, ext)
    if _native(lam, mu, t, ext):
        from . import gradops
        return gradops.tkf_trans(lam, mu, t, ext, kind="tkf92")
    t91 = tkf91_trans(lam, mu, t)
    (ext,) = _common(ext) if isinstance(ext, torch.Tensor) else (torch.as_tensor(ext, dtype=torch.float64, device=t91.device),)
    ext = ext.to(t91.device)
    ext = ext.expand(t91.shape[:-2]) if ext.dim() <= t91.dim() - 2 else ext
    e = ext[..., None, None]
    scale = torch.ones_like(t91)
    scale[..., M:D + 1, :] = 0
    out = t91 * torch.where(scale > 0, torch.ones_like(t91), (1 - e).expand_as(t91))
    eye = torch.zeros_like(t91)
    for s_ in (M, I, D):
        eye[..., s_, s_] = 1.0
    return out + eye * e
