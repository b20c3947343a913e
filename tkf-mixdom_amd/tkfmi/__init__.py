"""tkfmi -- MI355X-native pair-HMM forward/backward for TKF91/TKF92/MixDom2.

Drop-in replacements for the 2D dynamic-programming path of tkfmixdom.jax
(dp/hmm.py forward_backward_2d and its callers), running hand-written HIP
kernels for gfx950 through a C ABI (libtkfmi.so, bound with ctypes).
"""
from . import _abi, tables  # noqa: F401
from .hmm import (NEG_INF, _pad_to_bin, fb2d_batch,  # noqa: F401
                  forward_backward_2d, safe_log)

__all__ = ["fb2d_batch", "forward_backward_2d", "safe_log", "NEG_INF", "_pad_to_bin"]
