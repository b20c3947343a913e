"""ctypes binding of libtkfmi.so (the C ABI declared in include/tkfmi.h).

The shared library is built in-tree (``tkfmi/_lib/libtkfmi.so``) by
``tkfmi._build.build()``.  There is no fallback: if the library is missing,
:func:`lib` raises, so a GPU run can never silently use another code path.
"""
import ctypes
import os
import threading

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libtkfmi.so")

ABI_VERSION = 6
FORWARD_ONLY = 1
WANT_POSTERIORS = 2
NO_COUNTS = 4
GENERIC_KERNEL = 8
WANT_EMIT_COUNTS = 16
TIME_KERNELS = 32
CLASS_KERNEL = 64
REUSE_DESCRIPTORS = 128
CHECKPOINT = 256
A_PAD = 21
TAB_SZ = A_PAD * A_PAD + 2 * A_PAD

_c_p = ctypes.c_void_p


class Problem(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int32),
        ("S", ctypes.c_int32),
        ("state_types", _c_p),
        ("state_group", _c_p),
        ("n_groups", ctypes.c_int32),
        ("n_trans", ctypes.c_int32),
        ("log_trans", _c_p),
        ("n_tab", ctypes.c_int32),
        ("log_tab", _c_p),
        ("x_cat", _c_p),
        ("y_cat", _c_p),
        ("h_x_off", _c_p),
        ("h_y_off", _c_p),
        ("h_x_len", _c_p),
        ("h_y_len", _c_p),
        ("log_emit", _c_p),
    ]


class Problem1D(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int32),
        ("S", ctypes.c_int32),
        ("s_idx", ctypes.c_int32),
        ("e_idx", ctypes.c_int32),
        ("n_trans", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("log_trans", _c_p),
        ("log_emit", _c_p),
        ("state_types", _c_p),
        ("state_group", _c_p),
        ("n_groups", ctypes.c_int32),
        ("n_tab", ctypes.c_int32),
        ("log_tab", _c_p),
        ("col_type", _c_p),
        ("anc", _c_p),
        ("desc", _c_p),
        ("h_off", _c_p),
        ("h_len", _c_p),
    ]


class Outputs(ctypes.Structure):
    _fields_ = [
        ("log_prob", _c_p),
        ("n_trans", _c_p),
        ("posteriors", _c_p),
        ("emit_counts", _c_p),
        ("status", _c_p),
    ]


class MixdomModel(ctypes.Structure):
    _fields_ = [
        ("K", ctypes.c_int32),
        ("F", ctypes.c_int32),
        ("main_ins", ctypes.c_double),
        ("main_del", ctypes.c_double),
        ("dom_ins", _c_p),
        ("dom_del", _c_p),
        ("dom_w", _c_p),
        ("frag_w", _c_p),
        ("ext", _c_p),
    ]


TKF92_STATS = 35


class Builder(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32),
        ("emission", ctypes.c_int32),
        ("log_floor", ctypes.c_int32),
        ("C", ctypes.c_int32),
        ("A", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("lam", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("ext", ctypes.c_double),
        ("mix", ctypes.POINTER(MixdomModel)),
        ("Q", _c_p),
        ("pi", _c_p),
        ("class_dist", _c_p),
    ]


class PairText(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("x_text", _c_p),
        ("y_text", _c_p),
        ("a_text", _c_p),
        ("h_x_off", _c_p),
        ("h_x_len", _c_p),
        ("h_y_off", _c_p),
        ("h_y_len", _c_p),
        ("h_a_off", _c_p),
        ("h_a_len", _c_p),
    ]


class PairColumns(ctypes.Structure):
    _fields_ = [
        ("x_codes", _c_p),
        ("y_codes", _c_p),
        ("col_type", _c_p),
        ("anc", _c_p),
        ("desc", _c_p),
        ("n_cols", _c_p),
        ("status", _c_p),
    ]


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("layout", ctypes.c_int32),
        ("length_bin", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("best", ctypes.c_int32),
        ("long_w", ctypes.c_int32 * 16),
        ("main_w", ctypes.c_int32 * 16),
        ("long_frac", ctypes.c_double * 16),
        ("ms", ctypes.c_double * 16),
        ("long_one", ctypes.c_int32 * 16),
    ]


class KernelTime(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("n_pairs", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("cells", ctypes.c_int64),
        ("ms", ctypes.c_double),
    ]


_lib = None
_lock = threading.Lock()

# Every symbol include/tkfmi.h declares (checked by the CPU test-suite).
EXPORTED = ("tkfmi_abi_version", "tkfmi_last_error",
            "tkfmi_fb2d_workspace_bytes", "tkfmi_fb2d_batched",
            "tkfmi_fb1d_workspace_bytes", "tkfmi_fb1d_batched",
            "tkfmi_last_kernel_times", "tkfmi_chain_config", "tkfmi_chain_schedule", "tkfmi_suffstats_size",
            "tkfmi_restore_workspace_bytes", "tkfmi_restore_counts_batched",
            "tkfmi_bdi_stats_batched", "tkfmi_tkf92_estep_stats", "tkfmi_hr_stats_batched",
            "tkfmi_build_workspace_bytes", "tkfmi_build_params_batched",
            "tkfmi_fb2d_batched_noalloc", "tkfmi_pairs_decode_workspace_bytes",
            "tkfmi_pairs_decode", "tkfmi_autotune_workspace_bytes", "tkfmi_autotune",
            "tkfmi_checkpoint_budget", "tkfmi_checkpoint_segment", "tkfmi_tkf_trans_jac", "tkfmi_chi_workspace_bytes",
            "tkfmi_chi_forward", "tkfmi_chi_backward", "tkfmi_expm_batched", "tkfmi_expm_vjp_batched")
EXPM_MAX_N = 40  # TKFMI_EXPM_MAX_N
NOALLOC_SCRATCH_BYTES = 576 << 20  # TKFMI_NOALLOC_SCRATCH_BYTES


class TkfmiError(RuntimeError):
    pass


def lib():
    """Load libtkfmi.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise TkfmiError(
                    f"libtkfmi.so not found at {LIB_PATH}; build it with "
                    "tkfmi._build.build() (hipcc, gfx950)")
            L = ctypes.CDLL(LIB_PATH)
            L.tkfmi_abi_version.restype = ctypes.c_int
            L.tkfmi_last_error.restype = ctypes.c_char_p
            L.tkfmi_fb2d_workspace_bytes.restype = ctypes.c_size_t
            L.tkfmi_fb2d_workspace_bytes.argtypes = [ctypes.POINTER(Problem), ctypes.c_uint32]
            L.tkfmi_fb2d_batched.restype = ctypes.c_int
            L.tkfmi_fb2d_batched.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(Outputs),
                                             _c_p, ctypes.c_size_t, _c_p, ctypes.c_uint32]
            L.tkfmi_fb2d_batched_noalloc.restype = ctypes.c_int
            L.tkfmi_fb2d_batched_noalloc.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(Outputs),
                                                     _c_p, ctypes.c_uint32]
            L.tkfmi_checkpoint_budget.restype = ctypes.c_int
            L.tkfmi_checkpoint_budget.argtypes = [ctypes.c_size_t]
            L.tkfmi_checkpoint_segment.restype = ctypes.c_int
            L.tkfmi_checkpoint_segment.argtypes = [ctypes.c_int]
            L.tkfmi_autotune_workspace_bytes.restype = ctypes.c_size_t
            L.tkfmi_autotune_workspace_bytes.argtypes = [ctypes.POINTER(Problem), ctypes.c_uint32]
            L.tkfmi_autotune.restype = ctypes.c_int
            L.tkfmi_autotune.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(Outputs), _c_p,
                                         ctypes.c_size_t, _c_p, ctypes.c_uint32, ctypes.c_int32,
                                         ctypes.POINTER(TuneResult)]
            L.tkfmi_pairs_decode_workspace_bytes.restype = ctypes.c_size_t
            L.tkfmi_pairs_decode_workspace_bytes.argtypes = [ctypes.POINTER(PairText)]
            L.tkfmi_pairs_decode.restype = ctypes.c_int
            L.tkfmi_pairs_decode.argtypes = [ctypes.POINTER(PairText), ctypes.POINTER(PairColumns),
                                             _c_p, ctypes.c_size_t, _c_p]
            L.tkfmi_fb1d_workspace_bytes.restype = ctypes.c_size_t
            L.tkfmi_fb1d_workspace_bytes.argtypes = [ctypes.POINTER(Problem1D), ctypes.c_uint32]
            L.tkfmi_fb1d_batched.restype = ctypes.c_int
            L.tkfmi_fb1d_batched.argtypes = [ctypes.POINTER(Problem1D), ctypes.POINTER(Outputs),
                                             _c_p, ctypes.c_size_t, _c_p, ctypes.c_uint32]
            L.tkfmi_chain_config.restype = ctypes.c_int
            L.tkfmi_chain_config.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32]
            L.tkfmi_chain_schedule.restype = ctypes.c_int
            L.tkfmi_chain_schedule.argtypes = [ctypes.c_double, ctypes.c_int32]
            L.tkfmi_last_kernel_times.restype = ctypes.c_int
            L.tkfmi_last_kernel_times.argtypes = [ctypes.POINTER(KernelTime), ctypes.c_int32]
            L.tkfmi_suffstats_size.restype = ctypes.c_int32
            L.tkfmi_suffstats_size.argtypes = [ctypes.c_int32, ctypes.c_int32]
            L.tkfmi_restore_workspace_bytes.restype = ctypes.c_size_t
            L.tkfmi_restore_workspace_bytes.argtypes = [ctypes.POINTER(MixdomModel), ctypes.c_int32]
            L.tkfmi_restore_counts_batched.restype = ctypes.c_int
            L.tkfmi_restore_counts_batched.argtypes = [ctypes.POINTER(MixdomModel), ctypes.c_int32,
                                                       _c_p, _c_p, _c_p, _c_p, _c_p,
                                                       ctypes.c_size_t, _c_p]
            L.tkfmi_bdi_stats_batched.restype = ctypes.c_int
            L.tkfmi_bdi_stats_batched.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                                  _c_p, _c_p, _c_p, _c_p, _c_p]
            L.tkfmi_tkf92_estep_stats.restype = ctypes.c_int
            L.tkfmi_tkf92_estep_stats.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_double, _c_p, _c_p, _c_p, _c_p]
            L.tkfmi_hr_stats_batched.restype = ctypes.c_int
            L.tkfmi_hr_stats_batched.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]
            L.tkfmi_build_workspace_bytes.restype = ctypes.c_size_t
            L.tkfmi_build_workspace_bytes.argtypes = [ctypes.POINTER(Builder), ctypes.c_int32]
            L.tkfmi_build_params_batched.restype = ctypes.c_int
            L.tkfmi_build_params_batched.argtypes = [ctypes.POINTER(Builder), ctypes.c_int32, _c_p, _c_p,
                                                     _c_p, _c_p, ctypes.c_size_t, _c_p]
            L.tkfmi_tkf_trans_jac.restype = ctypes.c_int
            L.tkfmi_tkf_trans_jac.argtypes = [_c_p, _c_p, _c_p, _c_p, ctypes.c_int64, ctypes.c_int32,
                                              _c_p, _c_p, _c_p]
            L.tkfmi_chi_workspace_bytes.restype = ctypes.c_size_t
            L.tkfmi_chi_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32]
            chi_args = [_c_p, ctypes.c_int64, ctypes.c_int32, _c_p, _c_p, _c_p, ctypes.c_int32, _c_p, _c_p,
                        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
            L.tkfmi_chi_forward.restype = ctypes.c_int
            L.tkfmi_chi_forward.argtypes = chi_args + [_c_p, _c_p, _c_p, ctypes.c_size_t, _c_p]
            L.tkfmi_chi_backward.restype = ctypes.c_int
            L.tkfmi_chi_backward.argtypes = chi_args + [_c_p, _c_p, _c_p, _c_p, ctypes.c_size_t, _c_p]
            L.tkfmi_expm_batched.restype = ctypes.c_int
            L.tkfmi_expm_batched.argtypes = [_c_p, ctypes.c_int64, ctypes.c_int32, _c_p, _c_p]
            L.tkfmi_expm_vjp_batched.restype = ctypes.c_int
            L.tkfmi_expm_vjp_batched.argtypes = [_c_p, _c_p, ctypes.c_int64, ctypes.c_int32, _c_p, _c_p]
            v = L.tkfmi_abi_version()
            if v != ABI_VERSION:
                raise TkfmiError(f"libtkfmi ABI {v} != expected {ABI_VERSION}")
            _lib = L
        return _lib


def last_kernel_times(max_entries=16):
    """Kernel durations of this thread's last call made with TIME_KERNELS:
    list of dicts (name, n_pairs, cells, ms).  Waits for those kernels."""
    L = lib()
    buf = (KernelTime * max_entries)()
    n = L.tkfmi_last_kernel_times(buf, max_entries)
    if n < 0:
        raise TkfmiError(L.tkfmi_last_error().decode() or f"tkfmi error {n}")
    return [{"name": buf[i].name.decode(), "n_pairs": int(buf[i].n_pairs),
             "cells": int(buf[i].cells), "ms": float(buf[i].ms)} for i in range(n)]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class PackedPairs:
    """A ragged batch of (x, y) code sequences resident on the device."""

    def __init__(self, xs, ys, device):
        if len(xs) != len(ys):
            raise ValueError("xs and ys must have the same length")
        self.B = len(xs)
        xs = [_flat_codes(x) for x in xs]
        ys = [_flat_codes(y) for y in ys]
        self.x_len = np.array([len(x) for x in xs], dtype=np.int32)
        self.y_len = np.array([len(y) for y in ys], dtype=np.int32)
        self.x_off = np.zeros(self.B, dtype=np.int64)
        self.y_off = np.zeros(self.B, dtype=np.int64)
        if self.B:
            self.x_off[1:] = np.cumsum(self.x_len[:-1], dtype=np.int64)
            self.y_off[1:] = np.cumsum(self.y_len[:-1], dtype=np.int64)
        # one normalisation over the concatenation (per-sequence numpy calls
        # cost ~10 us each)
        xcat = normalise_codes(np.concatenate(xs + [np.zeros(1, np.int64)]))
        ycat = normalise_codes(np.concatenate(ys + [np.zeros(1, np.int64)]))
        if torch.device(device).type == "cuda":
            # one stream-ordered copy from pinned memory: a copy from pageable
            # memory would wait for all the GPU work queued before it
            host = torch.empty(len(xcat) + len(ycat), dtype=torch.uint8, pin_memory=True)
            hn = host.numpy()
            hn[:len(xcat)] = xcat
            hn[len(xcat):] = ycat
            both = host.to(device, non_blocking=True)
            self.x_cat, self.y_cat = both[:len(xcat)], both[len(xcat):]
        else:
            self.x_cat = torch.from_numpy(xcat)
            self.y_cat = torch.from_numpy(ycat)
        self.cells = (self.x_len.astype(np.int64) + 1) * (self.y_len.astype(np.int64) + 1)


def _flat_codes(seq):
    return np.asarray(seq.detach().cpu() if isinstance(seq, torch.Tensor) else seq).ravel()


def normalise_codes(seq):
    """Residue codes as uint8 in 0..20.  The reference indexes its padded
    (21,)-tables with jnp gathers, which clamp indices above the table; a
    negative index wraps once (Python semantics)."""
    a = np.asarray(seq.detach().cpu() if isinstance(seq, torch.Tensor) else seq).astype(np.int64).ravel()
    a = np.where(a < 0, a + A_PAD, a)
    return np.clip(a, 0, A_PAD - 1).astype(np.uint8)


def kernel_flags(kernel=None, generic=False):
    """Flag bits that select the 2D forward/backward kernel.

    kernel: None or "chain" (default: chained workgroups per pair -- the
    cell-per-lane path where the state layout allows, the banded MFMA kernel
    for 16..160 emitting states -- else the generic kernel), "class" (one
    workgroup per pair: cell-per-lane, or the antidiagonal MFMA kernel for
    16..160 states) or "generic" (any layout with at most 15 emitting
    states).  ``generic=True`` means kernel="generic".
    """
    if generic or kernel == "generic":
        return GENERIC_KERNEL
    if kernel == "class":
        return CLASS_KERNEL
    if kernel not in (None, "chain"):
        raise ValueError(f"unknown kernel {kernel!r}")
    return 0


def fb2d_batched(log_trans, state_types, state_group, log_tab, pairs, *,
                 forward_only=False, want_posteriors=False, want_counts=True,
                 want_emit_counts=False, generic=False, kernel=None, checkpoint=False,
                 log_emit=None):
    """Run the batched forward/backward kernel once.

    log_trans: float64 device tensor (S,S) or (B,S,S) of log transitions.
    state_types, state_group: int sequences of length S.
    log_tab: float64 device tensor (G, TAB_SZ) or (B, G, TAB_SZ).
    pairs: PackedPairs.
    Returns a dict: 'log_p' (B,), and unless forward_only 'n' (B,S,S)
    (want_counts), 'posteriors' (flat, want_posteriors), 'emit_counts'
    (B,G,TAB_SZ, want_emit_counts).  checkpoint: lattice checkpointing
    (TKFMI_CHECKPOINT: less lattice memory, one extra forward sweep; used by
    the 17..162-state kernels: banded recomputation per band, or
    antidiagonal segments in the one-workgroup-per-pair kernel).  log_emit: flat dense log emissions in the
    posterior layout instead of the tables (see BatchRunner).
    """
    r = BatchRunner(log_trans, state_types, state_group, log_tab, pairs,
                    forward_only=forward_only, want_posteriors=want_posteriors,
                    want_counts=want_counts, want_emit_counts=want_emit_counts,
                    generic=generic, kernel=kernel, checkpoint=checkpoint, log_emit=log_emit)
    r.run(zero_outputs=False)  # outputs are freshly zeroed
    return r.results()


CHECKPOINT_BUDGET_DEFAULT = 256 << 20


def set_checkpoint_budget(nbytes):
    """Lattice bytes of one pair above which the 16..160-state kernel
    checkpoints it automatically (0: only with checkpoint=True)."""
    check(lib().tkfmi_checkpoint_budget(int(nbytes)))


def set_checkpoint_segment(k):
    """Antidiagonals per recomputed segment of the antidiagonal kernel's
    checkpoints (0: the per-pair sqrt rule; otherwise at least 4)."""
    check(lib().tkfmi_checkpoint_segment(int(k)))


def check_status(status):
    """Raise TkfmiError if a call's device status word (tkfmi_outputs.status,
    or the 'status' entry of a results dict) is set: a chained launch gave
    up on a bounded internal wait, and the affected pairs' results are NaN.
    Waits for the call."""
    if isinstance(status, dict):
        status = status["status"]
    if int(status.item()) != 0:
        raise TkfmiError("a chained launch timed out on an internal wait: the affected pairs' "
                         "results are NaN")


def split_posteriors(post_flat, pairs, S):
    """Views (Lx+1, Ly+1, S) of each pair in a flat posterior buffer."""
    out = []
    off = 0
    for b in range(pairs.B):
        c = int(pairs.cells[b]) * S
        out.append(post_flat[off:off + c].view(int(pairs.x_len[b]) + 1, int(pairs.y_len[b]) + 1, S))
        off += c
    return out


class BatchRunner:
    """Reusable launcher: device inputs, workspace and outputs are allocated
    once; :meth:`run` only enqueues the kernel on the current stream (the
    per-call host work is the descriptor sort/upload inside the library)."""

    def __init__(self, log_trans, state_types, state_group, log_tab, pairs, *,
                 forward_only=False, want_posteriors=False, want_counts=True,
                 want_emit_counts=False, generic=False, time_kernels=False, kernel=None,
                 checkpoint=False, log_emit=None):
        """log_emit: optional flat FP64 device tensor of dense log emissions,
        pair b's (Lx+1, Ly+1, S) block at offset sum_{b'<b} (Lx+1)(Ly+1) S
        (the posterior layout); log_tab is then ignored and may be None."""
        self.L = lib()
        dev = log_trans.device
        if dev.type != "cuda":
            raise TkfmiError("tkfmi kernels need a CUDA/HIP device tensor")
        self.dev = dev
        self.pairs = pairs
        self.S = S = int(log_trans.shape[-1])
        self.st = np.ascontiguousarray(np.asarray(state_types, dtype=np.int32))
        self.sg = np.ascontiguousarray(np.asarray(state_group, dtype=np.int32))
        if self.st.shape != (S,) or self.sg.shape != (S,):
            raise ValueError("state_types/state_group must have length S")
        self.log_trans = log_trans.to(torch.float64).contiguous()
        self.log_emit = None
        if log_emit is not None:
            self.log_emit = log_emit.to(torch.float64).contiguous().reshape(-1)
            need = int(pairs.cells.sum()) * S
            if self.log_emit.numel() != need:
                raise ValueError(f"log_emit must hold sum (Lx+1)(Ly+1)S = {need} values, "
                                 f"got {self.log_emit.numel()}")
            if want_emit_counts:
                raise ValueError("want_emit_counts needs table emissions; with log_emit, "
                                 "d logP / d log_emit is the posterior lattice")
            if log_tab is None:  # never read: a placeholder covering the state groups
                log_tab = torch.zeros((int(self.sg.max(initial=0)) + 1, TAB_SZ), dtype=torch.float64,
                                      device=dev)
        self.log_tab = log_tab.to(torch.float64).contiguous()
        if self.log_tab.shape[-1] != TAB_SZ:
            raise ValueError(f"log_tab last dim must be {TAB_SZ}")
        n_trans = 1 if self.log_trans.dim() == 2 else int(self.log_trans.shape[0])
        if self.log_tab.dim() == 2:
            n_tab, G = 1, int(self.log_tab.shape[0])
        else:
            n_tab, G = int(self.log_tab.shape[0]), int(self.log_tab.shape[1])
        self.G = G
        fwd = bool(forward_only)
        self.flags = ((FORWARD_ONLY if fwd else 0)
                      | (WANT_POSTERIORS if (want_posteriors and not fwd) else 0)
                      | (0 if want_counts else NO_COUNTS)
                      | (WANT_EMIT_COUNTS if (want_emit_counts and not fwd) else 0)
                      | kernel_flags(kernel, generic)
                      | (CHECKPOINT if checkpoint else 0)
                      | (TIME_KERNELS if time_kernels else 0))
        B = pairs.B
        self.prob = Problem(B, S, _np_ptr(self.st), _np_ptr(self.sg), G, n_trans,
                            _ptr(self.log_trans), n_tab, _ptr(self.log_tab),
                            _ptr(pairs.x_cat), _ptr(pairs.y_cat), _np_ptr(pairs.x_off),
                            _np_ptr(pairs.y_off), _np_ptr(pairs.x_len), _np_ptr(pairs.y_len),
                            _ptr(self.log_emit))
        self.ws_bytes = int(self.L.tkfmi_fb2d_workspace_bytes(ctypes.byref(self.prob), self.flags))
        if self.ws_bytes == 0 and B > 0:
            # size query failed: run the call anyway to obtain the message
            rc = self.L.tkfmi_fb2d_batched(ctypes.byref(self.prob), ctypes.byref(Outputs()),
                                           None, 0, None, self.flags)
            raise TkfmiError(self.L.tkfmi_last_error().decode() or f"tkfmi error {rc}")
        self.ws = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=dev)
        self.log_p = torch.empty(B, dtype=torch.float64, device=dev)
        self.n = (torch.zeros((B, S, S), dtype=torch.float64, device=dev)
                  if (want_counts and not fwd) else None)
        self.post = (torch.zeros(int(pairs.cells.sum()) * S + 1, dtype=torch.float64, device=dev)
                     if (self.flags & WANT_POSTERIORS) else None)
        self.ecnt = (torch.zeros((B, G, TAB_SZ), dtype=torch.float64, device=dev)
                     if (self.flags & WANT_EMIT_COUNTS) else None)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.out = Outputs(_ptr(self.log_p), _ptr(self.n), _ptr(self.post), _ptr(self.ecnt),
                           _ptr(self.status))

    def run(self, zero_outputs=True, stream=None):
        """Enqueue one call on ``stream`` (default: torch's current stream);
        the output zeroing goes on the same stream."""
        if stream is None:
            stream = torch.cuda.current_stream(self.dev)
        with torch.cuda.stream(stream):
            self.status.zero_()
            if zero_outputs:
                for t in (self.n, self.post, self.ecnt):
                    if t is not None:
                        t.zero_()
        rc = self.L.tkfmi_fb2d_batched(ctypes.byref(self.prob), ctypes.byref(self.out),
                                       _ptr(self.ws), ctypes.c_size_t(self.ws_bytes),
                                       ctypes.c_void_p(stream.cuda_stream), self.flags)
        if rc != 0:
            raise TkfmiError(self.L.tkfmi_last_error().decode() or f"tkfmi error {rc}")
        return self.log_p, self.n, self.post

    def capture(self, zero_outputs=True):
        """Capture one call into a HIP graph and return it (a
        ``torch.cuda.CUDAGraph``; ``graph.replay()`` reruns the call with no
        host work).  The graph holds the output zeroing (torch fill kernels,
        no memset nodes) and the library's launches on its class streams,
        which join the capture through events.  The descriptors are uploaded
        by an ordinary call first; the captured call reuses them
        (REUSE_DESCRIPTORS), so the runner's inputs must not change shape
        while the graph is in use."""
        if self.flags & TIME_KERNELS:
            raise TkfmiError("per-kernel timing cannot be captured")
        self.run(zero_outputs=zero_outputs)
        torch.cuda.synchronize(self.dev)
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(self.dev)
        flags = self.flags | REUSE_DESCRIPTORS
        with torch.cuda.graph(graph, stream=side, capture_error_mode="relaxed"):
            self.status.zero_()
            if zero_outputs:
                for t in (self.n, self.post, self.ecnt):
                    if t is not None:
                        t.zero_()
            rc = self.L.tkfmi_fb2d_batched(ctypes.byref(self.prob), ctypes.byref(self.out),
                                           _ptr(self.ws), ctypes.c_size_t(self.ws_bytes),
                                           ctypes.c_void_p(side.cuda_stream), flags)
        if rc != 0:
            raise TkfmiError(self.L.tkfmi_last_error().decode() or f"tkfmi error {rc}")
        return graph

    def resize_workspace(self):
        """Query the workspace again (after tkfmi_chain_config or
        tkfmi_autotune changed the configuration of this batch's key) and
        grow it if needed."""
        need = int(self.L.tkfmi_fb2d_workspace_bytes(ctypes.byref(self.prob), self.flags))
        if need > self.ws.numel():
            self.ws = torch.empty(need, dtype=torch.uint8, device=self.dev)
        self.ws_bytes = max(need, 1)

    def check_status(self):
        """Raise if the last call's chained launches gave up on an internal
        wait (their pairs' log_prob is then NaN, which must not be mistaken
        for an impossible pair).  Waits for the call."""
        check_status(self.status)

    def results(self):
        out = {"log_p": self.log_p, "status": self.status}
        if self.n is not None:
            out["n"] = self.n
        if self.post is not None:
            out["posteriors"] = self.post
        if self.ecnt is not None:
            out["emit_counts"] = self.ecnt
        return out


def check(rc):
    """Raise TkfmiError with the library's message for a nonzero return."""
    if rc != 0:
        raise TkfmiError(lib().tkfmi_last_error().decode() or f"tkfmi error {rc}")


def current_stream_ptr(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class PackedColumns:
    """A ragged batch of alignments resident on the device: per column the
    type (1=M, 2=I, 3=D) and the ancestor / descendant codes (0..20; the
    code of the side a column does not consume is ignored)."""

    def __init__(self, col_types, ancs, descs, device):
        if not (len(col_types) == len(ancs) == len(descs)):
            raise ValueError("col_types, ancs and descs must have the same length")
        self.B = len(col_types)
        cts = [np.asarray(c.detach().cpu() if isinstance(c, torch.Tensor) else c).astype(np.int64).ravel()
               for c in col_types]
        for c in cts:
            if c.size and (c.min() < 1 or c.max() > 3):
                raise ValueError("column types must be 1 (M), 2 (I) or 3 (D)")
        an = [normalise_codes(a) for a in ancs]
        de = [normalise_codes(d) for d in descs]
        self.len = np.array([len(c) for c in cts], dtype=np.int32)
        for c, a, d in zip(cts, an, de):
            if len(a) != len(c) or len(d) != len(c):
                raise ValueError("anc / desc must have one code per column")
        self.off = np.zeros(self.B, dtype=np.int64)
        if self.B:
            self.off[1:] = np.cumsum(self.len[:-1], dtype=np.int64)
        cat = lambda arrs: torch.from_numpy(
            np.concatenate([x.astype(np.uint8) for x in arrs] + [np.zeros(1, np.uint8)])).to(device)
        self.col = cat(cts)
        self.anc = cat(an)
        self.desc = cat(de)
        self.rows = int(self.len.sum())


class Runner1D:
    """Launcher of the batched 1D forward/backward (tkfmi_fb1d_batched).

    log_trans: float64 device tensor (S,S) or (B,S,S).  Emissions either
    dense (``log_emit``: float64 device tensor (rows, S) with rows =
    sum of lengths, pair b's columns consecutive from ``offsets[b]``) or
    constrained (``columns``: PackedColumns, with ``state_types``,
    ``state_group`` and ``log_tab`` (G, TAB_SZ) or (B, G, TAB_SZ))."""

    def __init__(self, log_trans, *, init_state, final_state, lengths=None, offsets=None,
                 log_emit=None, columns=None, state_types=None, state_group=None, log_tab=None,
                 forward_only=False, want_posteriors=False, want_counts=True,
                 want_emit_counts=False, time_kernels=False):
        self.L = lib()
        dev = log_trans.device
        if dev.type != "cuda":
            raise TkfmiError("tkfmi kernels need a CUDA/HIP device tensor")
        self.dev = dev
        self.log_trans = log_trans.to(torch.float64).contiguous()
        S = self.S = int(self.log_trans.shape[-1])
        n_trans = 1 if self.log_trans.dim() == 2 else int(self.log_trans.shape[0])
        self.keep = []
        if (log_emit is None) == (columns is None):
            raise ValueError("give exactly one of log_emit and columns")
        if columns is not None:
            B = columns.B
            self.h_len, self.h_off = columns.len, columns.off
            st = np.ascontiguousarray(np.asarray(state_types, dtype=np.int32))
            sg = np.ascontiguousarray(np.asarray(state_group, dtype=np.int32))
            if st.shape != (S,) or sg.shape != (S,):
                raise ValueError("state_types/state_group must have length S")
            self.log_tab = log_tab.to(torch.float64).contiguous()
            if self.log_tab.shape[-1] != TAB_SZ:
                raise ValueError(f"log_tab last dim must be {TAB_SZ}")
            if self.log_tab.dim() == 2:
                n_tab, G = 1, int(self.log_tab.shape[0])
            else:
                n_tab, G = int(self.log_tab.shape[0]), int(self.log_tab.shape[1])
            self.G = G
            self.keep += [st, sg, columns]
            prob_args = dict(log_emit=None, state_types=_np_ptr(st), state_group=_np_ptr(sg),
                             n_groups=G, n_tab=n_tab, log_tab=_ptr(self.log_tab),
                             col_type=_ptr(columns.col), anc=_ptr(columns.anc), desc=_ptr(columns.desc))
        else:
            self.h_len = np.ascontiguousarray(np.asarray(lengths, dtype=np.int32))
            B = len(self.h_len)
            if offsets is None:
                offsets = np.zeros(B, dtype=np.int64)
                if B:
                    offsets[1:] = np.cumsum(self.h_len[:-1], dtype=np.int64)
            self.h_off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
            self.log_emit = log_emit.to(torch.float64).contiguous()
            if self.log_emit.dim() != 2 or self.log_emit.shape[1] != S:
                raise ValueError("log_emit must be (rows, S)")
            if B and int((self.h_off + self.h_len).max()) > self.log_emit.shape[0]:
                raise ValueError("log_emit has fewer rows than the lengths / offsets need")
            self.G = 0
            prob_args = dict(log_emit=_ptr(self.log_emit), state_types=None, state_group=None,
                             n_groups=0, n_tab=0, log_tab=None, col_type=None, anc=None, desc=None)
        self.B = B
        fwd = bool(forward_only)
        self.flags = ((FORWARD_ONLY if fwd else 0)
                      | (WANT_POSTERIORS if (want_posteriors and not fwd) else 0)
                      | (0 if want_counts else NO_COUNTS)
                      | (WANT_EMIT_COUNTS if (want_emit_counts and not fwd and columns is not None) else 0)
                      | (TIME_KERNELS if time_kernels else 0))
        self.prob = Problem1D(B, S, int(init_state), int(final_state), n_trans, 0,
                              _ptr(self.log_trans), prob_args["log_emit"], prob_args["state_types"],
                              prob_args["state_group"], prob_args["n_groups"], prob_args["n_tab"],
                              prob_args["log_tab"], prob_args["col_type"], prob_args["anc"],
                              prob_args["desc"], _np_ptr(self.h_off), _np_ptr(self.h_len))
        self.ws_bytes = int(self.L.tkfmi_fb1d_workspace_bytes(ctypes.byref(self.prob), self.flags))
        if self.ws_bytes == 0:
            raise TkfmiError(self.L.tkfmi_last_error().decode() or "invalid 1D problem")
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
        rows = int(self.h_len.sum())
        self.rows = rows
        self.log_p = torch.empty(B, dtype=torch.float64, device=dev)
        self.n = (torch.zeros((B, S, S), dtype=torch.float64, device=dev)
                  if (want_counts and not fwd) else None)
        self.post = (torch.zeros(rows * S + 1, dtype=torch.float64, device=dev)
                     if (self.flags & WANT_POSTERIORS) else None)
        self.ecnt = (torch.zeros((B, self.G, TAB_SZ), dtype=torch.float64, device=dev)
                     if (self.flags & WANT_EMIT_COUNTS) else None)
        self.out = Outputs(_ptr(self.log_p), _ptr(self.n), _ptr(self.post), _ptr(self.ecnt))

    def run(self, zero_outputs=True, stream=None):
        if zero_outputs:
            for t in (self.n, self.post, self.ecnt):
                if t is not None:
                    t.zero_()
        if stream is None:
            stream = torch.cuda.current_stream(self.dev)
        rc = self.L.tkfmi_fb1d_batched(ctypes.byref(self.prob), ctypes.byref(self.out),
                                       _ptr(self.ws), ctypes.c_size_t(self.ws_bytes),
                                       ctypes.c_void_p(stream.cuda_stream), self.flags)
        if rc != 0:
            raise TkfmiError(self.L.tkfmi_last_error().decode() or f"tkfmi error {rc}")
        return self.log_p

    def posteriors(self):
        """Per-pair (L_b, S) views of the posterior buffer (pair order)."""
        out, off = [], 0
        for b in range(self.B):
            n = int(self.h_len[b]) * self.S
            out.append(self.post[off:off + n].view(int(self.h_len[b]), self.S))
            off += n
        return out
