"""ctypes wrapper of the C oracle restatement (oracle/c/fb2d_oracle.c).

Test infrastructure only (checker and timed CPU baseline).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "build", "liboracle_fb2d.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(LIB)
        L.oracle_fb2d_batch.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else ctypes.c_void_p(0)


def fb2d_batch(log_trans, state_types, state_group, log_tab, xs, ys, forward_only=False,
               want_posteriors=False, nthreads=1):
    """log_tab: (G, 483) log tables (21x21 match, 21 insert, 21 delete).
    Returns (logP (B,), n (B,S,S) or None, posteriors list or None)."""
    lt = np.ascontiguousarray(log_trans, dtype=np.float64)
    S = lt.shape[-1]
    n_trans = 1 if lt.ndim == 2 else lt.shape[0]
    tab = np.ascontiguousarray(log_tab, dtype=np.float64).reshape(-1, 483)
    G = tab.shape[0]
    tabM = np.ascontiguousarray(tab[:, :441])
    tabI = np.ascontiguousarray(tab[:, 441:462])
    tabD = np.ascontiguousarray(tab[:, 462:483])
    st = np.ascontiguousarray(state_types, dtype=np.int32)
    sg = np.ascontiguousarray(state_group, dtype=np.int32)
    B = len(xs)
    xl = np.array([len(x) for x in xs], dtype=np.int32)
    yl = np.array([len(y) for y in ys], dtype=np.int32)
    xo = np.zeros(B, np.int64)
    yo = np.zeros(B, np.int64)
    if B:
        xo[1:] = np.cumsum(xl[:-1])
        yo[1:] = np.cumsum(yl[:-1])
    xc = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint8) for x in xs] + [np.zeros(1, np.uint8)]))
    yc = np.ascontiguousarray(np.concatenate([np.asarray(y, np.uint8) for y in ys] + [np.zeros(1, np.uint8)]))
    logp = np.zeros(B)
    n = None if forward_only else np.zeros((B, S, S))
    cells = (xl.astype(np.int64) + 1) * (yl.astype(np.int64) + 1) * S
    post = np.zeros(int(cells.sum()) + 1) if (want_posteriors and not forward_only) else None
    poff = np.zeros(B, np.int64)
    if B:
        poff[1:] = np.cumsum(cells[:-1])
    rc = lib().oracle_fb2d_batch(
        ctypes.c_int(B), ctypes.c_int(S), _p(st), _p(sg), _p(lt), ctypes.c_int(n_trans),
        _p(tabM), _p(tabI), _p(tabD), _p(xc), _p(xo), _p(xl), _p(yc), _p(yo), _p(yl),
        ctypes.c_int(1 if forward_only else 0), _p(logp), _p(n), _p(post), _p(poff),
        ctypes.c_int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_fb2d_batch failed: {rc}")
    posts = None
    if post is not None:
        posts = [post[poff[b]:poff[b] + cells[b]].reshape(xl[b] + 1, yl[b] + 1, S) for b in range(B)]
    return logp, n, posts


_LIBS_SCALED = {}


def _lib_scaled(long_double):
    name = "liboracle_fb2d_ld.so" if long_double else "liboracle_fb2d_sc.so"
    if name not in _LIBS_SCALED:
        path = os.path.join(_HERE, "build", name)
        if not os.path.exists(path):
            subprocess.run(["make", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(path)
        fn = L.oracle_fb2d_scaled_ld if long_double else L.oracle_fb2d_scaled
        fn.restype = ctypes.c_int
        _LIBS_SCALED[name] = fn
    return _LIBS_SCALED[name]


def fb2d_scaled(log_trans, state_types, state_group, log_tab, x, y, samples=None,
                long_double=True, nthreads=8):
    """One pair through oracle/c/fb2d_scaled.c (max-shifted product form of
    the log-space recursion, in 80-bit long double by default): returns
    (logP, n (S,S), emission counts (G,483), posteriors at the sampled cells
    (n_samp,S)).  samples: (n_samp, 2) int array of (i, j) cells.  The FP64
    build needs |logP| < ~700 (the count scale factor is exp(-logP))."""
    lt = np.ascontiguousarray(log_trans, dtype=np.float64)
    S = lt.shape[-1]
    tab = np.ascontiguousarray(log_tab, dtype=np.float64).reshape(-1, 483)
    G = tab.shape[0]
    st = np.ascontiguousarray(state_types, dtype=np.int32)
    sg = np.ascontiguousarray(state_group, dtype=np.int32)
    xc = np.ascontiguousarray(np.concatenate([np.minimum(np.asarray(x, np.int64), 20), [0]]).astype(np.uint8))
    yc = np.ascontiguousarray(np.concatenate([np.minimum(np.asarray(y, np.int64), 20), [0]]).astype(np.uint8))
    samples = np.zeros((0, 2), np.int32) if samples is None else np.asarray(samples, np.int32)
    si = np.ascontiguousarray(samples[:, 0])
    sj = np.ascontiguousarray(samples[:, 1])
    logp = np.zeros(1)
    n = np.zeros((S, S))
    ec = np.zeros((G, 483))
    post = np.zeros((max(len(samples), 1), S))
    rc = _lib_scaled(long_double)(
        ctypes.c_int(S), _p(st), _p(sg), ctypes.c_int(G), _p(lt), _p(tab), _p(xc),
        ctypes.c_int(len(x)), _p(yc), ctypes.c_int(len(y)), ctypes.c_int(nthreads), _p(logp),
        _p(n), _p(ec), ctypes.c_int(len(samples)), _p(si), _p(sj), _p(post))
    if rc != 0:
        raise RuntimeError(f"oracle_fb2d_scaled failed: {rc}")
    return float(logp[0]), n, ec, post[:len(samples)]
