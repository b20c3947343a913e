"""tkfmi_fb2d_batched_noalloc (no workspace, no allocation) against
tkfmi_fb2d_batched with its own workspace, through the C ABI: two calls in a
row on the shared device scratch, the second with a smaller batch, on the
chained 17-state path (with emission counts and posteriors), the TKF path,
and the 16..160-state kernel (82 states).  log P must be bit-identical (the
same kernels run in the same order per pair); counts, posteriors and
emission counts agree to summation order (1e-13).  A batch larger than the
scratch is refused with a message, before any device work."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _runner(w, want_post=True, want_ecnt=True):
    from tkfmi import _abi
    dev = torch.device("cuda", 0)
    pairs = _abi.PackedPairs(list(w.xs), list(w.ys), dev)
    return _abi.BatchRunner(w.log_trans.to(dev), w.state_types, w.state_group, w.log_tab.to(dev),
                            pairs, want_posteriors=want_post, want_emit_counts=want_ecnt)


def _noalloc(r):
    """The same problem as runner r, through the allocation-free entry point,
    into fresh outputs."""
    from tkfmi import _abi
    L = _abi.lib()
    dev = r.log_p.device
    lp = torch.empty_like(r.log_p)
    n = torch.zeros_like(r.n)
    post = torch.zeros_like(r.post) if r.post is not None else None
    ec = torch.zeros_like(r.ecnt) if r.ecnt is not None else None
    st = torch.zeros(1, dtype=torch.int32, device=dev)
    out = _abi.Outputs(_abi._ptr(lp), _abi._ptr(n), _abi._ptr(post), _abi._ptr(ec), _abi._ptr(st))
    s = torch.cuda.current_stream(dev)
    rc = L.tkfmi_fb2d_batched_noalloc(ctypes.byref(r.prob), ctypes.byref(out),
                                      ctypes.c_void_p(s.cuda_stream), r.flags)
    assert rc == 0, L.tkfmi_last_error().decode()
    torch.cuda.synchronize()
    assert int(st.item()) == 0
    return lp, n, post, ec


def _same(ref, got, what):
    ref, got = ref.cpu().numpy(), got.cpu().numpy()
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref) / (np.abs(ref) + 1e-12 * scale)
    assert err.max() <= 1e-13, (what, float(err.max()))


def _check(r):
    r.run()
    torch.cuda.synchronize()
    r.check_status()
    lp, n, post, ec = _noalloc(r)
    assert torch.equal(lp, r.log_p), "log P differs"
    _same(r.n, n, "n")
    if post is not None:
        _same(r.post, post, "posteriors")
    if ec is not None:
        _same(r.ecnt, ec, "emission counts")


@pytest.mark.parametrize("K,F,B,hi", [(3, 1, 40, 330), (8, 2, 16, 200)])
def test_noalloc_matches_workspace_entry_point(K, F, B, hi):
    from tkfmi import synth
    big = synth.make_mixdom2(seed=81, B=B, K=K, F=F, lengths=list(np.random.default_rng(2).integers(20, hi, B)))
    small = synth.make_mixdom2(seed=82, B=7, K=K, F=F, lengths=[5, 60, 0, 140, 33, 9, 210])
    _check(_runner(big))
    _check(_runner(small))


def test_noalloc_tkf_then_smaller():
    from tkfmi import synth
    big = synth.make_tkf(seed=83, B=30, lengths=list(np.random.default_rng(3).integers(10, 400, 30)))
    small = synth.make_tkf(seed=84, B=3, lengths=[1, 40, 8])
    _check(_runner(big, want_ecnt=False))
    _check(_runner(small, want_ecnt=False))


def test_noalloc_refuses_oversized_batch():
    from tkfmi import _abi, synth
    L = _abi.lib()
    w = synth.make_mixdom2(seed=85, B=1, lengths=[2200])
    w.ys[0] = np.resize(w.ys[0], 2300)
    r = _runner(w, want_post=False, want_ecnt=False)
    assert r.ws_bytes > _abi.NOALLOC_SCRATCH_BYTES
    out = _abi.Outputs(_abi._ptr(r.log_p), _abi._ptr(r.n), None, None, None)
    rc = L.tkfmi_fb2d_batched_noalloc(ctypes.byref(r.prob), ctypes.byref(out), None, r.flags)
    assert rc == -1 and "scratch" in L.tkfmi_last_error().decode()
    rc = L.tkfmi_fb2d_batched_noalloc(ctypes.byref(r.prob), ctypes.byref(out), None,
                                      r.flags | _abi.TIME_KERNELS)
    assert rc == -1
