"""Wide state spaces (16..160 emitting states; the banded MFMA kernel, the
default, and the one-workgroup-per-pair antidiagonal kernel, kernel="class",
which also runs every checkpointed call): MixDom2 layouts up to K=8 F=4 C=4
(S = 162) against the FP64 oracle (log-likelihood, transition counts,
posteriors, emission counts: 1e-10 relative), and the exact count
identities at lengths up to 1000."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def _close(a, b, rtol=RTOL, floor=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b)
    return bool(np.all(err <= rtol * np.abs(b) + floor * scale)), float((err / (np.abs(b) + floor * scale)).max())


def _oracle(w, b, x, y):
    from oracle import dp2d, emissions
    p = w.params
    em = emissions.pair_hmm_emissions_per_class(np.asarray(w.state_types), x, y, p["class_subs"][b],
                                                p["class_pi"], p["class_dist"], p["K"], p["F"])
    return dp2d.forward_backward(w.log_trans[b].numpy(), np.asarray(w.state_types), em)


@pytest.mark.parametrize("kernel", [None, "class"])
@pytest.mark.parametrize("K,F,C", [(2, 2, 2), (4, 2, 1), (5, 4, 2), (8, 2, 2), (8, 4, 4)])
def test_wide_matches_oracle(K, F, C, kernel):
    from oracle import emissions
    from tkfmi import hmm, synth
    lengths = [0, 1, 5, 17, 30, 40]
    w = synth.make_mixdom2(seed=50 + K + F + C, B=len(lengths), K=K, F=F, C=C, lengths=lengths)
    rng = np.random.default_rng(K * F * C)
    w.ys[0] = w.ys[0][:0]
    w.ys[2] = rng.integers(0, 20, 33)
    res = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                         want_posteriors=True, want_emit_counts=True, kernel=kernel)
    G = res["emit_counts"].shape[1]
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        print(f"K{K} F{F} C{C} kernel={kernel} pair {b} Lx={len(x)} Ly={len(y)}", flush=True)
        lp_o, post_o, n_o, _ = _oracle(w, b, x, y)
        assert abs(float(res["log_p"][b]) - lp_o) <= RTOL * abs(lp_o), (b, float(res["log_p"][b]), lp_o)
        ok, worst = _close(res["n"][b].cpu().numpy(), n_o)
        assert ok, ("n", b, worst)
        ok, worst = _close(res["posteriors"][b].cpu().numpy(), post_o)
        assert ok, ("post", b, worst)
        ec_o = emissions.emission_counts(post_o, w.state_types, w.state_group, x, y, G)
        ok, worst = _close(res["emit_counts"][b].cpu().numpy(), ec_o)
        assert ok, ("ecnt", b, worst)


def test_wide_forward_only_and_batched_equals_alone():
    from tkfmi import hmm, synth
    w = synth.make_mixdom2(seed=9, B=5, K=8, F=4, C=4, lengths=[60, 3, 90, 45, 120])
    full = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group)
    fo = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                        forward_only=True)
    assert np.array_equal(full["log_p"].cpu().numpy(), fo["log_p"].cpu().numpy())
    for b in (0, 2, 4):
        alone = hmm.fb2d_batch(w.log_trans[b], w.state_types, [w.xs[b]], [w.ys[b]], w.log_tab[b],
                               w.state_group)
        assert float(alone["log_p"][0]) == float(full["log_p"][b])
        assert np.array_equal(alone["n"][0].cpu().numpy(), full["n"][b].cpu().numpy())


def test_wide_count_identities_long_pairs():
    """K=8 F=4 C=4 (S=162) at lengths up to 1000: sum of counts into M/D
    states == Lx, into M/I == Ly, start row == 1, flow conservation."""
    from tkfmi import hmm, synth
    w = synth.make_mixdom2(seed=2, B=3, K=8, F=4, C=4, lengths=[1000, 400, 730])
    rng = np.random.default_rng(2)
    w.ys[0] = rng.integers(0, 20, 950)
    res = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group)
    types = np.asarray(w.state_types)
    n = res["n"].cpu().numpy()
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        nb = n[b]
        assert np.isfinite(float(res["log_p"][b]))
        assert abs(nb[:, np.isin(types, (1, 3))].sum() - len(x)) <= 1e-9 * len(x)
        assert abs(nb[:, np.isin(types, (1, 2))].sum() - len(y)) <= 1e-9 * len(y)
        assert abs(nb[0].sum() - 1.0) < 1e-11
        body = np.isin(types, (1, 2, 3))
        np.testing.assert_allclose(nb[:, body].sum(0), nb[body, :].sum(1), rtol=1e-9, atol=1e-9)


def test_wide_checkpoint_matches_full_lattice():
    """TKFMI_CHECKPOINT: the backward recomputes each K-antidiagonal segment
    from its checkpoint with the forward's own arithmetic, so log P, counts
    and posteriors equal the antidiagonal kernel's full-lattice run bit for bit (emission counts to
    the FP64 atomics' summation order); a small layout also against the
    oracle."""
    from oracle import emissions
    from tkfmi import hmm, synth
    w = synth.make_mixdom2(seed=21, B=5, K=8, F=4, C=4, lengths=[60, 3, 90, 0, 150])
    kw = dict(want_posteriors=True, want_emit_counts=True)
    full = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group, kernel="class", **kw)
    ck = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group, checkpoint=True, **kw)
    assert np.array_equal(full["log_p"].cpu().numpy(), ck["log_p"].cpu().numpy())
    assert np.array_equal(full["n"].cpu().numpy(), ck["n"].cpu().numpy())
    for b in range(5):
        print(f"K8 F4 C4 pair {b} Lx={len(w.xs[b])} Ly={len(w.ys[b])}", flush=True)
        assert np.array_equal(full["posteriors"][b].cpu().numpy(), ck["posteriors"][b].cpu().numpy()), b
    ok, worst = _close(ck["emit_counts"].cpu().numpy(), full["emit_counts"].cpu().numpy(), rtol=1e-12)
    assert ok, worst
    w = synth.make_mixdom2(seed=77, B=3, K=2, F=2, C=2, lengths=[0, 17, 40])
    res = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group, checkpoint=True, **kw)
    G = res["emit_counts"].shape[1]
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        lp_o, post_o, n_o, _ = _oracle(w, b, x, y)
        assert abs(float(res["log_p"][b]) - lp_o) <= RTOL * abs(lp_o), b
        assert _close(res["n"][b].cpu().numpy(), n_o)[0], b
        assert _close(res["posteriors"][b].cpu().numpy(), post_o)[0], b
        ec_o = emissions.emission_counts(post_o, w.state_types, w.state_group, x, y, G)
        assert _close(res["emit_counts"][b].cpu().numpy(), ec_o)[0], b


def test_wide_checkpoint_segment_length_does_not_change_results():
    """tkfmi_checkpoint_segment: with K fixed at 4 (the smallest), 7 and 50
    antidiagonals per segment instead of the per-pair sqrt rule, the
    checkpointed antidiagonal kernel still equals its full-lattice run bit
    for bit (emission counts to the FP64 atomics' summation order)."""
    from tkfmi import _abi, hmm, synth
    w = synth.make_mixdom2(seed=23, B=4, K=8, F=4, C=4, lengths=[70, 5, 120, 33])
    kw = dict(want_posteriors=True, want_emit_counts=True, kernel="class")
    full = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group, **kw)
    try:
        for k in (4, 7, 50):
            print(f"checkpoint segment K={k}", flush=True)
            _abi.set_checkpoint_segment(k)
            ck = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                                checkpoint=True, **kw)
            assert np.array_equal(full["log_p"].cpu().numpy(), ck["log_p"].cpu().numpy()), k
            assert np.array_equal(full["n"].cpu().numpy(), ck["n"].cpu().numpy()), k
            for b in range(4):
                assert np.array_equal(full["posteriors"][b].cpu().numpy(),
                                      ck["posteriors"][b].cpu().numpy()), (k, b)
            ok, worst = _close(ck["emit_counts"].cpu().numpy(), full["emit_counts"].cpu().numpy(), rtol=1e-12)
            assert ok, (k, worst)
    finally:
        _abi.set_checkpoint_segment(0)


def test_budget_checkpointing_is_bit_identical():
    """Pairs over the checkpoint budget are checkpointed automatically (here
    a 2 MB budget catches the three longer pairs of the batch but not the
    shortest); every result equals the run without checkpointing bit for
    bit (the recomputation repeats the forward's arithmetic in its order)."""
    from tkfmi import _abi, hmm, synth
    w = synth.make_mixdom2(seed=44, B=4, K=4, F=2, C=2, lengths=[12, 90, 150, 60])
    run = lambda: hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                                 want_posteriors=True, want_emit_counts=True, kernel="class")
    try:
        _abi.set_checkpoint_budget(0)
        full = run()
        _abi.set_checkpoint_budget(2 << 20)
        auto = run()
    finally:
        _abi.set_checkpoint_budget(_abi.CHECKPOINT_BUDGET_DEFAULT)
    assert np.array_equal(full["log_p"].cpu().numpy(), auto["log_p"].cpu().numpy())
    assert np.array_equal(full["n"].cpu().numpy(), auto["n"].cpu().numpy())
    for a, b in zip(full["posteriors"], auto["posteriors"]):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
