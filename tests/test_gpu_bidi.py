"""GPU parity of the chained kernels' one-launch schedule (a pair's forward
and backward meet at antidiagonal C = round(cut (Lx + Ly))): log P,
expected transition counts and expected emission counts against the FP64
oracle at 1e-10 relative, for meeting points from the origin (C = 0) to the
corner (C = Lx + Ly), and against the two-launch schedule (cut = 0).

Tolerance: 1e-10 relative; entries tiny next to the largest of their kind
get an absolute floor of 1e-12 x that maximum.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-10
CUTS = [pytest.param(0.45, id="cut0.45"), pytest.param(1e-9, id="origin"), pytest.param(0.2, id="cut0.2"),
        pytest.param(0.8, id="cut0.8"), pytest.param(1.0, id="corner")]


def _close(a, b, rtol=RTOL, floor=1e-12):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b)
    ok = err <= rtol * np.abs(b) + floor * scale
    return bool(ok.all()), float((err / (np.abs(b) + floor * scale)).max())


class _cut:
    """Forces the one-launch schedule (whatever the call's size) with the
    given meeting fraction for the duration of a block (0: two launches)."""

    def __init__(self, frac):
        self.frac = frac

    def __enter__(self):
        from tkfmi import autotune
        autotune.set_chain_schedule(self.frac, 1 << 30)

    def __exit__(self, *exc):
        from tkfmi import autotune
        autotune.set_chain_schedule(*autotune.DEFAULT_SCHEDULE)


def _run(w, cut, ec=True, kernel="chain"):
    from tkfmi import hmm
    with _cut(cut):
        return hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                              want_emit_counts=ec, kernel=kernel)


def _check_vs_oracle(w, res, emit, lt_of):
    from oracle import dp2d, emissions
    G = res["emit_counts"].shape[1]
    ec = res["emit_counts"].cpu().numpy()
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        print(f"pair {b} Lx={len(x)} Ly={len(y)}", flush=True)
        lp_o, post_o, n_o, _ = dp2d.forward_backward(lt_of(b), np.asarray(w.state_types), emit(b, x, y))
        lp = float(res["log_p"][b])
        assert abs(lp - lp_o) <= RTOL * abs(lp_o), (b, lp, lp_o)
        ok, worst = _close(res["n"][b].cpu().numpy(), n_o)
        assert ok, (b, "n", worst)
        want = emissions.emission_counts(post_o, w.state_types, w.state_group, x, y, G)
        ok, worst = _close(ec[b], want)
        assert ok, (b, "emit", worst)


def _tkf(lengths_x, lengths_y, seed, ext):
    from tkfmi import synth
    from oracle import emissions
    rng = np.random.default_rng(seed)
    w = synth.make_tkf(seed=seed, B=1, ext=ext)
    w.xs = [rng.integers(0, 20, size=n) for n in lengths_x]
    w.ys = [rng.integers(0, 20, size=n) for n in lengths_y]
    w.ts = np.full(len(w.xs), w.ts[0])
    emit = lambda b, x, y: emissions.pair_hmm_emissions(np.asarray(w.state_types), x, y,
                                                        w.params["sub"], w.params["pi"])
    return w, emit, (lambda b: w.log_trans.numpy())


def _mixdom(K, F, C, lengths, seed, ys=None):
    from tkfmi import synth
    from oracle import emissions
    w = synth.make_mixdom2(seed=seed, B=len(lengths), K=K, F=F, C=C, lengths=lengths)
    for b, n in (ys or {}).items():
        w.ys[b] = np.random.default_rng(seed + b).integers(0, 20, size=n)
    p = w.params
    emit = lambda b, x, y: emissions.pair_hmm_emissions_per_class(np.asarray(w.state_types), x, y,
                                                                 p["class_subs"][b], p["class_pi"],
                                                                 p["class_dist"], p["K"], p["F"])
    return w, emit, (lambda b: w.log_trans[b].numpy())


@pytest.mark.parametrize("cut", CUTS)
def test_tkf_edge_and_tall_pairs(cut):
    """Empty, one-cell, band-edge (63/64/65 rows), tall (one column over many
    bands) and flat pairs: every meeting point gives the oracle's results."""
    xs = [0, 0, 3, 1, 2, 37, 64, 65, 70, 5, 130, 300, 1000, 640, 2, 257]
    ys = [0, 4, 0, 1, 2, 64, 37, 3, 1, 140, 96, 520, 0, 3, 700, 2]
    w, emit, lt_of = _tkf(xs, ys, seed=3, ext=0.6)
    _check_vs_oracle(w, _run(w, cut), emit, lt_of)


@pytest.mark.parametrize("cut", CUTS)
def test_mixdom2_k3(cut):
    """MixDom2 K=3 F=1 C=2 (17 states), 1 to 8 bands per pair (chains of up
    to 3 workgroups of 3 waves each way)."""
    w, emit, lt_of = _mixdom(3, 1, 2, [0, 1, 31, 63, 64, 65, 130, 200, 140, 450], seed=17,
                             ys={0: 0, 1: 3, 7: 560, 8: 300})
    _check_vs_oracle(w, _run(w, cut), emit, lt_of)


@pytest.mark.parametrize("K,F,C", [(1, 3, 2), (2, 1, 3), (1, 2, 1), (1, 1, 2)])
def test_mixdom_layouts(K, F, C):
    w, emit, lt_of = _mixdom(K, F, C, [0, 5, 64, 65, 200, 330], seed=40 + K + F + C, ys={0: 7})
    _check_vs_oracle(w, _run(w, 0.45), emit, lt_of)


@pytest.mark.parametrize("cfg", [(4, 3, 0.3), (2, 2, 0.0), (3, 2, 0.3)])
def test_workgroup_widths(cfg):
    """Other workgroup widths per length class (1- to 4-wave chains), every
    class on the one-launch schedule."""
    from tkfmi import autotune
    w, emit, lt_of = _mixdom(3, 1, 1, [5, 64, 65, 130, 260, 390, 520, 700], seed=23)
    try:
        autotune.set_chain_config(*cfg)
        res = _run(w, 0.45)
    finally:
        autotune.set_chain_config(*autotune.DEFAULT)
    _check_vs_oracle(w, res, emit, lt_of)


def test_one_launch_equals_two_launches():
    """The benchmark workload's first 128 pairs: one launch (default meeting
    point, forced) == two launches; log P is the forward's own sum in both."""
    from tkfmi import synth
    w = synth.make_mixdom2(seed=1, B=128)
    a = _run(w, 0.45)
    b = _run(w, 0.0)
    la, lb = a["log_p"].cpu().numpy(), b["log_p"].cpu().numpy()
    assert np.array_equal(la, lb)
    ok, worst = _close(a["n"].cpu().numpy(), b["n"].cpu().numpy(), rtol=1e-12)
    assert ok, worst
    ok, worst = _close(a["emit_counts"].cpu().numpy(), b["emit_counts"].cpu().numpy(), rtol=1e-12)
    assert ok, worst


def test_counts_only_and_emissions_only():
    """Transition counts without emission counts, and emission counts
    without transition counts (TKFMI_NO_COUNTS), on the one-launch schedule."""
    from tkfmi import hmm
    w, emit, lt_of = _mixdom(3, 1, 1, [40, 200, 350], seed=29)
    full = _run(w, 0.45)
    a = _run(w, 0.45, ec=False)
    ok, worst = _close(a["n"].cpu().numpy(), full["n"].cpu().numpy(), rtol=1e-14)
    assert ok, worst
    with _cut(0.45):
        b = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                           want_counts=False, want_emit_counts=True, kernel="chain")
    ok, worst = _close(b["emit_counts"].cpu().numpy(), full["emit_counts"].cpu().numpy(), rtol=1e-14)
    assert ok, worst


def test_wait_timeout_marks_pairs_nan(monkeypatch):
    """Waits that time out (a spin limit of 1, forced from the host) end the
    launch: every pair is exact or NaN, never a wrong number, the status word
    is raised, and the next call with the normal limit is exact."""
    from tkfmi import _abi, synth
    w = synth.make_mixdom2(seed=31, B=8, lengths=[40, 700, 20, 650, 60, 800, 30, 500])
    ref = _run(w, 0.0, ec=False, kernel="generic")
    monkeypatch.setenv("TKFMI_CHAIN_SPIN_LIMIT", "1")
    bad = _run(w, 0.45, ec=False)
    monkeypatch.delenv("TKFMI_CHAIN_SPIN_LIMIT")
    lr, lb = ref["log_p"].cpu().numpy(), bad["log_p"].cpu().numpy()
    nr, nb = ref["n"].cpu().numpy(), bad["n"].cpu().numpy()
    for b in range(w.B):
        print(f"pair {b} Lx={len(w.xs[b])}: {lb[b]} vs {lr[b]}", flush=True)
        if not np.isnan(lb[b]):
            assert abs(lb[b] - lr[b]) <= 1e-12 * abs(lr[b]), b
            ok, worst = _close(nb[b], nr[b])
            assert ok, (b, worst)
    assert np.isnan(lb).any()
    assert int(bad["status"].item()) == 1
    with pytest.raises(_abi.TkfmiError):
        _abi.check_status(bad)
    good = _run(w, 0.45, ec=False)
    assert int(good["status"].item()) == 0
    lg = good["log_p"].cpu().numpy()
    assert np.all(np.abs(lg - lr) <= 1e-12 * np.abs(lr))
    ok, worst = _close(good["n"].cpu().numpy(), nr)
    assert ok, worst


def test_schedule_follows_call_size():
    """By default a call that the GPU holds at once (here 6 pairs) runs in
    one launch; in a call that does not (the 1024-pair benchmark batch) only
    the long class does (its few longest pairs), the others in two launches;
    both give the two-launch results."""
    from tkfmi import _abi, autotune, hmm, synth
    autotune.set_chain_schedule(*autotune.DEFAULT_SCHEDULE)
    w = synth.make_mixdom2(seed=5, B=6, lengths=[30, 200, 700, 64, 65, 400])
    res = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                         want_emit_counts=True, kernel="chain")
    ref = _run(w, 0.0)
    assert np.array_equal(res["log_p"].cpu().numpy(), ref["log_p"].cpu().numpy())
    ok, worst = _close(res["n"].cpu().numpy(), ref["n"].cpu().numpy(), rtol=1e-12)
    assert ok, worst
    r = _abi.BatchRunner(w.log_trans.to("cuda"), w.state_types, w.state_group, w.log_tab.to("cuda"),
                         _abi.PackedPairs(list(w.xs), list(w.ys), torch_device()), want_emit_counts=True,
                         time_kernels=True)
    r.run()
    kt = _abi.last_kernel_times()
    assert any("fbc_bidi_kernel" in k["name"] for k in kt), kt
    big = synth.make_mixdom2(seed=1, B=1024)
    r = _abi.BatchRunner(big.log_trans.to("cuda"), big.state_types, big.state_group, big.log_tab.to("cuda"),
                         _abi.PackedPairs(list(big.xs), list(big.ys), torch_device()), want_emit_counts=True,
                         time_kernels=True)
    r.run()
    kt = _abi.last_kernel_times()
    assert any(k["name"].startswith("fbc_fwd_kernel") for k in kt), kt
    assert all(k["name"].endswith(" long") for k in kt if "fbc_bidi_kernel" in k["name"]), kt


def torch_device():
    import torch
    return torch.device("cuda:0")


@pytest.mark.parametrize("cfg", [(3, 3, 0.5, 1), (4, 3, 0.3, 1)])
def test_long_class_one_launch(cfg):
    """The long class on the one-launch schedule and the other classes in two
    launches, in one call (default schedule limits): the oracle's results."""
    from tkfmi import autotune, hmm
    w, emit, lt_of = _mixdom(3, 1, 1, [5, 64, 65, 130, 260, 390, 520, 700, 800], seed=27,
                             ys={7: 760, 8: 640})
    autotune.set_chain_schedule(*autotune.DEFAULT_SCHEDULE)
    try:
        autotune.set_chain_config(*cfg)
        res = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                             want_emit_counts=True, kernel="chain")
    finally:
        autotune.set_chain_config(*autotune.DEFAULT)
    _check_vs_oracle(w, res, emit, lt_of)
