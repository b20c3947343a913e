"""GPU parity: libtkfmi batched forward/backward vs the FP64 oracle.

Tolerance: 1e-10 relative (logP, counts, posteriors); entries that are
tiny compared with the largest of their kind get an absolute floor of
1e-12 x that maximum.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _close(a, b, rtol=RTOL, floor=1e-12):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b)
    ok = err <= rtol * np.abs(b) + floor * scale
    return bool(ok.all()), float((err / (np.abs(b) + floor * scale)).max())


def _oracle_pair(log_trans, types, emit):
    from oracle import dp2d
    return dp2d.forward_backward(log_trans, np.asarray(types), emit)


def _run(w, forward_only=False, want_post=True, kernel=None):
    from tkfmi import hmm
    return hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                          forward_only=forward_only, want_posteriors=want_post and not forward_only,
                          kernel=kernel)


KERNELS = [pytest.param("chain", id="chain"), pytest.param("class", id="class"),
           pytest.param("generic", id="generic")]


def _tkf_workload(lengths_x, lengths_y, seed=0, ext=None):
    from tkfmi import synth
    import torch
    rng = np.random.default_rng(seed)
    w = synth.make_tkf(seed=seed, B=1, ext=ext)
    w.xs = [rng.integers(0, 20, size=n) for n in lengths_x]
    w.ys = [rng.integers(0, 20, size=n) for n in lengths_y]
    w.ts = np.full(len(w.xs), w.ts[0])
    return w


def _tkf_oracle_emit(w, x, y):
    from oracle import emissions
    return emissions.pair_hmm_emissions(np.asarray(w.state_types), x, y,
                                        w.params["sub"], w.params["pi"])


EDGE_X = [0, 0, 3, 1, 2, 37, 64, 65, 70, 5, 130]
EDGE_Y = [0, 4, 0, 1, 2, 64, 37, 3, 1, 140, 96]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("ext", [None, 0.6])
def test_tkf_edge_lengths_match_oracle(ext, kernel):
    w = _tkf_workload(EDGE_X + [300, 20], EDGE_Y + [520, 700], seed=3, ext=ext)
    res = _run(w, kernel=kernel)
    lt = w.log_trans.numpy()
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        print(f"pair {b} Lx={len(x)} Ly={len(y)}", flush=True)
        lp_o, post_o, n_o, _ = _oracle_pair(lt, w.state_types, _tkf_oracle_emit(w, x, y))
        lp = float(res["log_p"][b])
        assert abs(lp - lp_o) <= RTOL * abs(lp_o), (b, lp, lp_o)
        ok, worst = _close(res["n"][b].cpu().numpy(), n_o)
        assert ok, (b, worst)
        ok, worst = _close(res["posteriors"][b].cpu().numpy(), post_o)
        assert ok, (b, worst)


@pytest.mark.parametrize("kernel", KERNELS)
def test_tkf91_forward_only_matches_oracle(kernel):
    from oracle import dp2d
    w = _tkf_workload([10, 50, 150, 0], [12, 47, 160, 9], seed=5)
    res = _run(w, forward_only=True, kernel=kernel)
    assert "n" not in res
    lt = w.log_trans.numpy()
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        lp_o, _ = dp2d.forward(lt, np.asarray(w.state_types), _tkf_oracle_emit(w, x, y))
        assert abs(float(res["log_p"][b]) - lp_o) <= RTOL * abs(lp_o)


def _mixdom_oracle_emit(w, b, x, y):
    from oracle import emissions
    p = w.params
    return emissions.pair_hmm_emissions_per_class(np.asarray(w.state_types), x, y,
                                                  p["class_subs"][b], p["class_pi"],
                                                  p["class_dist"], p["K"], p["F"])


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("K,F,C", [(3, 1, 1), (3, 1, 2), (1, 3, 2), (2, 1, 3), (1, 2, 1), (1, 1, 2)])
def test_mixdom2_matches_oracle(K, F, C, kernel):
    from tkfmi import synth
    lengths = [0, 1, 31, 63, 64, 65, 130, 200, 140]
    w = synth.make_mixdom2(seed=11 + K + F + C, B=len(lengths), K=K, F=F, C=C, lengths=lengths)
    # include empty / very short / long descendants (1..8 waves per pair)
    w.ys[0] = w.ys[0][:0]
    w.ys[1] = w.ys[1][:3]
    rng = np.random.default_rng(K * 100 + F * 10 + C)
    w.ys[7] = rng.integers(0, 20, size=560)
    w.ys[8] = rng.integers(0, 20, size=300)
    res = _run(w, kernel=kernel)
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        print(f"pair {b} Lx={len(x)} Ly={len(y)}", flush=True)
        lt = w.log_trans[b].numpy()
        lp_o, post_o, n_o, _ = _oracle_pair(lt, w.state_types, _mixdom_oracle_emit(w, b, x, y))
        lp = float(res["log_p"][b])
        assert abs(lp - lp_o) <= RTOL * abs(lp_o), (b, lp, lp_o)
        ok, worst = _close(res["n"][b].cpu().numpy(), n_o)
        assert ok, (b, worst)
        ok, worst = _close(res["posteriors"][b].cpu().numpy(), post_o)
        assert ok, (b, worst)


@pytest.mark.parametrize("kernel", KERNELS)
def test_batched_equals_alone(kernel):
    """A pair's results do not depend on the rest of the batch (bitwise)."""
    from tkfmi import hmm, synth
    w = synth.make_mixdom2(seed=4, B=12, lengths=[40, 200, 7, 90, 150, 33, 64, 65, 2, 180, 120, 99])
    full = _run(w, want_post=False, kernel=kernel)
    for b in (0, 3, 8):
        alone = hmm.fb2d_batch(w.log_trans[b], w.state_types, [w.xs[b]], [w.ys[b]],
                               w.log_tab[b], w.state_group, kernel=kernel)
        assert float(alone["log_p"][0]) == float(full["log_p"][b])
        assert np.array_equal(alone["n"][0].cpu().numpy(), full["n"][b].cpu().numpy())


@pytest.mark.parametrize("kernel", KERNELS)
def test_count_identities_full_size(kernel):
    """Size-independent identities at the benchmark's length range:
    sum of counts into M/D states == Lx, into M/I states == Ly, start row sums
    to 1, and flow conservation for body states."""
    from tkfmi import synth
    w = synth.make_mixdom2(seed=7, B=6, lengths=[800, 30, 400, 650, 200, 777])
    res = _run(w, want_post=False, kernel=kernel)
    types = np.asarray(w.state_types)
    n = res["n"].cpu().numpy()
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        nb = n[b]
        into_x = nb[:, np.isin(types, (1, 3))].sum()
        into_y = nb[:, np.isin(types, (1, 2))].sum()
        assert abs(into_x - len(x)) <= 1e-9 * len(x) + 1e-9
        assert abs(into_y - len(y)) <= 1e-9 * len(y) + 1e-9
        assert abs(nb[0].sum() - 1.0) < 1e-12
        body = np.isin(types, (1, 2, 3))
        inflow = nb[:, body].sum(axis=0)
        outflow = nb[body, :].sum(axis=1)
        assert np.allclose(inflow, outflow, rtol=1e-10, atol=1e-10)
        assert np.isfinite(float(res["log_p"][b]))


@pytest.mark.parametrize("kernel", KERNELS)
def test_long_rows_global_row_buffer_matches_golden(kernel):
    """Descendants too long for the LDS row buffer (1200 and 2000 residues)
    use the global one, in one launch with a short pair that uses LDS.
    Expected values: tests/golden/long_pairs.npz, the oracle in 80-bit long
    double (the float64 log-space recursion itself drifts ~1e-10 here)."""
    import os
    from tkfmi import hmm
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "long_pairs.npz"))
    B = 3
    xs = [g[f"x{b}"] for b in range(B)]
    ys = [g[f"y{b}"] for b in range(B)]
    res = hmm.fb2d_batch(g["log_trans"], g["state_types"], xs, ys, g["log_tab"],
                         list(g["state_group"]), want_posteriors=True, kernel=kernel)
    stride = int(g["post_stride"])
    for b in range(B):
        print(f"pair {b} Lx={len(xs[b])} Ly={len(ys[b])}", flush=True)
        lp_o = float(g[f"logp{b}"])
        lp = float(res["log_p"][b])
        assert abs(lp - lp_o) <= RTOL * abs(lp_o), (b, lp, lp_o)
        ok, worst = _close(res["n"][b].cpu().numpy(), g[f"n{b}"])
        assert ok, (b, worst)
        post = res["posteriors"][b].cpu().numpy()[:, ::stride, :]
        ok, worst = _close(post, g[f"post{b}"])
        assert ok, (b, worst)


@pytest.mark.parametrize("kernel", ["chain", "class"])
def test_cell_per_lane_matches_generic_kernel_long_rows(kernel):
    """Long rows: in the one-workgroup kernel the wrap-around ring does not
    fit in LDS (8 waves, Ly >= 1500) and goes through global memory; in the
    chained kernel pairs of 10-12 bands span 3 workgroups.  Compare with the
    generic kernel and the exact count identities (sum into M/D == Lx, into
    M/I == Ly)."""
    from tkfmi import synth
    w = synth.make_mixdom2(seed=33, B=4, lengths=[600, 700, 130, 520])
    rng = np.random.default_rng(5)
    w.ys[0] = rng.integers(0, 20, size=2000)
    w.ys[1] = rng.integers(0, 20, size=1500)
    w.ys[2] = rng.integers(0, 20, size=900)
    a = _run(w, want_post=True, kernel=kernel)
    b = _run(w, want_post=True, kernel="generic")
    types = np.asarray(w.state_types)
    for p in range(w.B):
        print(f"pair {p} Lx={len(w.xs[p])} Ly={len(w.ys[p])}", flush=True)
        la, lb = float(a["log_p"][p]), float(b["log_p"][p])
        assert abs(la - lb) <= 1e-12 * abs(lb), (p, la, lb)
        na, nb_ = a["n"][p].cpu().numpy(), b["n"][p].cpu().numpy()
        ok, worst = _close(na, nb_)
        assert ok, (p, worst)
        ok, worst = _close(a["posteriors"][p].cpu().numpy(), b["posteriors"][p].cpu().numpy())
        assert ok, (p, worst)
        assert abs(na[:, np.isin(types, (1, 3))].sum() - len(w.xs[p])) < 1e-9 * len(w.xs[p])
        assert abs(na[:, np.isin(types, (1, 2))].sum() - len(w.ys[p])) < 1e-9 * len(w.ys[p])


@pytest.mark.parametrize("kernel", ["chain", "class"])
def test_c1_subset_both_kernels_agree(kernel):
    """The benchmark workload's first 96 pairs: cell-per-lane == generic."""
    from tkfmi import synth
    w = synth.make_mixdom2(seed=1, B=96)
    a = _run(w, want_post=False, kernel=kernel)
    b = _run(w, want_post=False, kernel="generic")
    la, lb = a["log_p"].cpu().numpy(), b["log_p"].cpu().numpy()
    assert np.all(np.abs(la - lb) <= 1e-12 * np.abs(lb))
    ok, worst = _close(a["n"].cpu().numpy(), b["n"].cpu().numpy())
    assert ok, worst


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("model", ["tkf92", "mixdom2_k3c2", "mixdom2_k1f3"])
def test_emit_counts_match_oracle(model, kernel):
    """Expected emission counts (d logP / d log_tab) vs the oracle's
    posterior reduction, over edge and long lengths."""
    from oracle import emissions
    from tkfmi import synth
    if model == "tkf92":
        w = _tkf_workload([0, 3, 1, 64, 65, 130, 300], [4, 0, 1, 63, 140, 96, 310], seed=9, ext=0.5)
        emit = lambda b, x, y: _tkf_oracle_emit(w, x, y)
        lt_of = lambda b: w.log_trans.numpy()
    else:
        K, F, C = (3, 1, 2) if model == "mixdom2_k3c2" else (1, 3, 1)
        w = synth.make_mixdom2(seed=21, B=7, K=K, F=F, C=C, lengths=[0, 1, 40, 64, 65, 200, 330])
        emit = lambda b, x, y: _mixdom_oracle_emit(w, b, x, y)
        lt_of = lambda b: w.log_trans[b].numpy()
    res = _run_ec(w, kernel)
    G = res["emit_counts"].shape[1]
    ec = res["emit_counts"].cpu().numpy()
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        print(f"pair {b} Lx={len(x)} Ly={len(y)}", flush=True)
        lp_o, post_o, n_o, _ = _oracle_pair(lt_of(b), w.state_types, emit(b, x, y))
        want = emissions.emission_counts(post_o, w.state_types, w.state_group, x, y, G)
        ok, worst = _close(ec[b], want)
        assert ok, (b, worst)
        assert abs(float(res["log_p"][b]) - lp_o) <= RTOL * abs(lp_o)


def _run_ec(w, kernel):
    from tkfmi import hmm
    return hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                          want_emit_counts=True, kernel=kernel)


@pytest.mark.parametrize("kernel", KERNELS)
def test_emit_counts_full_size_identities(kernel):
    """At the benchmark's lengths: emission totals per type equal the
    expected visits in n, and every residue of x (of y) is emitted once."""
    from tkfmi import synth
    w = synth.make_mixdom2(seed=8, B=5, lengths=[800, 650, 30, 420, 777])
    res = _run_ec(w, kernel)
    types = np.asarray(w.state_types)
    n = res["n"].cpu().numpy()
    ec = res["emit_counts"].cpu().numpy()
    nm = 21 * 21
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        e = ec[b]
        assert abs(e[:, :nm].sum() - n[b][:, types == 1].sum()) <= 1e-10 * len(x)
        assert abs(e[:, nm:nm + 21].sum() - n[b][:, types == 2].sum()) <= 1e-10 * len(y) + 1e-12
        m = e[:, :nm].reshape(-1, 21, 21)
        np.testing.assert_allclose(m.sum(axis=(0, 2)) + e[:, nm + 21:].sum(axis=0),
                                   np.bincount(np.minimum(x, 20), minlength=21), rtol=1e-10, atol=1e-9)
        np.testing.assert_allclose(m.sum(axis=(0, 1)) + e[:, nm:nm + 21].sum(axis=0),
                                   np.bincount(np.minimum(y, 20), minlength=21), rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("cfg", [(4, 3, 0.3), (4, 4, 0.0), (3, 3, 0.3), (2, 4, 0.05), (3, 2, 0.3)])
def test_chain_configs_agree_with_generic(cfg):
    """Every chained-kernel configuration (workgroup widths per length class)
    gives the generic kernel's results."""
    from tkfmi import autotune, synth
    w = synth.make_mixdom2(seed=12, B=24, lengths=[5, 40, 64, 65, 130, 200, 260, 300, 390, 450, 520, 700] * 2)
    b = _run(w, want_post=False, kernel="generic")
    try:
        autotune.set_chain_config(*cfg)
        a = _run(w, want_post=False, kernel="chain")
    finally:
        autotune.set_chain_config(*autotune.DEFAULT)
    la, lb = a["log_p"].cpu().numpy(), b["log_p"].cpu().numpy()
    assert np.all(np.abs(la - lb) <= 1e-12 * np.abs(lb)), cfg
    ok, worst = _close(a["n"].cpu().numpy(), b["n"].cpu().numpy())
    assert ok, (cfg, worst)


@pytest.mark.parametrize("kernel", KERNELS)
def test_tall_and_flat_pairs_match_oracle(kernel):
    """Pairs whose lattice is one cell wide or high over many bands (the
    chained kernel's cross-workgroup rows carry a single column)."""
    w = _tkf_workload([1000, 900, 640, 0, 2, 257], [0, 1, 3, 900, 700, 2], seed=4, ext=0.4)
    res = _run(w, want_post=False, kernel=kernel)
    lt = w.log_trans.numpy()
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        print(f"pair {b} Lx={len(x)} Ly={len(y)}", flush=True)
        lp_o, _post, n_o, _ = _oracle_pair(lt, w.state_types, _tkf_oracle_emit(w, x, y))
        lp = float(res["log_p"][b])
        assert abs(lp - lp_o) <= RTOL * abs(lp_o), (b, lp, lp_o)
        ok, worst = _close(res["n"][b].cpu().numpy(), n_o)
        assert ok, (b, worst)


def test_chain_wait_timeout_aborts_cleanly(monkeypatch):
    """A wait that times out (forced from the host with a spin limit of 1)
    must end the launch: pairs computed without a wait keep their values,
    pairs whose chain waited either completed or report NaN, never a wrong
    number; the next call with the normal limit is exact.  Two-launch
    schedule (forward, then backward), where one-band pairs never wait; the
    one-launch schedule's version is in test_gpu_bidi.py."""
    from tkfmi import autotune, synth
    w = synth.make_mixdom2(seed=31, B=8, lengths=[40, 700, 20, 650, 60, 800, 30, 500])
    ref = _run(w, want_post=False, kernel="generic")
    monkeypatch.setenv("TKFMI_CHAIN_SPIN_LIMIT", "1")
    autotune.set_chain_schedule(0.0)
    try:
        bad = _run(w, want_post=False, kernel="chain")
    finally:
        autotune.set_chain_schedule(*autotune.DEFAULT_SCHEDULE)
    monkeypatch.delenv("TKFMI_CHAIN_SPIN_LIMIT")
    lr, lb = ref["log_p"].cpu().numpy(), bad["log_p"].cpu().numpy()
    for b in range(w.B):
        print(f"pair {b} Lx={len(w.xs[b])}: {lb[b]} vs {lr[b]}", flush=True)
        assert np.isnan(lb[b]) or abs(lb[b] - lr[b]) <= 1e-12 * abs(lr[b]), b
    for b in (0, 2, 4, 6):  # one band: no cross-workgroup wait
        assert abs(lb[b] - lr[b]) <= 1e-12 * abs(lr[b]), b
    # a chain's first wait comes ~200 steps before its producer can publish,
    # so with one spin it times out: the abort path did run
    assert np.isnan(lb[[1, 3, 5, 7]]).any()
    # the call's status word says so, and the callers' check raises on it
    from tkfmi import _abi
    assert int(bad["status"].item()) == 1
    with pytest.raises(_abi.TkfmiError):
        _abi.check_status(bad)
    good = _run(w, want_post=False, kernel="chain")
    assert int(good["status"].item()) == 0
    _abi.check_status(good)
    lg = good["log_p"].cpu().numpy()
    assert np.all(np.abs(lg - lr) <= 1e-12 * np.abs(lr))
    ok, worst = _close(good["n"].cpu().numpy(), ref["n"].cpu().numpy())
    assert ok, worst
