"""The banded 16..160-state kernel (a pair's lattice rows in 16-row bands,
one workgroup of 8 waves per band, bands chained through boundary rows in
global memory) against the one-workgroup-per-pair antidiagonal kernel
(kernel="class") and the FP64 oracle.

Covers every padded state width (NBP = 32, 64, 96, 128, 160), band edges
(Lx + 1 = 16, 17, 32, 33 rows), one-row, flat and tall pairs, shared and
per-pair parameters, dense emissions, determinism (the count accumulation
runs in a fixed band order) and the bounded-wait abort path.  The two
kernels sum in different orders, so they agree to rounding: log P to 1e-12
relative, counts, posteriors and emission counts to 1e-10 relative with a
floor of 1e-12 x the largest entry."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
LENGTHS = [0, 15, 16, 17, 31, 32, 33, 100, 250, 1, 2]


def _close(a, b, rtol=1e-10, floor=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b)
    return bool(np.all(err <= rtol * np.abs(b) + floor * scale)), float((err / (np.abs(b) + floor * scale)).max())


def _workload(K, F, C, seed, lengths=LENGTHS):
    from tkfmi import synth
    w = synth.make_mixdom2(seed=seed, B=len(lengths), K=K, F=F, C=C, lengths=lengths)
    if len(lengths) < 11:
        return w
    rng = np.random.default_rng(seed)
    w.ys[-2] = rng.integers(0, 20, 300)   # flat: one ancestor residue, 300 descendant residues
    w.ys[-1] = w.ys[-1][:0]               # empty descendant
    w.ys[0] = rng.integers(0, 20, 7)      # empty ancestor, 7 descendant residues
    w.ys[7] = w.ys[7][:3]                 # tall: 100 x 3
    return w


def _run(w, kernel=None, shared=False, **kw):
    from tkfmi import _abi, hmm
    lt = w.log_trans[0] if shared else w.log_trans
    tab = w.log_tab[0] if shared else w.log_tab
    res = hmm.fb2d_batch(lt, w.state_types, w.xs, w.ys, tab, w.state_group, kernel=kernel, **kw)
    _abi.check_status(res)
    return res


def _compare(a, b, what, B):
    la, lb = a["log_p"].cpu().numpy(), b["log_p"].cpu().numpy()
    for q in range(B):
        assert abs(la[q] - lb[q]) <= 1e-12 * abs(lb[q]), (what, q, la[q], lb[q])
    if "n" in a:
        ok, worst = _close(a["n"].cpu().numpy(), b["n"].cpu().numpy())
        assert ok, (what, "n", worst)
    if "posteriors" in a:
        for q in range(B):
            ok, worst = _close(a["posteriors"][q].cpu().numpy(), b["posteriors"][q].cpu().numpy())
            assert ok, (what, "posteriors", q, worst)
    if "emit_counts" in a:
        ok, worst = _close(a["emit_counts"].cpu().numpy(), b["emit_counts"].cpu().numpy())
        assert ok, (what, "emit counts", worst)


@pytest.mark.parametrize("shared", [False, True])
@pytest.mark.parametrize("K,F,C", [pytest.param(2, 2, 2, id="NB20-NBP32"), pytest.param(3, 4, 1, id="NB60-NBP64"),
                                   pytest.param(4, 4, 2, id="NB80-NBP96"), pytest.param(5, 4, 2, id="NB100-NBP128"),
                                   pytest.param(8, 4, 4, id="NB160-NBP160")])
def test_band_equals_antidiagonal_kernel(K, F, C, shared):
    w = _workload(K, F, C, seed=300 + 10 * K + F)
    kw = dict(want_posteriors=True, want_emit_counts=True, shared=shared)
    print(f"K{K} F{F} C{C} shared={shared}: lengths {[(len(x), len(y)) for x, y in zip(w.xs, w.ys)]}", flush=True)
    band = _run(w, **kw)
    anti = _run(w, kernel="class", **kw)
    _compare(band, anti, f"K{K}F{F}", w.B)
    assert np.isfinite(band["log_p"].cpu().numpy()).all()


def test_band_edges_against_oracle():
    """Pairs whose ancestor fills exactly one band, one band plus a row,
    two bands and two bands plus a row, against the FP64 oracle."""
    from oracle import dp2d, emissions
    from tkfmi import hmm, synth
    w = synth.make_mixdom2(seed=333, B=4, K=3, F=2, C=2, lengths=[15, 16, 31, 32])
    res = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group,
                         want_posteriors=True, want_emit_counts=True)
    G = res["emit_counts"].shape[1]
    p = w.params
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        print(f"pair {b} Lx={len(x)} Ly={len(y)}", flush=True)
        em = emissions.pair_hmm_emissions_per_class(np.asarray(w.state_types), x, y, p["class_subs"][b],
                                                    p["class_pi"], p["class_dist"], p["K"], p["F"])
        lp_o, post_o, n_o, _ = dp2d.forward_backward(w.log_trans[b].numpy(), np.asarray(w.state_types), em)
        assert abs(float(res["log_p"][b]) - lp_o) <= 1e-10 * abs(lp_o), b
        assert _close(res["n"][b].cpu().numpy(), n_o)[0], b
        assert _close(res["posteriors"][b].cpu().numpy(), post_o)[0], b
        ec_o = emissions.emission_counts(post_o, w.state_types, w.state_group, x, y, G)
        assert _close(res["emit_counts"][b].cpu().numpy(), ec_o)[0], b


def test_band_is_deterministic_and_batch_independent():
    """Counts are summed over a pair's bands in a fixed order, so two runs,
    and a pair run alone or inside a batch, agree bit for bit (log P,
    transition counts, posteriors)."""
    w = _workload(8, 2, 2, seed=41, lengths=[300, 17, 120, 64, 500])
    a = _run(w, want_posteriors=True)
    b = _run(w, want_posteriors=True)
    assert np.array_equal(a["log_p"].cpu().numpy(), b["log_p"].cpu().numpy())
    assert np.array_equal(a["n"].cpu().numpy(), b["n"].cpu().numpy())
    for q in range(w.B):
        assert np.array_equal(a["posteriors"][q].cpu().numpy(), b["posteriors"][q].cpu().numpy()), q
    from tkfmi import hmm
    for q in (0, 2, 4):
        alone = hmm.fb2d_batch(w.log_trans[q], w.state_types, [w.xs[q]], [w.ys[q]], w.log_tab[q], w.state_group)
        assert float(alone["log_p"][0]) == float(a["log_p"][q]), q
        assert np.array_equal(alone["n"][0].cpu().numpy(), a["n"][q].cpu().numpy()), q
    fo = _run(w, forward_only=True)
    assert np.array_equal(fo["log_p"].cpu().numpy(), a["log_p"].cpu().numpy())


def test_band_dense_emissions_equal_antidiagonal_kernel():
    from oracle import emissions
    from tkfmi import hmm
    w = _workload(8, 2, 2, seed=52, lengths=[40, 0, 17, 90, 16])
    rng = np.random.default_rng(3)
    st = np.asarray(w.state_types)
    emitting = (st == 1) | (st == 2) | (st == 3)
    dense = []
    for b in range(w.B):
        e = emissions.emit_from_tables(st, w.state_group, w.log_tab[b].numpy(), w.xs[b], w.ys[b])
        e = e + rng.normal(0.0, 0.3, e.shape)
        e[(rng.random(e.shape) < 0.05) & emitting[None, None, :]] = -1e30
        dense.append(torch.as_tensor(e))
    run = lambda kernel: hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, None, w.state_group,
                                        want_posteriors=True, log_emit=dense, kernel=kernel)
    _compare(run(None), run("class"), "dense", w.B)


def test_band_long_pairs_count_identities():
    """K=8 F=4 C=4 (S = 162) at lengths up to 1000 with full lattices (the
    checkpoint budget switched off, so the banded kernel runs): sum of counts
    into M/D states == Lx, into M/I == Ly, start row == 1."""
    from tkfmi import _abi, hmm, synth
    w = synth.make_mixdom2(seed=2, B=3, K=8, F=4, C=4, lengths=[1000, 400, 730])
    try:
        _abi.set_checkpoint_budget(0)
        res = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group)
    finally:
        _abi.set_checkpoint_budget(_abi.CHECKPOINT_BUDGET_DEFAULT)
    _abi.check_status(res)
    types = np.asarray(w.state_types)
    n = res["n"].cpu().numpy()
    for b, (x, y) in enumerate(zip(w.xs, w.ys)):
        nb = n[b]
        print(f"pair {b} Lx={len(x)} Ly={len(y)} logP={float(res['log_p'][b])}", flush=True)
        assert np.isfinite(float(res["log_p"][b]))
        assert abs(nb[:, np.isin(types, (1, 3))].sum() - len(x)) <= 1e-9 * len(x)
        assert abs(nb[:, np.isin(types, (1, 2))].sum() - len(y)) <= 1e-9 * len(y)
        assert abs(nb[0].sum() - 1.0) <= 1e-12


def test_band_wait_timeout_aborts_cleanly(monkeypatch):
    """A boundary wait that times out (forced from the host with a spin
    limit of 1) ends the launch: one-band pairs (no wait) keep their values,
    multi-band pairs either completed or report NaN, the status word is set;
    the next call with the normal limit is exact."""
    from tkfmi import _abi
    w = _workload(3, 2, 1, seed=61, lengths=[10, 400, 5, 300, 12, 350, 0, 200, 14, 1, 2])
    ref = _run(w)
    monkeypatch.setenv("TKFMI_CHAIN_SPIN_LIMIT", "1")
    from tkfmi import hmm
    bad = hmm.fb2d_batch(w.log_trans, w.state_types, w.xs, w.ys, w.log_tab, w.state_group)
    monkeypatch.delenv("TKFMI_CHAIN_SPIN_LIMIT")
    lr, lb = ref["log_p"].cpu().numpy(), bad["log_p"].cpu().numpy()
    for b in range(w.B):
        print(f"pair {b} Lx={len(w.xs[b])}: {lb[b]} vs {lr[b]}", flush=True)
        assert np.isnan(lb[b]) or lb[b] == lr[b], b
    for b in (0, 2, 4, 6, 8, 9, 10):  # one band: no cross-workgroup wait
        assert lb[b] == lr[b], b
    assert np.isnan(lb[[1, 3, 5, 7]]).any()
    assert int(bad["status"].item()) == 1
    with pytest.raises(_abi.TkfmiError):
        _abi.check_status(bad)
    good = _run(w)
    assert int(good["status"].item()) == 0
    assert np.array_equal(good["log_p"].cpu().numpy(), lr)


@pytest.mark.parametrize("K,F,C,n,lo,hi", [pytest.param(3, 2, 1, 240, 150, 400, id="S32"),
                                           pytest.param(8, 4, 4, 300, 120, 260, id="S162")])
def test_band_checkpoint_recomputation_is_bit_identical(K, F, C, n, lo, hi):
    """TKFMI_CHECKPOINT on the banded kernel: the forward keeps no lattice,
    each backward band recomputes its own from the forward's boundary row
    into a slot (batches with many more bands than resident workgroups, so
    that the slots save most of the lattice memory); log P, transition
    counts and posteriors equal the full-lattice run bit for bit, emission
    counts to the atomics' summation order."""
    from tkfmi import _abi, synth
    rng = np.random.default_rng(70 + K)
    w = synth.make_mixdom2(seed=80 + K, B=n, K=K, F=F, C=C, lengths=rng.integers(lo, hi, n))
    dev = torch.device("cuda", 0)
    pairs = _abi.PackedPairs(w.xs, w.ys, dev)
    out = {}
    try:
        _abi.set_checkpoint_budget(0)
        for ck in (False, True):
            r = _abi.BatchRunner(w.log_trans.to(dev), w.state_types, w.state_group, w.log_tab.to(dev), pairs,
                                 want_posteriors=True, want_emit_counts=True, checkpoint=ck, time_kernels=True)
            r.run()
            names = [k["name"] for k in _abi.last_kernel_times()]
            print(f"K{K} F{F} checkpoint={ck}: workspace {r.ws_bytes / 2**30:.2f} GiB, kernels {names}", flush=True)
            assert any(nm.startswith("band_bwd_kernel") for nm in names), names
            res = r.results()
            _abi.check_status(res)
            out[ck] = ({k: (v.cpu().numpy() if torch.is_tensor(v) else v) for k, v in res.items()}, r.ws_bytes)
            del r
    finally:
        _abi.set_checkpoint_budget(_abi.CHECKPOINT_BUDGET_DEFAULT)
    full, ws_full = out[False]
    ck, ws_ck = out[True]
    assert ws_ck < ws_full / 2, (ws_ck, ws_full)
    assert np.array_equal(full["log_p"], ck["log_p"])
    assert np.array_equal(full["n"], ck["n"])
    assert np.array_equal(full["posteriors"], ck["posteriors"])
    ok, worst = _close(ck["emit_counts"], full["emit_counts"], rtol=1e-12)
    assert ok, worst
