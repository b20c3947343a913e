"""tkfmi_autotune (tkfmi.autotune.tune): times the chained kernels'
candidate configurations on the caller's batch and keeps the fastest per
(layout, length bin).  The results of the tuned calls equal the generic
kernel's (parity is not a function of the tuning); a batch of another
length bin keeps the default; a layout without a tunable configuration
reports none."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _runner(w, kernel=None):
    from tkfmi import _abi
    dev = torch.device("cuda", 0)
    pairs = _abi.PackedPairs(list(w.xs), list(w.ys), dev)
    return _abi.BatchRunner(w.log_trans.to(dev), w.state_types, w.state_group, w.log_tab.to(dev),
                            pairs, want_emit_counts=True, kernel=kernel)


def test_autotune_keeps_results_and_keys():
    from tkfmi import autotune, synth
    autotune.set_chain_config(*autotune.DEFAULT)  # no tuned keys
    w = synth.make_mixdom2(seed=31, B=96, lengths=list(np.random.default_rng(1).integers(40, 700, 96)))
    r = _runner(w)
    res = autotune.tune(r, reps=2)
    print(res, flush=True)
    assert res["key"][1] == 3 and len(res["ms"]) >= 6  # longest sequence in [512, 1024)
    assert all(v > 0 for v in res["ms"].values())
    assert res["best"] in res["ms"]
    r.run()
    ref = _runner(w, kernel="generic")
    ref.run()
    torch.cuda.synchronize()
    r.check_status()
    assert np.allclose(r.log_p.cpu().numpy(), ref.log_p.cpu().numpy(), rtol=1e-12, atol=0)
    assert np.allclose(r.n.cpu().numpy(), ref.n.cpu().numpy(), rtol=1e-11, atol=1e-12)
    # the wide kernel has one configuration
    w2 = synth.make_mixdom2(seed=32, B=2, K=4, F=2, lengths=[30, 40])
    res2 = autotune.tune(_runner(w2))
    assert res2["ms"] == {} and res2["best"] is None
    autotune.set_chain_config(*autotune.DEFAULT)
