"""HIP-graph capture of a fixed-shape batched call (BatchRunner.capture):
every replay must reproduce the eager call, with other eager work between
replays.  logP bit-exact; counts within 1e-12 relative (FP64 atomics make the
summation order vary between runs)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _close(a, b, rtol=1e-12):
    a = a.detach().cpu().numpy()
    b = b.detach().cpu().numpy()
    scale = max(np.abs(b).max(), 1e-300)
    return bool(np.all(np.abs(a - b) <= rtol * np.abs(b) + 1e-14 * scale))


@pytest.mark.parametrize("kernel", ["chain", "class", "generic"])
def test_graph_replays_match_eager(kernel):
    from tkfmi import _abi, synth
    dev = torch.device("cuda", 0)
    w = synth.make_mixdom2(seed=7, B=40, lengths=[30 + 17 * k for k in range(40)])
    pairs = _abi.PackedPairs(w.xs, w.ys, dev)
    r = _abi.BatchRunner(w.log_trans.to(dev), w.state_types, w.state_group, w.log_tab.to(dev), pairs,
                         want_emit_counts=True, kernel=kernel)
    r.run()
    torch.cuda.synchronize()
    ref = {k: v.clone() for k, v in r.results().items()}
    graph = r.capture()
    # other eager work on the same library streams between replays
    w2 = synth.make_mixdom2(seed=8, B=5, lengths=[100, 7, 260, 64, 65])
    other = _abi.BatchRunner(w2.log_trans.to(dev), w2.state_types, w2.state_group, w2.log_tab.to(dev),
                             _abi.PackedPairs(w2.xs, w2.ys, dev), want_emit_counts=True, kernel=kernel)
    for rep in range(3):
        print(f"{kernel} replay {rep}", flush=True)
        for t in (r.log_p, r.n, r.ecnt):
            t.fill_(-7.0)  # stale values must all be overwritten / re-zeroed by the replay
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(r.log_p, ref["log_p"]), rep
        assert _close(r.n, ref["n"]), rep
        assert _close(r.ecnt, ref["emit_counts"]), rep
        other.run()
        torch.cuda.synchronize()
        assert torch.isfinite(other.log_p).all()
