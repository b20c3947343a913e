"""Concurrent calls from several host threads.

Each caller stream gets its own set of library class streams and fork/join
events (created on first use, one mutex per set held from the fork record to
the last join), so calls on different streams overlap on the device and
never pick up each other's event records.  Two threads, each with its own
runner (inputs, workspace, outputs), call repeatedly -- on two different
streams, and on one shared stream -- and every result must equal the serial
call's bit for bit in log P and to summation order (1e-13) in the counts."""
import threading

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _runner(w, kernel):
    from tkfmi import _abi
    dev = torch.device("cuda", 0)
    pairs = _abi.PackedPairs(list(w.xs), list(w.ys), dev)
    return _abi.BatchRunner(w.log_trans.to(dev), w.state_types, w.state_group, w.log_tab.to(dev),
                            pairs, want_emit_counts=True, kernel=kernel)


def _snap(r):
    return r.log_p.cpu().numpy().copy(), r.n.cpu().numpy().copy(), r.ecnt.cpu().numpy().copy()


def _same(a, b):
    assert np.array_equal(a[0], b[0])
    for x, y in ((a[1], b[1]), (a[2], b[2])):
        err = np.abs(x - y) / (np.abs(y) + 1e-12 * np.abs(y).max())
        assert err.max() <= 1e-13, float(err.max())


@pytest.mark.parametrize("kernel", ["chain", "class"])
@pytest.mark.parametrize("shared_stream", [False, True])
def test_two_threads_match_serial(kernel, shared_stream):
    from tkfmi import synth
    ws = [synth.make_mixdom2(seed=60 + k, B=24, lengths=list(np.random.default_rng(k).integers(30, 420, 24)))
          for k in range(2)]
    runners = [_runner(w, kernel) for w in ws]
    ref = []
    for r in runners:
        r.run()
        torch.cuda.synchronize()
        r.check_status()
        ref.append(_snap(r))
    streams = [torch.cuda.Stream()] if shared_stream else [torch.cuda.Stream(), torch.cuda.Stream()]
    errors, start = [], threading.Barrier(2)

    def work(k):
        try:
            s = streams[k % len(streams)]
            r = runners[k]
            start.wait()
            for _ in range(12):
                with torch.cuda.stream(s):
                    r.run(stream=s)
                s.synchronize()
                r.check_status()
                _same(_snap(r), ref[k])
        except Exception as e:  # reported on the main thread
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th), "a worker thread did not finish"
    assert not errors, errors
