"""Continuous-batching request loop for pair scoring and E-step requests.

Requests (a pair x, y at branch length t, under TKF92 or MixDom2, either
"score" = log-likelihood only or "estep" = log-likelihood, expected
transition counts and the model's sufficient statistics) arrive over time.
One scheduler thread:

* admits arrived requests in arrival order while the HBM budget holds their
  lattices (the forward lattice of a pair is (Lx+64)(Ly+64) x 128 bytes;
  scoring needs none) and the batch has room;
* groups requests by (model, kind, size bucket) -- a call finishes when its
  longest pair does, so short pairs never ride with long ones -- and, when
  a stream slot is free, launches the group of the oldest waiting request
  (up to ``max_batch_pairs`` of its requests, in arrival order) as one
  forward/backward call, followed for E-steps by the statistics kernel
  (TKF92 self-loop split + BDI, or MixDom2 null-state restoration + BDI),
  then records an event;
* polls the events of in-flight calls; a finished call releases its memory
  and slot, and its requests' latencies (completion - arrival) are
  recorded.

So work that arrives while earlier batches are on the GPU is admitted as
soon as a slot and memory free up, and overlaps with them on the device.
Parameters stay device-resident; each call builds its per-pair transition
matrices and emission tables for the pairs' own t on the device.

A pair's results do not depend on the batch it rides in (the kernels give
every pair its own workgroup and fixed summation orders), which
``tests/test_gpu_serve.py`` checks request by request.
"""
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi, build as dbuild, mixdom, params, tables

LATTICE_BYTES_PER_CELL = 128


@dataclass
class Request:
    rid: int
    model: str        # "tkf92" or "mixdom2"
    kind: str         # "score" or "estep"
    x: np.ndarray
    y: np.ndarray
    t: float
    arrival: float = 0.0  # seconds from the start of the run

    def bucket(self):
        """Size class: the number of waves the kernel gives the pair grows
        with its length, and so does its serial step chain."""
        c = (len(self.x) + 1) * (len(self.y) + 1)
        return 0 if c <= 160 * 160 else (1 if c <= 400 * 400 else (2 if c <= 900 * 900 else 3))

    def key(self):
        return (self.model, self.kind, self.bucket())

    def footprint(self):
        if self.kind == "score":
            return 0
        return (len(self.x) + 64) * (len(self.y) + 64) * LATTICE_BYTES_PER_CELL

    def cell_states(self, S):
        return (len(self.x) + 1) * (len(self.y) + 1) * S


@dataclass
class Result:
    log_p: float
    n: np.ndarray = None
    stats: np.ndarray = None
    latency: float = 0.0


def gtr_Q(S_exch, pi):
    """Q[a,b] = S[a,b] pi[b] off the diagonal, rows summing to zero."""
    Q = np.asarray(S_exch, np.float64) * np.asarray(pi, np.float64)[..., None, :]
    A = Q.shape[-1]
    Q = Q * (1.0 - np.eye(A))
    return Q - np.eye(A) * Q.sum(-1)[..., :, None]


class _Model:
    """Device-resident parameters of one model; its per-pair transition
    matrices (log floor 1e-30) and emission tables are built on the device
    for each call's t (tkfmi.build)."""

    def __init__(self, name, p, dev):
        self.name = name
        self.dev = dev
        if name == "tkf92":
            self.lam, self.mu, self.ext = float(p["lam"]), float(p["mu"]), float(p["ext"])
            self.builder = dbuild.DeviceBuilder("tkf92", "shared", p["Q"], p["pi"], lam=self.lam,
                                                mu=self.mu, ext=self.ext, log_floor=True, device=dev)
        elif name == "mixdom2":
            self.p = p
            mix = dict(main_ins=p["main_lam"], main_del=p["main_mu"], dom_ins=p["dom_lam"],
                       dom_del=p["dom_mu"], dom_w=p["dom_w"], frag_w=p["frag_w"], ext=p["ext"])
            self.builder = dbuild.DeviceBuilder("mixdom", "per_class", gtr_Q(p["class_S"], p["class_pi"]),
                                                p["class_pi"], mix=mix, class_dist=p["class_dist"],
                                                log_floor=True, device=dev)
            self.model = None
        else:
            raise ValueError(f"unknown model {name}")
        self.st = self.builder.state_types
        self.groups = self.builder.state_group
        self.S = len(self.st)

    def build(self, ts):
        """(log_trans (B,S,S), log_tab (B,G,TAB), t) at per-pair t."""
        return self.builder.build(ts)


@dataclass
class _Call:
    reqs: list
    runner: object
    t: torch.Tensor = None
    stats: tuple = ()      # device outputs of the statistics kernel
    keep: tuple = ()       # other device buffers the queued kernels use


@dataclass
class _Batch:
    calls: list
    event: object
    bytes: int
    slot: int
    keep: list = field(default_factory=list)


class ServingEngine:
    def __init__(self, tkf92_params, mixdom2_params, device=None, budget_bytes=16 << 30,
                 max_batch_pairs=512, slots=6):
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        self.models = {"tkf92": _Model("tkf92", tkf92_params, self.dev),
                       "mixdom2": _Model("mixdom2", mixdom2_params, self.dev)}
        self.budget = int(budget_bytes)
        self.max_batch_pairs = int(max_batch_pairs)
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(slots)]

    # ---------------------------------------------------------------- launch
    def _launch_group(self, model, kind, reqs, stream):
        m = self.models[model]
        with torch.cuda.stream(stream):
            lt, tab, t = m.build([r.t for r in reqs])
            pairs = _abi.PackedPairs([r.x for r in reqs], [r.y for r in reqs], self.dev)
            r = _abi.BatchRunner(lt, m.st, m.groups, tab, pairs, forward_only=(kind == "score"))
            r.run(zero_outputs=False, stream=stream)
            call = _Call(reqs, r, t)
            if kind == "estep":
                call.stats, call.keep = self._estep_stats(m, r, t, stream)
        return call

    def _estep_stats(self, m, r, t, stream):
        """Statistics kernel of an E-step group; returns (outputs, buffers)."""
        L = _abi.lib()
        sp = _abi.ctypes.c_void_p(stream.cuda_stream)
        B = r.pairs.B
        if m.name == "tkf92":
            out = torch.empty((B, _abi.TKF92_STATS), dtype=torch.float64, device=self.dev)
            _abi.check(L.tkfmi_tkf92_estep_stats(B, m.lam, m.mu, m.ext, _abi._ptr(t), _abi._ptr(r.n),
                                                 _abi._ptr(out), sp))
            return (out,), ()
        from .suffstats import _Model as SSModel
        p = m.p
        if m.model is None:
            m.model = SSModel(p["main_lam"], p["main_mu"], p["dom_lam"], p["dom_mu"], p["dom_w"],
                              p["frag_w"], p["ext"])
        mm = m.model
        nst = int(L.tkfmi_suffstats_size(mm.K, mm.F))
        wsb = int(L.tkfmi_restore_workspace_bytes(_abi.ctypes.byref(mm.c), B))
        ws = torch.empty(wsb, dtype=torch.uint8, device=self.dev)
        st_buf = torch.empty((B, nst), dtype=torch.float64, device=self.dev)
        bdi_buf = torch.empty((B, (mm.K + 1) * 4), dtype=torch.float64, device=self.dev)
        _abi.check(L.tkfmi_restore_counts_batched(_abi.ctypes.byref(mm.c), B, _abi._ptr(t),
                                                  _abi._ptr(r.n), _abi._ptr(st_buf), _abi._ptr(bdi_buf),
                                                  _abi._ptr(ws), _abi.ctypes.c_size_t(wsb), sp))
        return (st_buf, bdi_buf), (ws,)

    def _launch(self, reqs, slot):
        stream = self.streams[slot]
        call = self._launch_group(reqs[0].model, reqs[0].kind, reqs, stream)
        ev = torch.cuda.Event()
        ev.record(stream)
        return _Batch([call], ev, sum(r.footprint() for r in reqs), slot)

    def _collect(self, batch, results, t0, clock):
        now = clock() - t0
        for c in batch.calls:
            lp = c.runner.log_p.cpu().numpy()
            n = c.runner.n.cpu().numpy() if c.runner.n is not None else None
            st = np.concatenate([x.cpu().numpy() for x in c.stats], axis=1) if c.stats else None
            for i, r in enumerate(c.reqs):
                results[r.rid] = Result(float(lp[i]), None if n is None else n[i],
                                        None if st is None else st[i], now - r.arrival)

    # ----------------------------------------------------------------- serve
    def serve(self, requests, realtime=True, clock=time.perf_counter):
        """Serve ``requests`` (sorted by arrival).  With ``realtime`` a
        request becomes visible at its arrival time on the wall clock;
        otherwise all are visible at once.  Returns {rid: Result}."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if r.footprint() > self.budget:
                raise ValueError(f"request {r.rid} needs {r.footprint()} B > budget {self.budget}")
        results = {}
        inflight = []
        free_slots = list(range(len(self.streams)))
        used = 0
        nxt = 0
        pending = []
        t0 = clock()
        while len(results) < len(reqs):
            now = clock() - t0
            while nxt < len(reqs) and (not realtime or reqs[nxt].arrival <= now):
                pending.append(reqs[nxt])
                nxt += 1
            # completions
            still = []
            for b in inflight:
                if b.event.query():
                    self._collect(b, results, t0, clock)
                    used -= b.bytes
                    free_slots.append(b.slot)
                else:
                    still.append(b)
            inflight = still
            # admission: the group of the oldest waiting request
            if pending and free_slots:
                key = pending[0].key()
                take, rest, nbytes = [], [], 0
                for r in pending:
                    fp = r.footprint()
                    if (r.key() == key and len(take) < self.max_batch_pairs
                            and used + nbytes + fp <= self.budget):
                        take.append(r)
                        nbytes += fp
                    else:
                        rest.append(r)
                pending = rest
                if take:
                    slot = free_slots.pop(0)
                    inflight.append(self._launch(take, slot))
                    used += nbytes
                    continue
            if not inflight and not pending and nxt < len(reqs):
                time.sleep(max(0.0, min(0.001, reqs[nxt].arrival - (clock() - t0))))
            elif inflight:
                time.sleep(0.0001)
        return results


def poisson_requests(seed, n, rate, tkf_frac=0.5, score_frac=0.5, median=250, sigma=0.5,
                     lo=20, hi=2000, t_range=(0.05, 2.0), A=20):
    """Synthetic request stream: Poisson arrivals at ``rate`` per second,
    lengths round(LogNormal(ln median, sigma)) clipped to [lo, hi] for both
    sequences (the descendant drawn around its ancestor's length), a
    ``tkf_frac`` share of TKF92 requests (the rest MixDom2) and a
    ``score_frac`` share of scoring requests (the rest E-steps)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n)
    arr = np.cumsum(gaps) - gaps[0]
    Lx = np.clip(np.round(np.exp(np.log(median) + sigma * rng.standard_normal(n))), lo, hi).astype(int)
    Ly = np.clip(np.round(Lx * np.exp(0.1 * rng.standard_normal(n))), lo, hi).astype(int)
    out = []
    for i in range(n):
        out.append(Request(i, "tkf92" if rng.random() < tkf_frac else "mixdom2",
                           "score" if rng.random() < score_frac else "estep",
                           rng.integers(0, A, Lx[i]), rng.integers(0, A, Ly[i]),
                           float(rng.uniform(*t_range)), float(arr[i])))
    return out


def summarize(requests, results, wall, S_of):
    lat = np.array([results[r.rid].latency for r in requests])
    cs = sum(r.cell_states(S_of[r.model]) for r in requests)
    return {"requests": len(requests), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "max_ms": float(lat.max() * 1e3),
            "wall_s": wall, "achieved_req_per_s": len(requests) / wall,
            "cell_states_per_s": cs / wall}
