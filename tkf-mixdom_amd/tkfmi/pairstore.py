"""Precompiled X/A/Y training pairs, decoded on the device.

Reference: python/tkfmixdom/jax/util/pair_format.py (the record format and
its host codec) and python/train_pfam.py:984-1075 (PrecompiledPairSource,
which reads JSONL shards of records and decodes every pair on the host).

A record is {'x': ancestor string, 'a': run-length alignment "M5I1M3D2",
'y': descendant string, 't': time, 'fam': family, 'id': pair id} over the
20 amino acids "ACDEFGHIKLMNPQRSTVWY".  The host codec below keeps the
reference's function names and results (rle_encode, rle_decode,
encode_pair, decode_pair, validate_record).  :class:`PairStore` is the
device-resident form: the host only joins the record strings into three
byte buffers (their lengths give the residue counts) and uploads them once;
libtkfmi's tkfmi_pairs_decode turns them, in one launch, into the residue
codes the 2D forward/backward reads (a PairStore can be passed wherever a
PackedPairs is) and the alignment columns the 1D alignment-constrained path
reads (:meth:`PairStore.columns`).
"""
import hashlib
import io
import itertools
import json
import re

import numpy as np
import torch

from . import _abi, hmm

AMINO_ACIDS = "ACDEFGHIKLMNPQRSTVWY"
AA_TO_IDX = {a: i for i, a in enumerate(AMINO_ACIDS)}
M_STATE, I_STATE, D_STATE = 1, 2, 3
_RUN_LETTER = {M_STATE: "M", I_STATE: "I", D_STATE: "D"}
_LETTER_RUN = {v: k for k, v in _RUN_LETTER.items()}
_RUN_RE = re.compile(r"([MID])(\d+)")


# ---------------------------------------------------------------- host codec
def rle_encode(states):
    """'M3I1M1D2' from the state codes (1=M, 2=I, 3=D) of the alignment
    columns (pair_format.py:38-66)."""
    return "".join(f"{_RUN_LETTER[int(s)]}{len(list(g))}" for s, g in itertools.groupby(states))


def rle_decode(rle_string):
    """State codes of 'M3I1M1D2'; text that is not a letter M/I/D followed by
    digits is skipped (pair_format.py:69-90)."""
    out = []
    for letter, count in _RUN_RE.findall(rle_string or ""):
        out += [_LETTER_RUN[letter]] * int(count)
    return out


def encode_pair(anc_seq, desc_seq, alignment_states, t_est, family_id, row1_name="", row2_name="",
                gamma_labels=None, gamma_log_posteriors=None):
    """A record for an aligned pair (pair_format.py:93-137): the id is the
    first 16 hex digits of sha256('family:row1:row2'), t is rounded to 6
    decimals, gamma annotations are optional."""
    rec = {"x": "".join(AMINO_ACIDS[int(c)] for c in anc_seq),
           "a": rle_encode(alignment_states),
           "y": "".join(AMINO_ACIDS[int(c)] for c in desc_seq),
           "t": round(float(t_est), 6),
           "fam": family_id,
           "id": hashlib.sha256(f"{family_id}:{row1_name}:{row2_name}".encode()).hexdigest()[:16]}
    if gamma_labels is not None:
        rec["gamma"] = [int(g) for g in gamma_labels]
    if gamma_log_posteriors is not None:
        rec["gamma_lp"] = [[round(float(v), 4) for v in row] for row in gamma_log_posteriors]
    return rec


def decode_pair(record):
    """(x_int, y_int, states, anc_chars, desc_chars, t_est, gamma, gamma_lp)
    of a record, on the host (pair_format.py:140-190)."""
    x = np.array([AA_TO_IDX[c] for c in record["x"]], dtype=np.int32)
    y = np.array([AA_TO_IDX[c] for c in record["y"]], dtype=np.int32)
    states = rle_decode(record["a"])
    st = np.asarray(states, dtype=np.int64)
    anc = x[:int(np.isin(st, (M_STATE, D_STATE)).sum())].tolist()
    desc = y[:int(np.isin(st, (M_STATE, I_STATE)).sum())].tolist()
    return (x, y, states, anc, desc, float(record["t"]), record.get("gamma"), record.get("gamma_lp"))


def validate_record(record):
    """(ok, message): M+D run lengths = |X|, M+I = |Y|, only the 20 amino
    acids (pair_format.py:193-223)."""
    try:
        st = np.asarray(rle_decode(record["a"]), dtype=np.int64)
        nm, ni, nd = (int((st == s).sum()) for s in (M_STATE, I_STATE, D_STATE))
        if nm + nd != len(record["x"]):
            return False, f"M+D={nm + nd} != len(X)={len(record['x'])}"
        if nm + ni != len(record["y"]):
            return False, f"M+I={nm + ni} != len(Y)={len(record['y'])}"
        for name in ("x", "y"):
            bad = set(record[name]) - set(AMINO_ACIDS)
            if bad:
                return False, f"Invalid AA in {name.upper()}: {sorted(bad)[0]}"
        return True, ""
    except Exception as e:  # malformed record
        return False, str(e)


def read_jsonl(path):
    """Records of a JSONL shard (.jsonl, or .jsonl.zst as precompile_pairs.py
    writes them when the zstandard module is importable)."""
    with open(path, "rb") as f:
        data = f.read()
    if path.endswith(".zst"):
        try:
            import zstandard
        except ImportError as e:
            raise ImportError("reading .zst shards needs the zstandard module") from e
        data = zstandard.ZstdDecompressor().decompress(data)
    return [json.loads(line) for line in io.StringIO(data.decode("utf-8")) if line.strip()]


# -------------------------------------------------------------- device store
_STATUS_TEXT = ((1, "invalid character in X"), (2, "invalid character in Y"),
                (4, "M+D run lengths != len(X)"), (8, "M+I run lengths != len(Y)"),
                (16, "run length above 2^24"))


class _Columns:
    """The store's alignment columns in the form the 1D runner reads
    (PackedColumns' attributes): per pair ``len`` columns at ``off``."""

    def __init__(self, store):
        self.B = store.B
        self.len = store.n_cols
        self.off = store.col_off
        self.col, self.anc, self.desc = store.col_type, store.anc, store.desc
        self.rows = int(self.len.sum())


class PairStore:
    """B precompiled pairs resident on the device.

    Attributes (those of a PackedPairs, so a store goes straight into the
    2D batched calls): B, x_cat / y_cat (device uint8 codes), x_off, y_off,
    x_len, y_len, cells (host).  Also t (host FP64 times), fam, ids, and,
    with ``columns=True``, the decoded alignment columns (col_type, anc,
    desc on the device; n_cols, col_off on the host).

    Invalid records (validate_record's checks) raise ValueError, as
    decode_pair would fail on them."""

    def __init__(self, records, device=None, columns=True):
        dev = device or hmm.device()
        self.dev = dev
        self.B = B = len(records)
        xs = [r["x"] for r in records]
        ys = [r["y"] for r in records]
        as_ = [r["a"] for r in records]
        self.t = np.array([float(r.get("t", 0.0)) for r in records], dtype=np.float64)
        self.fam = [r.get("fam", "unknown") for r in records]
        self.ids = [r.get("id", "") for r in records]
        self.x_len = np.array([len(x) for x in xs], dtype=np.int32)
        self.y_len = np.array([len(y) for y in ys], dtype=np.int32)
        a_len = np.array([len(a) for a in as_], dtype=np.int32)
        off = lambda n: np.concatenate([[0], np.cumsum(n[:-1], dtype=np.int64)]).astype(np.int64) \
            if len(n) else np.zeros(0, np.int64)  # noqa: E731
        self.x_off, self.y_off, a_off = off(self.x_len), off(self.y_len), off(a_len)
        self.cells = (self.x_len.astype(np.int64) + 1) * (self.y_len.astype(np.int64) + 1)
        cap = self.x_len.astype(np.int64) + self.y_len
        self.col_off = off(cap)
        up = lambda strs: torch.from_numpy(  # noqa: E731
            np.frombuffer(("".join(strs) + "\0").encode("ascii", errors="replace"), dtype=np.uint8).copy()).to(dev)
        x_text, y_text, a_text = up(xs), up(ys), up(as_)
        self.x_cat = torch.empty(int(self.x_len.sum()) + 1, dtype=torch.uint8, device=dev)
        self.y_cat = torch.empty(int(self.y_len.sum()) + 1, dtype=torch.uint8, device=dev)
        ncap = int(cap.sum()) + 1
        mk = lambda: torch.zeros(ncap, dtype=torch.uint8, device=dev) if columns else None  # noqa: E731
        self.col_type, self.anc, self.desc = mk(), mk(), mk()
        ncols = torch.zeros(max(B, 1), dtype=torch.int32, device=dev)
        status = torch.zeros(max(B, 1), dtype=torch.int32, device=dev)
        np_ptr = _abi._np_ptr
        txt = _abi.PairText(B, 0, _abi._ptr(x_text), _abi._ptr(y_text), _abi._ptr(a_text),
                            np_ptr(self.x_off), np_ptr(self.x_len), np_ptr(self.y_off), np_ptr(self.y_len),
                            np_ptr(a_off), np_ptr(a_len))
        out = _abi.PairColumns(_abi._ptr(self.x_cat), _abi._ptr(self.y_cat), _abi._ptr(self.col_type),
                               _abi._ptr(self.anc), _abi._ptr(self.desc), _abi._ptr(ncols),
                               _abi._ptr(status))
        L = _abi.lib()
        wsb = int(L.tkfmi_pairs_decode_workspace_bytes(ctypes_byref(txt)))
        if wsb == 0:
            raise ValueError("invalid pair text")
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream(dev)
        with torch.cuda.stream(stream):
            self.x_cat[-1:].zero_()
            self.y_cat[-1:].zero_()
        _abi.check(L.tkfmi_pairs_decode(ctypes_byref(txt), ctypes_byref(out), _abi._ptr(ws),
                                        _abi.ctypes.c_size_t(wsb), _abi.ctypes.c_void_p(stream.cuda_stream)))
        self.status = status[:B].cpu().numpy()       # waits for the decode
        self.n_cols = ncols[:B].cpu().numpy()
        bad = np.nonzero(self.status)[0]
        if len(bad):
            b = int(bad[0])
            why = "; ".join(t for bit, t in _STATUS_TEXT if self.status[b] & bit)
            raise ValueError(f"{len(bad)} invalid record(s); first: #{b} (id {self.ids[b]!r}): {why}")
        self._keep = (x_text, y_text, a_text, ws)

    def __len__(self):
        return self.B

    @classmethod
    def from_jsonl(cls, paths, device=None, columns=True, max_alignment_len=None):
        """All records of one or more JSONL shards (PrecompiledPairSource's
        _decode_all, train_pfam.py:1017-1038: records with more alignment
        columns than max_alignment_len are dropped)."""
        if isinstance(paths, str):
            paths = [paths]
        recs = [r for p in paths for r in read_jsonl(p)]
        if max_alignment_len:
            recs = [r for r in recs if sum(int(c) for c in re.findall(r"[MID](\d+)", r["a"])) <= max_alignment_len]
        return cls(recs, device=device, columns=columns)

    def columns(self):
        """The alignment columns for tkfmi.hmm1d / tkfmi._abi.Runner1D."""
        if self.col_type is None:
            raise ValueError("the store was decoded without columns")
        return _Columns(self)

    def decoded(self, b):
        """Pair b as decode_pair returns it, read back from the device:
        (x_int, y_int, states, anc_chars, desc_chars, t_est)."""
        x = self.x_cat[self.x_off[b]:self.x_off[b] + self.x_len[b]].cpu().numpy().astype(np.int32)
        y = self.y_cat[self.y_off[b]:self.y_off[b] + self.y_len[b]].cpu().numpy().astype(np.int32)
        if self.col_type is None:
            return x, y, None, None, None, float(self.t[b])
        sl = slice(int(self.col_off[b]), int(self.col_off[b]) + int(self.n_cols[b]))
        st = self.col_type[sl].cpu().numpy().astype(np.int64)
        an = self.anc[sl].cpu().numpy().astype(np.int64)
        de = self.desc[sl].cpu().numpy().astype(np.int64)
        return (x, y, st.tolist(), an[np.isin(st, (M_STATE, D_STATE))].tolist(),
                de[np.isin(st, (M_STATE, I_STATE))].tolist(), float(self.t[b]))


def ctypes_byref(x):
    return _abi.ctypes.byref(x)
