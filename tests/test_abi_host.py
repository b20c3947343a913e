"""Host-side checks of the C ABI (no GPU calls): libtkfmi.so loads, exports
every entry point include/tkfmi.h declares, and its host-only paths (version,
workspace query, argument validation) behave as documented.  Also checks the
ragged packing and the geometric-bin helper of the Python mirror."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tkfmi.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tkfmi_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    decl = _declared()
    assert "tkfmi_fb2d_batched" in decl and "tkfmi_fb2d_workspace_bytes" in decl
    from tkfmi import _abi
    assert sorted(_abi.EXPORTED) == decl


def test_library_exports_every_declared_symbol():
    from tkfmi import _abi
    L = _abi.lib()
    for name in _declared():
        assert hasattr(L, name), name
    assert L.tkfmi_abi_version() == _abi.ABI_VERSION
    m = re.search(r"#define TKFMI_ABI_VERSION (\d+)", open(HEADER).read())
    assert int(m.group(1)) == _abi.ABI_VERSION


def test_struct_layout_matches_header():
    """Field order of the ctypes mirrors == the C structs."""
    from tkfmi import _abi
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for cname, py in (("tkfmi_problem", _abi.Problem), ("tkfmi_outputs", _abi.Outputs),
                      ("tkfmi_problem1d", _abi.Problem1D),
                      ("tkfmi_kernel_time", _abi.KernelTime),
                      ("tkfmi_mixdom_model", _abi.MixdomModel),
                      ("tkfmi_builder", _abi.Builder)):
        body = re.search(r"typedef struct %s \{(.*?)\}" % cname, text, flags=re.S).group(1)
        fields = []
        for decl in body.split(";"):
            parts = [x.strip() for x in decl.strip().split(",") if x.strip()]
            if not parts:
                continue
            fields.append(re.findall(r"(\w+)(?:\[\w+\])?$", parts[0])[0])
            fields += [re.sub(r"[\s*]|\[\w+\]", "", x) for x in parts[1:]]
        assert fields == [f[0] for f in py._fields_], cname
    flags = dict(re.findall(r"#define TKFMI_(\w+) (\d+)u", text))
    for k in ("FORWARD_ONLY", "WANT_POSTERIORS", "NO_COUNTS", "GENERIC_KERNEL", "CLASS_KERNEL", "REUSE_DESCRIPTORS", "WANT_EMIT_COUNTS",
              "TIME_KERNELS", "CHECKPOINT"):
        assert int(flags[k]) == getattr(_abi, k)


def _problem(st, sg, xl, yl, G=1, n_trans=1):
    from tkfmi import _abi
    B = len(xl)
    keep = dict(st=np.asarray(st, np.int32), sg=np.asarray(sg, np.int32),
                xl=np.asarray(xl, np.int32), yl=np.asarray(yl, np.int32),
                xo=np.zeros(B, np.int64), yo=np.zeros(B, np.int64))
    p = _abi.Problem(B, len(st), _abi._np_ptr(keep["st"]), _abi._np_ptr(keep["sg"]), G, n_trans,
                     None, 1, None, None, None, _abi._np_ptr(keep["xo"]), _abi._np_ptr(keep["yo"]),
                     _abi._np_ptr(keep["xl"]), _abi._np_ptr(keep["yl"]))
    return p, keep


def test_workspace_query_is_host_only_and_monotone():
    from tkfmi import _abi, mixdom
    L = _abi.lib()
    st = list(mixdom.state_types(3, 1))
    sg = [0, 0] + [(s - 2) // 5 for s in range(2, 17)]
    p_small, k1 = _problem(st, sg, [50, 80], [60, 70], G=3)
    p_big, k2 = _problem(st, sg, [500, 800], [600, 700], G=3)
    a = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p_small), 0)
    b = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p_big), 0)
    c = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p_big), _abi.FORWARD_ONLY)
    g = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p_big), _abi.GENERIC_KERNEL)
    assert 0 < a < b and 0 < c < b and g > 0


def test_checkpoint_shrinks_the_wide_workspace():
    """TKFMI_CHECKPOINT on the 162-state layout: the lattice part shrinks from
    (Ly+1) antidiagonal-row equivalents per pair to 2 (Lx+Ly)/K + K + 3 with
    K = ceil(sqrt(2 (Lx+Ly+1))); the cell-per-lane layouts ignore the flag."""
    from tkfmi import _abi, mixdom
    L = _abi.lib()
    st = list(mixdom.state_types(8, 4))
    p, k = _problem(st, [0] * len(st), [500, 800], [600, 700], G=1)
    _abi.set_checkpoint_budget(0)  # no automatic checkpointing
    full = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), _abi.CLASS_KERNEL)
    _abi.set_checkpoint_budget(_abi.CHECKPOINT_BUDGET_DEFAULT)
    ck = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), _abi.CHECKPOINT)
    fo = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), _abi.FORWARD_ONLY | _abi.CLASS_KERNEL)
    lat_full = sum((lx + 1) * (ly + 1) * 161 * 8 for lx, ly in ((500, 600), (800, 700)))
    assert abs((full - fo) - lat_full) < 4096
    for lx, ly in ((500, 600), (800, 700)):
        K = int(np.ceil(np.sqrt(2.0 * (lx + ly + 1))))
        lat_full -= ((lx + ly) // K + 1) * 2 * (lx + 1) * 161 * 8 + (K + 3) * (lx + 1) * 161 * 8
    assert abs((full - ck) - lat_full) < 4096 and ck < full / 4
    st17 = list(mixdom.state_types(3, 1))
    p17, k17 = _problem(st17, [0] * 17, [500, 800], [600, 700], G=1)
    assert L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p17), _abi.CHECKPOINT) == \
        L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p17), 0)


def test_banded_workspace_holds_lattice_boundaries_and_accumulators():
    """The banded 16..160-state kernel (the default when no pair is
    checkpointed): full minus forward-only = every band's lattice (stored
    by step: Ly + rows steps of 16 rows, NBP + 1 doubles per cell) plus one
    NBP x NBP count accumulator per pair of more than one 16-row band;
    forward-only still holds the band boundaries ((ceil((Lx+1)/16) - 1)
    (Ly+1)(NBP+1) doubles per pair)."""
    from tkfmi import _abi, mixdom
    L = _abi.lib()
    st = list(mixdom.state_types(8, 4))  # 160 emitting states: NBP = 160
    xl, yl = [15, 16, 200, 0], [30, 40, 150, 9]
    p, k = _problem(st, [0] * len(st), xl, yl, G=1)
    full = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), 0)
    fo = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), _abi.FORWARD_ONLY)
    lat = sum((((x + 16) // 16) * y + x + 1) * 16 * 161 * 8 for x, y in zip(xl, yl))
    acc = sum(160 * 160 * 8 for x in xl if x + 1 > 16)
    assert abs((full - fo) - lat - acc) < 4096
    bnd = sum(((x + 16) // 16 - 1) * (y + 1) * 161 * 8 for x, y in zip(xl, yl))
    assert fo > bnd and fo - bnd < 1 << 20
    # one workgroup per pair: the antidiagonal kernel's own layout
    assert L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), _abi.CLASS_KERNEL) != full


def test_banded_checkpoint_workspace_holds_slots_not_lattices():
    """TKFMI_CHECKPOINT on the banded kernel: no lattice is kept; the
    workspace holds the forward's boundary rows, a second set for the
    backward's pushes, the count accumulators and one lattice slot per
    workgroup the device holds at once (one band: (Ly + 16) steps of 16
    rows of NBP + 1 doubles; 256 slots when no device is visible, the
    MI355X's CU count).  A batch whose slots would save less than half of
    its lattices (one pair) goes to the antidiagonal kernel instead."""
    from tkfmi import _abi, mixdom
    L = _abi.lib()
    st = list(mixdom.state_types(8, 4))  # NBP = 160: one workgroup per CU
    n, x, y = 300, 200, 150
    p, k = _problem(st, [0] * len(st), [x] * n, [y] * n, G=1)
    ck = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), _abi.CHECKPOINT)
    fo = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), _abi.FORWARD_ONLY)
    full = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), 0)
    bnd = n * ((x + 16) // 16 - 1) * (y + 1) * 161 * 8
    slots = 256 * (y + 16) * 16 * 161 * 8
    acc = n * 160 * 160 * 8
    assert abs((ck - fo) - bnd - slots - acc) < 4096
    assert ck < full / 3
    p1, k1 = _problem(st, [0] * len(st), [x], [y], G=1)
    assert (L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p1), _abi.CHECKPOINT)
            == L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p1), _abi.CHECKPOINT | _abi.CLASS_KERNEL))


def test_checkpoint_budget_applies_per_pair():
    """Without TKFMI_CHECKPOINT, 162-state pairs whose lattice exceeds the
    budget (default 256 MiB) are checkpointed, smaller ones are not, and
    budget 0 turns it off.  Many such pairs: the banded kernel's
    recomputation slots, as with the flag.  One 3000 x 2500 pair (9.7 GB,
    which slots cannot halve): the antidiagonal kernel's checkpoints, as
    with the flag.  One 800 x 700 pair (0.9 GB, at most 16 budgets): its
    lattice is kept whole, as with budget 0.  A 200 x 150 pair (39 MB) is
    never checkpointed."""
    from tkfmi import _abi, mixdom
    L = _abi.lib()
    st = list(mixdom.state_types(8, 4))

    def ws(xl, yl, f=0):
        p, keep = _problem(st, [0] * len(st), xl, yl)  # keep: the arrays p points into
        return L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), f)

    many = ([600] * 300, [600] * 300)
    try:
        _abi.set_checkpoint_budget(_abi.CHECKPOINT_BUDGET_DEFAULT)
        auto = {"many": ws(*many), "huge": ws([3000], [2500]), "mid": ws([800], [700]),
                "small": ws([200], [150])}
        assert auto["many"] == ws(*many, _abi.CHECKPOINT)
        assert auto["huge"] == ws([3000], [2500], _abi.CHECKPOINT)
        assert auto["huge"] == ws([3000], [2500], _abi.CHECKPOINT | _abi.CLASS_KERNEL)
        assert auto["mid"] > ws([800], [700], _abi.CHECKPOINT)
        _abi.set_checkpoint_budget(0)
        assert ws(*many) > 5 * auto["many"]
        assert ws([3000], [2500]) > 4 * auto["huge"]
        assert ws([800], [700]) == auto["mid"]
        assert ws([200], [150]) == auto["small"]
    finally:
        _abi.set_checkpoint_budget(_abi.CHECKPOINT_BUDGET_DEFAULT)


def test_checkpoint_segment_sets_the_antidiagonal_kernels_k():
    """tkfmi_checkpoint_segment fixes K for the antidiagonal kernel's
    checkpoints: the workspace of a checkpointed 3000 x 2500 pair at 162
    states changes by exactly the restart states and segment rows that K
    sets ((Lx + Ly) / K + 1 checkpoints of two rows, K segment rows; a row is
    (Lx + 1)(NBP + 1) doubles); k = 0 restores the sqrt rule (K = 105 here);
    k in 1..3 is rejected with a message."""
    from tkfmi import _abi, mixdom
    L = _abi.lib()
    st = list(mixdom.state_types(8, 4))
    p, keep = _problem(st, [0] * len(st), [3000], [2500])
    flags = _abi.CHECKPOINT | _abi.CLASS_KERNEL
    row = 3001 * (160 + 1) * 8
    rows = lambda K: 2 * (5500 // K + 1) + K
    try:
        auto = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), flags)
        for k in (50, 400):
            _abi.set_checkpoint_segment(k)
            got = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), flags)
            assert abs((got - auto) - (rows(k) - rows(105)) * row) < 512, k
        _abi.set_checkpoint_segment(0)
        assert L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), flags) == auto
        assert L.tkfmi_checkpoint_segment(3) == -1  # TKFMI_EINVAL
        assert "at least 4" in L.tkfmi_last_error().decode()
    finally:
        _abi.set_checkpoint_segment(0)


def test_invalid_problems_rejected_with_message():
    from tkfmi import _abi
    L = _abi.lib()
    # no E state
    p, k = _problem([0, 1, 2, 3, 1], [0] * 5, [3], [4])
    assert L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), 0) == 0
    rc = L.tkfmi_fb2d_batched(ctypes.byref(p), ctypes.byref(_abi.Outputs()), None, 0, None, 0)
    assert rc == -1 and L.tkfmi_last_error().decode()
    # negative length
    p, k = _problem([0, 1, 2, 3, 4], [0] * 5, [-1], [4])
    rc = L.tkfmi_fb2d_batched(ctypes.byref(p), ctypes.byref(_abi.Outputs()), None, 0, None, 0)
    assert rc == -1
    # dense log emissions with emission counts: refused before any device work
    p, k = _problem([0, 1, 2, 3, 4], [0] * 5, [3], [4])
    p.log_emit = ctypes.c_void_p(8)  # never dereferenced: the call fails first
    rc = L.tkfmi_fb2d_batched(ctypes.byref(p), ctypes.byref(_abi.Outputs()), None, 0, None,
                              _abi.WANT_EMIT_COUNTS)
    assert rc == -1 and "TKFMI_WANT_EMIT_COUNTS" in L.tkfmi_last_error().decode()


def test_dense_problem_workspace_takes_generic_or_wide_kernel():
    """With dense emissions the layouts the cell-per-lane kernels would take
    (TKF, MixDom K*F <= 3) get the generic kernel's workspace, and the
    16..160-state layouts no emission-table staging."""
    from tkfmi import _abi, mixdom
    L = _abi.lib()
    st = list(mixdom.state_types(3, 1))
    p, k = _problem(st, [0] * 17, [100, 40], [90, 70], G=1)
    chained = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), 0)
    generic = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), _abi.GENERIC_KERNEL)
    p.log_emit = ctypes.c_void_p(8)
    assert L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), 0) == generic != chained
    st82 = list(mixdom.state_types(8, 2))
    p, k = _problem(st82, [0] * 82, [100, 40], [90, 70], G=8)
    table = L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), 0)
    p.log_emit = ctypes.c_void_p(8)
    assert 0 < L.tkfmi_fb2d_workspace_bytes(ctypes.byref(p), 0) < table


def _problem1d(S, lens, s_idx=0, e_idx=1, constrained=True, G=1, sg=None):
    from tkfmi import _abi
    B = len(lens)
    keep = dict(st=np.asarray([0, 4] + [1, 2, 3, 2, 3] * ((S - 2) // 5), np.int32)[:S],
                sg=np.asarray(sg if sg is not None else [0] * S, np.int32),
                ln=np.asarray(lens, np.int32), off=np.zeros(B, np.int64))
    if B:
        keep["off"][1:] = np.cumsum(keep["ln"][:-1])
    fake = ctypes.c_void_p(16)  # never dereferenced by the host-side checks
    p = _abi.Problem1D(B, S, s_idx, e_idx, 1, 0, fake, None if constrained else fake,
                       _abi._np_ptr(keep["st"]) if constrained else None,
                       _abi._np_ptr(keep["sg"]) if constrained else None, G, 1,
                       fake if constrained else None, fake, fake, fake,
                       _abi._np_ptr(keep["off"]), _abi._np_ptr(keep["ln"]))
    return p, keep


def test_fb1d_workspace_query_and_validation():
    from tkfmi import _abi
    L = _abi.lib()
    p_small, k1 = _problem1d(17, [50, 80])
    p_big, k2 = _problem1d(17, [500, 800])
    a = L.tkfmi_fb1d_workspace_bytes(ctypes.byref(p_small), 0)
    b = L.tkfmi_fb1d_workspace_bytes(ctypes.byref(p_big), 0)
    c = L.tkfmi_fb1d_workspace_bytes(ctypes.byref(p_big), _abi.FORWARD_ONLY)
    d = L.tkfmi_fb1d_workspace_bytes(ctypes.byref(p_big), _abi.WANT_EMIT_COUNTS)
    assert 0 < a < b and 0 < c < b < d
    dense, k3 = _problem1d(162, [10], constrained=False)
    assert L.tkfmi_fb1d_workspace_bytes(ctypes.byref(dense), 0) > 0
    for bad in (_problem1d(300, [4], constrained=False),          # S above 256
                _problem1d(17, [4], s_idx=17),                   # init state out of range
                _problem1d(17, [-2]),                            # negative length
                _problem1d(17, [4], G=1, sg=[0] * 16 + [3])):    # group out of range
        p, k = bad
        assert L.tkfmi_fb1d_workspace_bytes(ctypes.byref(p), 0) == 0
        rc = L.tkfmi_fb1d_batched(ctypes.byref(p), ctypes.byref(_abi.Outputs()), None, 0, None, 0)
        assert rc == -1 and L.tkfmi_last_error().decode()


def test_packed_columns_validate_alignments():
    import torch
    from tkfmi import _abi
    pc = _abi.PackedColumns([np.array([1, 2, 3]), np.array([], int)],
                            [np.array([5, 0, 25]), np.array([], int)],
                            [np.array([6, 7, 0]), np.array([], int)], torch.device("cpu"))
    assert list(pc.len) == [3, 0] and list(pc.off) == [0, 3] and pc.rows == 3
    assert pc.col.tolist()[:3] == [1, 2, 3] and pc.anc.tolist()[:3] == [5, 0, 20]
    with pytest.raises(ValueError):
        _abi.PackedColumns([np.array([1, 4])], [np.array([0, 0])], [np.array([0, 0])],
                           torch.device("cpu"))


def test_packed_pairs_and_code_normalisation():
    import torch
    from tkfmi import _abi
    pp = _abi.PackedPairs([np.array([1, 2, 25]), np.array([], int)],
                          [np.array([-1, 0]), np.array([4])], torch.device("cpu"))
    assert list(pp.x_len) == [3, 0] and list(pp.y_len) == [2, 1]
    assert list(pp.x_off) == [0, 3] and list(pp.y_off) == [0, 2]
    assert pp.x_cat.tolist()[:3] == [1, 2, 20]          # clamp to the wildcard
    assert pp.y_cat.tolist()[:2] == [20, 0]             # -1 wraps to index 20
    assert list(pp.cells) == [12, 2]


def test_pad_to_bin_matches_reference_bins(monkeypatch):
    """dp/hmm.py:46-91: [0,1,2,3] + strictly increasing round(4 * 1.2^k)
    while below 4096 (the last one overshoots to 4899), then powers of 2;
    TKFMIXDOM_MAX_PAD rounds up to multiples of 32, capped."""
    from tkfmi import hmm
    monkeypatch.delenv("TKFMIXDOM_MAX_PAD", raising=False)
    bins = hmm._GEOM_BINS
    assert bins[:8] == [0, 1, 2, 3, 4, 5, 6, 7] and bins[-1] == 4899 and len(bins) == 44
    assert all(b2 > b1 for b1, b2 in zip(bins, bins[1:]))
    assert [hmm._pad_to_bin(n) for n in (0, 1, 3, 8, 4096, 4899, 4900)] == [1, 1, 3, 8, 4899, 4899, 8192]
    assert hmm._pad_to_bin(200) in bins and 200 <= hmm._pad_to_bin(200) < 240
    monkeypatch.setenv("TKFMIXDOM_MAX_PAD", "100")
    assert hmm._pad_to_bin(70) == 96 and hmm._pad_to_bin(97) == 100 and hmm._pad_to_bin(0) == 32
    with pytest.raises(ValueError):
        hmm._pad_to_bin(101)


def test_chain_config_validates_arguments():
    from tkfmi import _abi, autotune
    L = _abi.lib()
    assert L.tkfmi_chain_config(5, 3, 0.3, 0) < 0
    assert L.tkfmi_chain_config(4, 1, 0.3, 0) < 0
    assert L.tkfmi_chain_config(4, 3, -1.0, 1) < 0
    assert b"tkfmi_chain_config" in L.tkfmi_last_error()
    assert L.tkfmi_chain_schedule(1.5, -1) < 0
    assert L.tkfmi_chain_schedule(-0.1, -1) < 0
    assert b"tkfmi_chain_schedule" in L.tkfmi_last_error()
    autotune.set_chain_config(*autotune.DEFAULT)


def test_gradient_builders_validate_arguments_on_the_host():
    """The gradient path's builder entry points reject bad sizes and missing
    pointers before any GPU call, with a message; the elimination's
    workspace size is nk^2 + nk nz + nz (nz + nk) doubles per pair."""
    from tkfmi import _abi, mixdom
    L = _abi.lib()
    P = ctypes.c_void_p
    fake = P(16)  # never dereferenced: validation fails first
    assert L.tkfmi_tkf_trans_jac(fake, fake, fake, None, 4, 3, fake, fake, None) == -1
    assert b"mode" in L.tkfmi_last_error()
    assert L.tkfmi_tkf_trans_jac(fake, fake, fake, None, 4, 2, fake, fake, None) == -1  # TKF92 needs ext
    assert b"pointer" in L.tkfmi_last_error()
    assert L.tkfmi_tkf_trans_jac(None, None, None, None, 0, 0, None, None, None) == 0  # n = 0: nothing to do
    assert L.tkfmi_expm_batched(fake, 3, 41, fake, None) == -1
    assert b"1..40" in L.tkfmi_last_error()
    assert L.tkfmi_expm_vjp_batched(fake, fake, 3, 21, fake, None) == -1
    assert b"1..20" in L.tkfmi_last_error()
    assert L.tkfmi_expm_batched(None, 0, 20, None, None) == 0
    pl = mixdom._plan(3, 1)
    nk, nz = len(pl.keep), len(pl.elim)
    assert (nk, nz, pl.N0) == (17, 60, 77)
    assert L.tkfmi_chi_workspace_bytes(200, nk, nz) == 200 * (nk * nk + nk * nz + nz * (nz + nk)) * 8
    idx = [P(16)] * 3
    args = [fake, 5, pl.n_atoms, idx[0], idx[1], idx[2], len(pl.flat), fake, fake, pl.N0, nk, nz]
    small = L.tkfmi_chi_workspace_bytes(5, nk, nz) - 8
    assert L.tkfmi_chi_forward(*args, fake, fake, fake, small, None) == -2
    assert b"workspace too small" in L.tkfmi_last_error()
    bad = list(args)
    bad[10], bad[11] = 70, 60  # nk + nz > N0
    assert L.tkfmi_chi_forward(*bad, fake, fake, fake, 1 << 30, None) == -1
    assert L.tkfmi_chi_backward(*args, None, fake, fake, fake, 1 << 30, None) == -1
    assert b"pointer" in L.tkfmi_last_error()
    # the plan's native view: every state kept or eliminated, maps inverse to the lists
    assert np.all((pl.kmap >= 0) != (pl.zmap >= 0))
    assert np.array_equal(pl.kmap[pl.keep], np.arange(nk)) and np.array_equal(pl.zmap[pl.elim], np.arange(nz))
    assert len(np.unique(pl.flat32)) == len(pl.flat32) and pl.flat32.max() < pl.N0 * pl.N0
