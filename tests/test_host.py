"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol its header declares, argument validation returns status codes
before touching a device, the Python layer keeps the reference's error
behaviour, and the delay-table helpers follow the tutorial's rules."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import beampower_amd as bp
from beampower_amd import core, network, synthetic

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "beamform_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))


def test_header_lists_expected_entry_points():
    names = declared_functions()
    assert set(core.GPU_SYMBOLS) == set(names), names


def test_library_exports_every_declared_symbol():
    lib = bp.load_library("gpu")
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.bp_version().decode().count(".") == 2


def test_every_declared_symbol_has_argument_types():
    """ctypes converts Python floats and structs only for declared argument
    types; an undeclared entry point would fail at its first call."""
    lib = bp.load_library("gpu")
    for name in core.GPU_SYMBOLS:
        assert getattr(lib, name).argtypes is not None, name


def test_library_is_gfx950_only():
    out = os.popen(f"/opt/rocm/llvm/bin/clang-offload-bundler --list --type=o "
                   f"--input={core.LIBRARIES['gpu']['filepath_library']} 2>/dev/null").read()
    if not out:
        pytest.skip("offload bundler listing unavailable")
    targets = [t for t in out.split() if "amdgcn" in t]
    assert targets and all("gfx950" in t for t in targets), targets


def test_load_library_errors_mirror_reference():
    with pytest.raises(NameError):
        bp.load_library("tpu")
    if not os.path.exists(core.LIBRARIES["cpu"]["filepath_library"]):
        with pytest.raises(OSError):
            bp.load_library("cpu")
    assert bp.load_library("GPU") is bp.load_library("gpu")


def test_bad_out_of_bounds_returns_none(capsys):
    w = synthetic.random_case(0, S=2, C=3, P=2, T=16, K=3)
    assert bp.beamform(w[0], w[2], w[1], w[3], device="gpu", out_of_bounds="loose") is None
    assert "strict" in capsys.readouterr().out


def test_bad_mode_returns_one(capsys):
    w = synthetic.random_case(0, S=2, C=3, P=2, T=16, K=3)
    assert bp.beamform(w[0], w[2], w[1], w[3], device="gpu", mode="sideways") == 1


def test_unknown_reduce_in_direct_mode_returns_one(capsys):
    """The reference's direct branch has no case for another reduce and
    falls through to its final else (beampower.py:225-227)."""
    w = synthetic.random_case(0, S=2, C=3, P=2, T=16, K=3)
    for mode in ("direct", "normal"):
        assert bp.beamform(w[0], w[2], w[1], w[3], device="gpu", mode=mode,
                           reduce="mean") == 1
        assert f"not {mode}." in capsys.readouterr().out


def test_differential_c_abi_checks_arguments_without_a_device():
    """beamform_differential (the CPU library's signature) validates before
    touching a device; shapes with nothing to compute leave `beam` as is."""
    lib = bp.load_library("gpu")
    P_F = ct.POINTER(ct.c_float)
    assert lib.beamform_differential(None, None, None, 10, 4, 2, 2, 1, None) == -22
    assert "NULL output" in core.last_error(lib)
    b = np.full(4, 3.0, np.float32)
    assert lib.beamform_differential(None, None, None, 10, 4, 2, 2, 1,
                                     b.ctypes.data_as(P_F)) == -22
    assert "NULL input" in core.last_error(lib)
    for T, K, S, P in ((0, 4, 2, 2), (10, 0, 2, 2), (10, 4, 0, 2), (10, 4, 2, 0)):
        assert lib.beamform_differential(None, None, None, T, K, S, P, 1,
                                         b.ctypes.data_as(P_F)) == 0
    assert np.all(b == 3.0)
    assert lib.beamform_differential(None, None, None, 1 << 40, 4, 2, 2, 1, None) == -22
    assert lib.bp_network_beamform_differential(None, None, None, None) == -22


def test_c_abi_rejects_bad_arguments_without_a_device():
    lib = bp.load_library("gpu")
    F = np.zeros(8, np.float32)
    tau = np.zeros(4, np.int32)
    beta = np.ones(2, np.float32)
    m = np.zeros(2, np.float32)
    a = np.zeros(2, np.int32)
    P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)
    rc = lib.beamform_max(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
                          beta.ctypes.data_as(P_F), 2, 2, 1, 2, 7, m.ctypes.data_as(P_F),
                          a.ctypes.data_as(P_I))
    assert rc == -22 and "out_of_bounds" in core.last_error(lib)
    rc = lib.beamform(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I), beta.ctypes.data_as(P_F),
                      2, 2, 1, 2, -1, m.ctypes.data_as(P_F))
    assert rc == -22
    rc = lib.beamform_max(None, None, None, 1 << 40, 2, 1, 2, 0, None, None)
    assert rc == -22 and "too large" in core.last_error(lib)
    out = ct.c_void_p()
    assert lib.bp_network_create(None, None, 0, 1, 1, ct.byref(out)) == -22
    assert lib.bp_batch_create(0, None, None, None, None, None, None, ct.byref(out)) == -22
    assert "n_windows > 0" in core.last_error(lib)
    assert lib.bp_batch_launch(None, None) == -22
    assert lib.bp_batch_tune(None, 1, None, 0, None, None, None) == -22
    assert lib.bp_batch_destroy(None) == 0
    assert lib.bp_detect_threshold(None, 10, 5, 0.66, 8.0, None, None) == -22
    n = ct.c_size_t(7)
    assert lib.bp_detect_peaks(None, None, 10, 1, None, 0, ct.byref(n), None) == -22
    assert lib.bp_network_create_ordered(tau.ctypes.data, beta.ctypes.data, 2, 1, 2, 7,
                                         ct.byref(out)) == -22
    assert "order" in core.last_error(lib)
    assert lib.bp_network_source_order(None, None, None) == -22


def test_network_order_names_are_checked():
    from beampower_amd import device as bpd
    assert bpd.ORDERS == {"natural": 0, "similarity": 1, "auto": 2}
    with pytest.raises(ValueError, match="order"):
        bpd.Network(np.zeros((2, 1, 2), np.int32), np.ones((2, 1), np.float32), order="sorted")


def test_degenerate_shapes_follow_reference_semantics():
    """No source or no station: strict keeps the caller's beam_max and
    writes argmax 0; flexible gives every source a 0 beam (beamform.c)."""
    lib = bp.load_library("gpu")
    P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)
    for K, S, oob, expect in ((0, 3, 0, 7.0), (0, 3, 1, 7.0), (4, 0, 0, 7.0), (4, 0, 1, 0.0)):
        m = np.full(5, 7.0, np.float32)
        a = np.full(5, 9, np.int32)
        rc = lib.beamform_max(None, None, None, 5, K, S, 2, oob, m.ctypes.data_as(P_F),
                              a.ctypes.data_as(P_I))
        assert rc == 0
        assert np.all(m == expect) and np.all(a == 0)


def test_fused_entry_points_check_arguments_without_a_device():
    """bp_beamform_max_fused / bp_beamform_fused refuse what beamform_max /
    beamform refuse, before touching a device, and keep the degenerate
    shapes' semantics."""
    lib = bp.load_library("gpu")
    P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)
    u = np.zeros(12, np.float32)
    alpha = np.ones(6, np.float32)
    tau = np.zeros(4, np.int32)
    beta = np.ones(2, np.float32)
    m = np.full(2, 7.0, np.float32)
    a = np.full(2, 9, np.int32)
    ins = (u.ctypes.data_as(P_F), alpha.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
           beta.ctypes.data_as(P_F))
    rc = lib.bp_beamform_max_fused(*ins, 2, 2, 1, 3, 2, 5, m.ctypes.data_as(P_F),
                                   a.ctypes.data_as(P_I))
    assert rc == -22 and "out_of_bounds" in core.last_error(lib)
    assert lib.bp_beamform_fused(*ins, 2, 2, 1, 3, 2, -1, m.ctypes.data_as(P_F)) == -22
    assert lib.bp_beamform_max_fused(None, None, None, None, 1 << 40, 2, 1, 3, 2, 0,
                                     None, None) == -22
    assert "too large" in core.last_error(lib)
    assert lib.bp_beamform_max_fused(None, None, None, None, 5, 2, 1, 1 << 40, 2, 0,
                                     None, None) == -22
    assert "n_channels" in core.last_error(lib)
    # channels present but no phase weights
    rc = lib.bp_beamform_max_fused(ins[0], None, ins[2], ins[3], 2, 2, 1, 3, 2, 0,
                                   m.ctypes.data_as(P_F), a.ctypes.data_as(P_I))
    assert rc == -22 and "NULL input" in core.last_error(lib)
    for K, S, oob, expect in ((0, 3, 0, 7.0), (4, 0, 1, 0.0)):
        m = np.full(5, 7.0, np.float32)
        a = np.full(5, 9, np.int32)
        assert lib.bp_beamform_max_fused(None, None, None, None, 5, K, S, 3, 2, oob,
                                         m.ctypes.data_as(P_F), a.ctypes.data_as(P_I)) == 0
        assert np.all(m == expect) and np.all(a == 0)


def test_wrapper_passes_library_typed_inputs_without_copies():
    from beampower_amd import beampower as bpw
    x = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    y = bpw._flat(x, np.float32)
    assert np.shares_memory(x, y) and y.shape == (24,)
    for z, dt in ((np.asfortranarray(x), np.float32), (x.astype(np.float64), np.float32),
                  (x[:, :, ::2], np.float32), (x * 1.7, np.int32)):
        got = bpw._flat(z, dt)
        want = z.flatten().astype(dt)
        assert got.dtype == want.dtype and np.array_equal(got, want)
    u = np.zeros((5, 3, 100), np.float32)
    tau = np.zeros((7, 5, 2), np.int32)
    assert bpw._fusable(u, tau, np.zeros((5, 3, 2)), np.ones((7, 5)))
    assert not bpw._fusable(u, tau, np.zeros((5, 3, 1)), np.ones((7, 5)))
    assert not bpw._fusable(u, tau, np.zeros((5, 2, 2)), np.ones((7, 5)))
    assert not bpw._fusable(u, tau, np.zeros((5, 3, 2)), np.ones((7, 4)))
    assert not bpw._fusable(u, np.zeros((7, 4, 2), np.int32), np.zeros((5, 3, 2)),
                            np.ones((7, 4)))


def test_sec_to_samp_follows_tutorial():
    # notebooks/3_beampower.ipynb:380-392: sign(t) * int(|t*sr| + 0.2)
    t = np.array([0.0, 0.03, 0.039, -0.039, 1.0, 2.99, -2.99, 0.0319])
    got = network.sec_to_samp(t, 25.0)
    expect = np.array([int(np.sign(x)) * int(abs(x * 25.0) + 0.2) for x in t])
    assert got.dtype == np.int32 and np.array_equal(got, expect)


def test_relative_moveouts_and_plain_geometry():
    w = synthetic.plain_c1()
    tau = w.time_delays
    assert tau.shape == (4000, 20, 2) and tau.dtype == np.int32
    assert np.all(tau.min(axis=(1, 2)) == 0)
    # S slower than P at every station
    assert np.all(tau[:, :, 1] >= tau[:, :, 0])
    # geometric upper bound: box diagonal + depth at Vs, at 25 Hz
    assert tau.max() <= int(np.sqrt(50 ** 2 * 2 + 20 ** 2) / 3.5 * 25) + 1
    assert w.waveform_features.shape == (20, 3, 6000)
    assert w.weights_phases[:, 0, 0].min() == 1 and w.weights_phases[:, 1:, 1].max() == 0.5


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asking for more exits with a message."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_tuner_candidates_and_build_flags():
    """The tuner's first candidate is the sources' default shape, every
    candidate differs, and build() turns a recorded shape into -D flags."""
    from beampower_amd import build as bp_build, tune
    src = ""
    for f in ("bp_kernels.h", "beamform_gpu.hip"):
        with open(os.path.join(bp_build.CSRC, f)) as fh:
            src += fh.read()
    for m in bp_build.SHAPE_MACROS:
        assert f"#ifndef {m}\n#define {m} " in src, m  # every tunable macro has a default
    defaults = {m: int(v) for m, v in re.findall(r"#define (BP_[A-Z0-9_]+) (\d+)\b", src)
                if m in bp_build.SHAPE_MACROS}
    for cands in tune.CANDIDATE_SETS.values():
        first = cands[0]
        assert all(defaults[m] == v for m, v in first.items())
        keys = [tuple(sorted(c.items())) for c in cands]
        assert len(set(keys)) == len(keys)
        for c in cands:  # register-resident accumulators, at most 16 waves per work-group
            kw, nacc = c.get("BP_KW", defaults["BP_KW"]), c.get("BP_NACC", defaults["BP_NACC"])
            w = c.get("BP_W", 8)
            assert kw * nacc <= 64 and w <= 16, c
    assert tune.name_of({"BP_KW": 6, "BP_W": 12}) == "kw=6,w=12"
    flags = bp_build.shape_flags({"BP_KW": 5, "BP_PIPE_DEPTH": 2})
    assert flags == ["-DBP_KW=5", "-DBP_PIPE_DEPTH=2"]
    assert all(f in bp_build.command("x.so", flags) for f in flags)
    with pytest.raises(ValueError):
        bp_build.shape_flags({"BP_ANYTHING": 1})


def test_reference_import_name():
    """`import beampower` exposes the reference's names and module layout
    (beampower/__init__.py:9-14, beampower/beampower.py, beampower/core.py),
    served by this engine."""
    import beampower
    import beampower_amd
    assert beampower.__all__ == ["load_library", "beamform"]
    assert beampower.beamform is beampower_amd.beamform
    assert beampower.beampower.beamform is beampower_amd.beampower.beamform
    assert beampower.beampower.prestack_traces is beampower_amd.beampower.prestack_traces
    assert beampower.load_library is beampower_amd.load_library
    assert beampower.core.load_library("gpu") is beampower_amd.load_library("gpu")
    assert beampower.core.LIBRARIES is beampower_amd.core.LIBRARIES
    w = synthetic.random_case(0, S=2, C=3, P=2, T=16, K=3)
    assert beampower.beamform(w[0], w[2], w[1], w[3], device="gpu", mode="sideways") == 1
