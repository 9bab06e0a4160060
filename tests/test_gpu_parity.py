"""GPU parity: the gfx950 library against the reference's golden vectors and
the CPU oracle, through the C ABI (drop-in symbols) and the Python API.

Bar: bit-identical beams, maxima and argmax to the reference's pip/CMake
build (FMA-contracted), which the oracle's "fma" flavour reproduces.
"""
import ctypes as ct

import numpy as np
import pytest

from conftest import GOLDEN_CASES, bits_equal, load_golden
from oracle import oracle
import beampower_amd as bp
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu

P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)


def abi_max(lib, F, tau, beta, oob, init=None):
    F = np.ascontiguousarray(F, np.float32)
    tau = np.ascontiguousarray(tau, np.int32)
    beta = np.ascontiguousarray(beta, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    m = np.zeros(T, np.float32) if init is None else init.astype(np.float32).copy()
    a = np.full(T, -5, np.int32)
    rc = lib.beamform_max(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
                          beta.ctypes.data_as(P_F), T, K, S, P, oob, m.ctypes.data_as(P_F),
                          a.ctypes.data_as(P_I))
    assert rc == 0, bp.core.last_error(lib)
    return m, a


def abi_none(lib, F, tau, beta, oob):
    F = np.ascontiguousarray(F, np.float32)
    tau = np.ascontiguousarray(tau, np.int32)
    beta = np.ascontiguousarray(beta, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    b = np.full((K, T), 3.0, np.float32)
    rc = lib.beamform(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I), beta.ctypes.data_as(P_F),
                      T, K, S, P, oob, b.ctypes.data_as(P_F))
    assert rc == 0, bp.core.last_error(lib)
    return b


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_golden_through_c_abi(gpu_lib, case):
    g = load_golden(case)
    F = g["F_cmake"]
    for oob in (0, 1):
        if f"max_{oob}_cmake" in g:
            m, a = abi_max(gpu_lib, F, g["tau"], g["beta"], oob)
            assert bits_equal(m, g[f"max_{oob}_cmake"]), (case, oob)
            assert np.array_equal(a, g[f"argmax_{oob}_cmake"]), (case, oob)
        if f"none_{oob}_cmake" in g:
            b = abi_none(gpu_lib, F, g["tau"], g["beta"], oob)
            assert bits_equal(b, g[f"none_{oob}_cmake"]), (case, oob)


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_golden_prestack_on_gpu(gpu_lib, case):
    g = load_golden(case)
    F = bp.prestack_traces(g["u"], g["alpha"], device="gpu")
    assert bits_equal(F, g["F_cmake"])


def test_golden_through_python_api(gpu_lib):
    g = load_golden("random_small")
    for oob, name in ((0, "strict"), (1, "flexible")):
        m, a = bp.beamform(g["u"], g["tau"], g["alpha"], g["beta"], device="gpu",
                           reduce="max", out_of_bounds=name)
        assert bits_equal(m, g[f"max_{oob}_cmake"]) and np.array_equal(a, g[f"argmax_{oob}_cmake"])
        b = bp.beamform(g["u"], g["tau"], g["alpha"], g["beta"], device="gpu", reduce="none",
                        out_of_bounds=name)
        assert bits_equal(b, g[f"none_{oob}_cmake"])


def test_plain_full_known_answer(gpu_lib):
    """Plain-method workload (20 stations, 4,000 sources, 6,000 samples):
    bit-identical to the reference's golden output, events recovered."""
    g = load_golden("plain_full")
    w = synthetic.plain_c1()
    m, a = bp.beamform(w.waveform_features, w.time_delays, w.weights_phases,
                       w.weights_sources, device="gpu", reduce="max")
    assert bits_equal(m, g["max_0_cmake"])
    assert np.array_equal(a, g["argmax_0_cmake"])
    for k_e, t_e in w.events:
        assert np.array_equal(w.time_delays[a[t_e]], w.time_delays[k_e])


def test_caller_beam_max_kept_where_no_source_is_valid(gpu_lib):
    g = load_golden("random_small")
    init = np.full(g["F_cmake"].shape[1], 42.0, np.float32)
    m, a = abi_max(gpu_lib, g["F_cmake"], g["tau"], g["beta"], 0, init=init)
    mo, ao = oracle.beamform_max(g["F_cmake"], g["tau"], g["beta"], 0, init=init)
    assert bits_equal(m, mo) and np.array_equal(a, ao)
    assert np.any(m == 42.0)


@pytest.mark.parametrize("seed,S,P,T,K,tmax,oob", [
    (31, 20, 2, 3000, 2500, 500, 0),
    (32, 20, 2, 3000, 2500, 500, 1),
    (33, 13, 1, 2000, 1000, 300, 0),
    (34, 9, 3, 1500, 700, 200, 1),      # three phases
    (38, 6, 5, 1200, 300, 150, 0),      # five phases -> direct kernel
    (35, 10, 2, 6000, 400, 4000, 0),    # spread beyond the LDS window -> direct kernel
    (36, 10, 2, 6000, 400, 4000, 1),
    (37, 50, 2, 1100, 130, 1030, 1),    # windows cross both ends
])
def test_random_vs_oracle(gpu_lib, seed, S, P, T, K, tmax, oob):
    u, alpha, tau, beta = synthetic.random_case(seed, S=S, C=3, P=P, T=T, K=K,
                                                tau_range=(0, tmax),
                                                negative_delays=bool(oob))
    F = oracle.prestack(u, alpha)
    m, a = abi_max(gpu_lib, F, tau, beta, oob)
    mo, ao = oracle.beamform_max(F, tau, beta, oob)
    assert bits_equal(m, mo) and np.array_equal(a, ao)
    b = abi_none(gpu_lib, F, tau[: min(K, 300)], beta[: min(K, 300)], oob)
    bo = oracle.beamform(F, tau[: min(K, 300)], beta[: min(K, 300)], oob)
    assert bits_equal(b, bo)


def test_nonfinite_weights_in_flexible_edges(gpu_lib):
    """An infinite weight must not turn a dropped out-of-range term into NaN."""
    u, alpha, tau, beta = synthetic.random_case(41, S=6, C=3, P=2, T=700, K=200,
                                                tau_range=(0, 300), negative_delays=True)
    beta[3, 2] = np.inf
    beta[100, 0] = -np.inf
    beta[150, 4] = np.nan
    F = oracle.prestack(u, alpha)
    m, a = abi_max(gpu_lib, F, tau, beta, 1)
    mo, ao = oracle.beamform_max(F, tau, beta, 1)
    assert bits_equal(m, mo) and np.array_equal(a, ao)
    b = abi_none(gpu_lib, F, tau, beta, 1)
    assert bits_equal(b, oracle.beamform(F, tau, beta, 1))


def test_all_tile_classes_vs_oracle(gpu_lib):
    """Moveout spreads that put source tiles in each kernel class: small LDS
    window, large LDS window and direct gathers."""
    from beampower_amd import device as bpd
    rng = np.random.default_rng(51)
    S, C, P, T, K = 12, 3, 2, 4600, 64 * 60
    u = rng.standard_normal((S, C, T)).astype(np.float32)
    alpha = rng.uniform(-1, 1, (S, C, P)).astype(np.float32)
    tau = rng.integers(0, 60, (K, S, P)).astype(np.int32)
    tau[64 * 5:64 * 6] = rng.integers(0, 1400, (64, S, P))    # large window
    tau[64 * 9:64 * 10] = rng.integers(0, 1900, (64, S, P))   # large window
    tau[64 * 20:64 * 21] = rng.integers(0, 4000, (64, S, P))  # direct
    beta = rng.uniform(0.5, 1.5, (K, S)).astype(np.float32)
    beta[rng.random((K, S)) < 0.1] = 0.0
    net = bpd.Network(tau, beta)
    n = (ct.c_int * 3)()
    lw = (ct.c_int * 3)()
    assert gpu_lib.bp_network_classes(ct.c_void_p(net._h.value), n, lw) == 0
    assert n[0] > 0 and n[1] > 0 and n[2] > 0, list(n)
    F = oracle.prestack(u, alpha)
    for oob in (0, 1):
        m, a = abi_max(gpu_lib, F, tau, beta, oob)
        mo, ao = oracle.beamform_max(F, tau, beta, oob)
        assert bits_equal(m, mo) and np.array_equal(a, ao), oob
        b = abi_none(gpu_lib, F, tau, beta, oob)
        assert bits_equal(b, oracle.beamform(F, tau, beta, oob)), oob


def test_device_api_matches_drop_in(gpu_lib):
    """Device-resident Network path equals the host drop-in on the same data,
    including reduce='none' over a source sub-range."""
    import torch
    from beampower_amd import device as bpd
    w = synthetic.plain_c1()
    F = oracle.prestack(w.waveform_features, w.weights_phases)
    m0, a0 = abi_max(gpu_lib, F, w.time_delays, w.weights_sources, 0)
    net = bpd.Network(w.time_delays, w.weights_sources)
    u = torch.from_numpy(w.waveform_features).cuda()
    Fd = bpd.prestack_planar(u, torch.from_numpy(w.weights_phases).cuda())
    m, a = net.beamform_max(Fd, 0)
    torch.cuda.synchronize()
    assert bits_equal(m.cpu().numpy(), m0) and np.array_equal(a.cpu().numpy(), a0)
    b = net.beamform(Fd, 1, source_begin=1000, source_end=1300).cpu().numpy()
    bo = oracle.beamform(F, w.time_delays[1000:1300], w.weights_sources[1000:1300], 1)
    assert bits_equal(b, bo)


@pytest.mark.parametrize("P", [3, 4])
def test_three_and_four_phases_in_every_tile_class(gpu_lib, P):
    """The LDS kernel takes one to four phases; with more phases its largest
    window shrinks to fit a CU's LDS, so tile spreads put tiles in the small
    window, the large window and the direct kernel."""
    from beampower_amd import device as bpd
    rng = np.random.default_rng(60 + P)
    S, C, T, K = 9, 3, 4000, 48 * 40
    u = rng.standard_normal((S, C, T)).astype(np.float32)
    alpha = rng.uniform(-1, 1, (S, C, P)).astype(np.float32)
    tau = rng.integers(0, 50, (K, S, P)).astype(np.int32)
    tau[48 * 7:48 * 9] = rng.integers(0, 600, (96, S, P))      # large window
    tau[48 * 20:48 * 21] = rng.integers(0, 2500, (48, S, P))   # direct
    beta = rng.uniform(0.5, 1.5, (K, S)).astype(np.float32)
    beta[rng.random((K, S)) < 0.1] = 0.0
    net = bpd.Network(tau, beta, order="natural")
    cls = net.info()["tiles_per_class"]
    assert cls[0] > 0 and cls[1] > 0 and cls[2] > 0, cls
    F = oracle.prestack(u, alpha)
    for oob in (0, 1):
        m, a = abi_max(gpu_lib, F, tau, beta, oob)
        mo, ao = oracle.beamform_max(F, tau, beta, oob)
        assert bits_equal(m, mo) and np.array_equal(a, ao), (P, oob)
        b = abi_none(gpu_lib, F, tau, beta, oob)
        assert bits_equal(b, oracle.beamform(F, tau, beta, oob)), (P, oob)


@pytest.mark.parametrize("zero", ["some_tiles", "every_tile"])
def test_tiles_whose_sources_use_no_station(gpu_lib, zero):
    """Whole source tiles with all-zero weights (no source of the tile uses
    a station): they read no feature, yet still stage a window; its start
    must stay inside the rows.  Bit-identical to the oracle through the C
    ABI and a resident network in the caller's order (tiles of consecutive
    sources), with padded rows (fast staging) and unpadded ones (guarded
    staging); in bfloat16 storage the rows of those sources are 0."""
    import torch
    from beampower_amd import device as bpd
    from beampower_amd import load_library
    tile_c = ct.c_int()
    assert load_library("gpu").bp_kernel_shape(ct.byref(tile_c), None, None) == 0
    tile = tile_c.value
    u, alpha, tau, beta = synthetic.random_case(61, S=9, C=2, P=2, T=1300, K=8 * tile,
                                                tau_range=(0, 60))
    zero_rows = slice(tile, 4 * tile) if zero == "some_tiles" else slice(None)
    beta[zero_rows] = 0.0
    F = oracle.prestack(u, alpha)
    K, T = tau.shape[0], F.shape[1]
    net = bpd.Network(tau, beta, order="natural")
    Fd = torch.from_numpy(np.ascontiguousarray(F.transpose(0, 2, 1))).cuda()
    for oob in (0, 1):
        mo, ao = oracle.beamform_max(F, tau, beta, oob)
        m, a = abi_max(gpu_lib, F, tau, beta, oob)
        assert bits_equal(m, mo) and np.array_equal(a, ao), oob
        bo = oracle.beamform(F, tau, beta, oob)
        assert bits_equal(abi_none(gpu_lib, F, tau, beta, oob), bo), oob
        for pad in (net.padding(), (0, 0)):
            V = bpd.pack(Fd, bpd.BP_F32, *pad)
            m, a = net.beamform_max(V, oob)
            assert bits_equal(m.cpu().numpy(), mo), (oob, pad)
            assert np.array_equal(a.cpu().numpy(), ao), (oob, pad)
            b = net.beamform(V, oob).cpu().numpy()
            assert bits_equal(b, bo), (oob, pad)
    for pad in (net.padding(), (0, 0)):
        V = bpd.pack(Fd, bpd.BP_BF16, *pad)
        b = net.beamform(V, 1).cpu().numpy()
        assert not np.any(b[zero_rows]), pad
        net.beamform_max(V, 1)
    torch.cuda.synchronize()
