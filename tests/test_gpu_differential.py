"""mode='differential' on the GPU (beamform.c:338-394) against the reference's
golden vectors, the CPU oracle and the direct beams.

Bar: bit-identical beams to the reference's pip/CMake build (FMA-contracted)
and to the oracle's "fma" flavour; sources out of range at the centre sample
keep the caller's value.  At the one-GPU shape (1,225 station pairs,
192,000 sources) the check is against the oracle in full and against the
size-independent property that a differential beam is the direct strict
beam at the centre sample (T - 1) // 2.
"""
import ctypes as ct

import numpy as np
import pytest

from conftest import bits_equal, load_golden
from oracle import oracle
import beampower_amd as bp
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu

P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)
DIFF_CASES = ("differential_odd", "differential_even", "differential_nonfinite")


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def abi_differential(lib, F, tau, beta, init):
    F = np.ascontiguousarray(F, np.float32)
    tau = np.ascontiguousarray(tau, np.int32)
    beta = np.ascontiguousarray(beta, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    b = np.ascontiguousarray(init, np.float32).copy()
    rc = lib.beamform_differential(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
                                   beta.ctypes.data_as(P_F), T, K, S, P, 4,
                                   b.ctypes.data_as(P_F))
    assert rc == 0, bp.core.last_error(lib)
    return b


@pytest.mark.parametrize("case", DIFF_CASES)
def test_golden_through_c_abi(gpu_lib, case):
    g = load_golden(case)
    F, tau, beta = g["F_cmake"], g["tau"], g["beta"]
    K = tau.shape[0]
    assert bits_equal(abi_differential(gpu_lib, F, tau, beta, np.zeros(K)), g["diff_cmake"])
    # sources left out keep the caller's values
    init = np.arange(K, dtype=np.float32) + 0.5
    got = abi_differential(gpu_lib, F, tau, beta, init)
    assert bits_equal(got, oracle.beamform_differential(F, tau, beta, init=init))


@pytest.mark.parametrize("case", DIFF_CASES)
def test_golden_through_python_api(gpu_lib, case):
    g = load_golden(case)
    args = (g["u"], g["tau"], g["alpha"], g["beta"])
    b = bp.beamform(*args, device="gpu", mode="differential", reduce="none")
    assert bits_equal(b, g["diff_cmake"])
    m, a = bp.beamform(*args, device="gpu", mode="differential", reduce="max")
    assert bits_equal(np.asarray(m), np.max(g["diff_cmake"], axis=0))
    assert a == np.argmax(g["diff_cmake"], axis=0)


@pytest.mark.parametrize("seed,S,P,T,K", [(31, 7, 1, 1, 50), (32, 5, 2, 2, 77), (33, 33, 1, 64, 700),
                                          (34, 11, 3, 101, 1000), (35, 4, 7, 40, 513),
                                          (36, 17, 2, 257, 4096), (37, 9, 4, 50, 300),
                                          (38, 3, 5, 31, 129)])
def test_random_shapes_vs_oracle(gpu_lib, seed, S, P, T, K):
    """One to seven phases (the 1-4 phase kernels load 16 / P stations at
    a time, others one station at a time), station counts off those chunks,
    source counts off the 64-source wave blocks."""
    half = T // 2 + 3
    u, alpha, tau, beta = synthetic.random_case(seed, S=S, C=2, P=P, T=T, K=K,
                                                tau_range=(-half, half))
    F = oracle.prestack(u, alpha)
    init = np.full(K, -9.0, np.float32)
    got = abi_differential(gpu_lib, F, tau, beta, init)
    assert bits_equal(got, oracle.beamform_differential(F, tau, beta, init=init))


@pytest.mark.parametrize("P,T", [(2, 1001), (2, 32768), (3, 1001)])
def test_host_call_with_large_tables(gpu_lib, P, T):
    """Delay tables of 32 MiB or more through the host-pointer call: with
    two phases and windows up to 32,768 samples the host threads that stage
    the upload pack the delays into saturated int16 pairs; otherwise the
    int32 delays go up through the pinned stage.  Delays at and beyond the
    int16 range (which saturate, and then still point outside the window)
    are planted.  Bit-identical to the oracle."""
    K, S = 30_000, 150
    rng = np.random.default_rng(40 + P + T)
    F = rng.standard_normal((S, T, P)).astype(np.float32)
    tau = rng.integers(-(T // 2), T // 2 + 1, size=(K, S, P)).astype(np.int32)
    plant = rng.random(tau.shape) < 0.01
    tau[plant] = rng.choice(np.array([-40_000, -32_769, -32_768, -32_767, 32_767, 32_768,
                                      40_000, 2**31 - 1, -2**31], np.int64),
                            size=int(plant.sum())).astype(np.int32)
    beta = rng.uniform(-1.0, 2.0, (K, S)).astype(np.float32)
    beta[rng.random((K, S)) < 0.1] = 0.0
    assert tau.nbytes >= 32 << 20
    init = np.full(K, -3.0, np.float32)
    got = abi_differential(gpu_lib, F, tau, beta, init)
    ref = oracle.beamform_differential(F, tau, beta, init=init, num_threads=8)
    assert bits_equal(got, ref)
    assert np.sum(got != init) > 100  # sources that stay in range are written


@pytest.mark.parametrize("kind", ["ones", "mask01", "const", "zeros", "nan", "late_second_value"])
def test_host_call_with_uniform_weights(gpu_lib, kind):
    """Weights of 32 MiB or more with one nonzero value go up as the
    used-bit table that the host threads build (the device's
    diff_weight_bits_kernel / diff_mask_kernel words); any second value,
    even in the last row, sends the float32 weights instead.  -0.0 counts
    as zero, NaN as a value.  Bit-identical to the oracle."""
    K, S, P, T = 60_000, 150, 2, 1001
    rng = np.random.default_rng(77)
    F = rng.standard_normal((S, T, P)).astype(np.float32)
    tau = rng.integers(-(T // 2), T // 2 + 1, size=(K, S, P)).astype(np.int32)
    beta = np.ones((K, S), np.float32)
    if kind == "mask01":
        beta[rng.random((K, S)) < 0.3] = 0.0
        beta[rng.random((K, S)) < 0.05] = -0.0
    elif kind == "const":
        beta[:] = 0.37
        beta[rng.random((K, S)) < 0.2] = -0.0
    elif kind == "zeros":
        beta[:] = 0.0
    elif kind == "nan":
        beta[:] = np.nan
        beta[rng.random((K, S)) < 0.5] = 0.0
    elif kind == "late_second_value":
        beta[-1, -1] = 2.0
    assert beta.nbytes >= 32 << 20
    init = np.full(K, -3.0, np.float32)
    got = abi_differential(gpu_lib, F, tau, beta, init)
    ref = oracle.beamform_differential(F, tau, beta, init=init, num_threads=8)
    assert bits_equal(got, ref)


def test_c5_device_api_vs_oracle_and_event(gpu_lib):
    """20 stations (190 pairs), 4,000 sources: the device-resident network,
    float32 planar features, bit-exact; the injected source is the argmax."""
    torch = _torch()
    from beampower_amd import device as bpd
    w = synthetic.differential_c5()
    F = oracle.prestack(w.waveform_features, w.weights_phases)
    ref = oracle.beamform_differential(F, w.time_delays, w.weights_sources)
    net = bpd.Network(w.time_delays, w.weights_sources)
    Fd = torch.from_numpy(np.ascontiguousarray(F.transpose(0, 2, 1))).cuda()
    b = net.beamform_differential(Fd).cpu().numpy()
    assert bits_equal(b, ref)
    assert int(np.argmax(b)) == w.events[0][0]


@pytest.mark.parametrize("fmt", ["f32", "bf16", "f16"])
def test_views_equal_direct_beam_at_centre(gpu_lib, fmt):
    """Every storage format and padded rows: the differential beam is the
    direct strict beam of the same view at the centre sample, and sources
    out of range are left untouched in both."""
    torch = _torch()
    from beampower_amd import device as bpd
    w = synthetic.differential_c5()
    net = bpd.Network(w.time_delays, w.weights_sources)
    Fd = bpd.prestack_planar(torch.from_numpy(w.waveform_features).cuda(),
                             torch.from_numpy(w.weights_phases).cuda())
    f = {"f32": bpd.BP_F32, "bf16": bpd.BP_BF16, "f16": bpd.BP_F16}[fmt]
    V = bpd.pack(Fd, f, 5, 8)
    K, T = w.time_delays.shape[0], w.waveform_features.shape[2]
    init = torch.full((K,), -3.0, device="cuda")
    b = net.beamform_differential(V, beam=init.clone()).cpu().numpy()
    direct = torch.full((K, T), -3.0, device="cuda")
    net.beamform(V, 0, beam=direct)
    col = direct[:, (T - 1) // 2].contiguous().cpu().numpy()
    assert bits_equal(b, col)
    if fmt == "f32":
        F = oracle.prestack(w.waveform_features, w.weights_phases)
        assert bits_equal(b, oracle.beamform_differential(F, w.time_delays, w.weights_sources,
                                                          init=np.full(K, -3.0, np.float32)))


def test_c6_full_size_vs_oracle_and_direct(gpu_lib):
    """50 stations (1,225 pairs), 192,000 sources, T = 1,581: bit-exact
    against the oracle, and equal to the direct strict beams at the centre
    sample, computed 16,000 sources at a time."""
    torch = _torch()
    from beampower_amd import device as bpd
    w = synthetic.differential_c6()
    F = oracle.prestack(w.waveform_features, w.weights_phases, num_threads=8)
    ref = oracle.beamform_differential(F, w.time_delays, w.weights_sources, num_threads=8)
    net = bpd.Network(w.time_delays, w.weights_sources)
    Fd = torch.from_numpy(np.ascontiguousarray(F.transpose(0, 2, 1))).cuda()
    b = net.beamform_differential(Fd)
    torch.cuda.synchronize()
    bh = b.cpu().numpy()
    assert bits_equal(bh, ref)
    assert int(np.argmax(bh)) == w.events[0][0]
    K, T = w.time_delays.shape[0], F.shape[1]
    t = (T - 1) // 2
    for k0 in range(0, K, 16000):
        k1 = min(K, k0 + 16000)
        d = net.beamform(Fd, 0, source_begin=k0, source_end=k1)[:, t].contiguous()
        assert bits_equal(d.cpu().numpy(), bh[k0:k1]), k0


@pytest.mark.parametrize("T", [32767, 32768, 40001])
def test_delay_table_widths_around_the_int16_limit(gpu_lib, T):
    """Two-phase delays go in paired int16 tables for windows of at most
    32,768 samples (saturated delays stay outside the window) and in int32
    tables for longer windows.  Delays beyond the int16 range, and sources
    with only in-range delays, both give the oracle's beams; one network
    serves a short and a long window in turn."""
    torch = _torch()
    from beampower_amd import device as bpd
    rng = np.random.default_rng(T)
    S, P, K = 5, 2, 300
    F = rng.standard_normal((S, T, P)).astype(np.float32)
    c = (T - 1) // 2
    tau = rng.integers(-c, T - 1 - c + 1, (K, S, P)).astype(np.int32)   # in range
    tau[:40] = rng.integers(-70000, 70000, (40, S, P))                  # mostly out of range
    tau[40:50, 0, 0] = 32767 + rng.integers(0, 3, 10)                   # around the int16 edge
    tau[50:60, 1, 1] = -32768 - rng.integers(0, 3, 10)
    beta = rng.uniform(0.5, 1.5, (K, S)).astype(np.float32)
    beta[rng.random((K, S)) < 0.2] = 0.0
    ref = oracle.beamform_differential(F, tau, beta, flavour="fma")
    assert bits_equal(abi_differential(gpu_lib, F, tau, beta, np.zeros(K)), ref)
    net = bpd.Network(tau, beta)
    for Fw, refw in ((F[:, :2001], None), (F, ref)):
        if refw is None:
            refw = oracle.beamform_differential(np.ascontiguousarray(Fw), tau, beta,
                                                flavour="fma")
        Fd = torch.from_numpy(np.ascontiguousarray(Fw.transpose(0, 2, 1))).cuda()
        assert bits_equal(net.beamform_differential(Fd).cpu().numpy(), refw)


def weights_of_kind(kind, rng, K, S):
    """Source weights whose nonzero entries share one value (the used-bit
    table) or not (the float32 table)."""
    if kind == "ones":
        return np.ones((K, S), np.float32)
    if kind == "all_zero":
        return np.zeros((K, S), np.float32)
    beta = np.full((K, S), 1.0 if kind == "mask01" else 0.37, np.float32)
    beta[rng.random((K, S)) < 0.3] = 0.0
    beta[rng.random((K, S)) < 0.05] = -0.0  # negative zero: unused
    beta[1, :] = 0.0                        # a source with no used station
    if kind == "two_values":
        beta[3, 5] = 2.0
    return beta


FORM_OF_KIND = {"ones": "used_bits", "mask01": "used_bits", "const": "used_bits",
                "all_zero": "used_bits", "two_values": "float32"}


@pytest.mark.parametrize("kind", sorted(FORM_OF_KIND))
@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_uniform_weights_as_used_bits(gpu_lib, kind, P):
    """Weights whose nonzero entries share one value are read as one bit per
    (source, station): bit-identical to the oracle through the C ABI and a
    resident network, equal to the direct strict beams at the centre sample
    in the 16-bit formats, and the weight table is 1/8 of the float32 one
    or less (70 stations: three-phase chunks of 5 stations straddle 32-bit
    words)."""
    torch = _torch()
    from beampower_amd import device as bpd
    S, K, T = 70, 700, 61
    u, alpha, tau, _ = synthetic.random_case(40 + P, S=S, C=2, P=P, T=T, K=K,
                                             negative_delays=True)
    beta = weights_of_kind(kind, np.random.default_rng(P), K, S)
    F = oracle.prestack(u, alpha, flavour="fma")
    init = np.full(K, -7.0, np.float32)
    ref = oracle.beamform_differential(F, tau, beta, init=init, flavour="fma")
    assert bits_equal(abi_differential(gpu_lib, F, tau, beta, init), ref)
    net = bpd.Network(tau, beta)
    Fd = torch.from_numpy(np.ascontiguousarray(F.transpose(0, 2, 1))).cuda()
    b = net.beamform_differential(Fd, beam=torch.from_numpy(init).cuda())
    assert bits_equal(b.cpu().numpy(), ref)
    info = net.diff_tables()
    assert info["weight_form"] == FORM_OF_KIND[kind]
    blocks = (K + 63) // 64
    if info["weight_form"] == "used_bits":
        assert 0 < info["weight_bytes"] * 8 <= blocks * 64 * 4 * S
    else:
        assert info["weight_bytes"] >= blocks * 64 * 4 * S
    for f in (bpd.BP_BF16, bpd.BP_F16):
        V = bpd.pack(Fd, f, 3, 4)
        bv = net.beamform_differential(V, beam=torch.full((K,), -7.0, device="cuda"))
        direct = torch.full((K, T), -7.0, device="cuda")
        net.beamform(V, 0, beam=direct)
        col = direct[:, (T - 1) // 2].contiguous().cpu().numpy()
        assert bits_equal(bv.cpu().numpy(), col), f


@pytest.mark.parametrize("layout", ["scattered", "smooth"])
def test_unused_entries_contribute_nothing(gpu_lib, layout):
    """Unused stations whose features are NaN or infinite and whose delays
    point anywhere in the int32 range add nothing, signed zeros included:
    beams whose every used product underflows end at -0.0, and sources
    with one used delay out of range keep the caller's value.  "smooth":
    delays that change by at most one sample per source, so that a wave's
    64 sources mostly fall in one 64-sample window per station and phase
    (gathered by one load and a cross-lane permute), with a few stations
    spread wider (gathered lane by lane), mixed within the chunks."""
    rng = np.random.default_rng(77)
    S, P, T, K = 13, 2, 301, 1100
    F = np.full((S, T, P), 1e-30, np.float32)
    F[:, ::7, :] = 0.5
    unused_rows = rng.random(S) < 0.4
    F[unused_rows] = np.nan
    F[unused_rows, ::3] = np.inf
    if layout == "scattered":
        tau = rng.integers(-T // 2 + 1, T // 2, size=(K, S, P)).astype(np.int32)
    else:
        slope = rng.choice([0, 1, 1, 3], size=(S, P))  # 3: wider than a window
        start = rng.integers(-60, 40, size=(S, P))
        tau = (start[None] + slope[None] * (np.arange(K) % 64)[:, None, None]).astype(np.int32)
    beta = np.where(rng.random((K, S)) < 0.7, 1.0, 0.0).astype(np.float32)
    beta[:, unused_rows] = np.where(rng.random((K, unused_rows.sum())) < 0.5, 0.0, -0.0)
    tau[:, unused_rows] = rng.integers(-2**31, 2**31 - 1, size=(K, unused_rows.sum(), P),
                                       dtype=np.int64).astype(np.int32)
    tiny = rng.random(K) < 0.3  # products of -1e-30 and 1e-30 underflow to -0.0
    beta[tiny] = np.where(beta[tiny] != 0, -1e-30, beta[tiny]).astype(np.float32)
    tau[tiny] = (tau[tiny] // 7) * 7 + 1  # off the 0.5 samples (t = 150 = 7 * 21 + 3)
    oob = rng.random(K) < 0.1
    for k in np.flatnonzero(oob):
        s = int(rng.choice(np.flatnonzero(~unused_rows)))
        beta[k, s] = 1.0
        tau[k, s, 1] = T  # out of range at the centre sample
    init = np.full(K, 7.0, np.float32)
    ref = oracle.beamform_differential(F, tau, beta, init=init)
    assert np.signbit(ref[tiny & ~oob & (ref == 0)]).any()  # -0.0 beams present
    assert np.isfinite(ref).all()
    got = abi_differential(gpu_lib, F, tau, beta, init)
    assert bits_equal(got, ref)
