"""The FP32 drop-in at its one-GPU shape: 50 stations, 192,000 sources,
90,000 samples (1 h at 25 Hz), through the host-pointer C ABI.

* reduce='max' over all sources and samples, checked bit for bit against the
  oracle on halo-padded time slices: a slice [t0, t0 + W + halo) with
  halo = max delay + 1 gives every source exactly the validity and the beam
  it has at t0 + t in the full call for t < W (delays are relative, >= 0);
* reduce='none' on a seeded 4,096-source subset, the full (4096, 90000)
  beams against the oracle;
* the size-independent link between the two: the maximum over all sources is
  at least every subset beam, with equality where the argmax is a subset
  source;
* the injected events are found.
"""
import ctypes as ct

import numpy as np
import pytest

from conftest import bits_equal
from oracle import oracle
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu

P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)


@pytest.fixture(scope="module")
def c2(gpu_lib):
    w = synthetic.dropin_c2()
    F = oracle.prestack(w.waveform_features, w.weights_phases, num_threads=16)
    return w, F


@pytest.fixture(scope="module")
def c2_max(gpu_lib, c2):
    """reduce='max' over all 192,000 sources through the drop-in C ABI."""
    w, F = c2
    S, T, P = F.shape
    tau, beta = w.time_delays, w.weights_sources
    m = np.zeros(T, np.float32)
    a = np.zeros(T, np.int32)
    rc = gpu_lib.beamform_max(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
                              beta.ctypes.data_as(P_F), T, tau.shape[0], S, P, 0,
                              m.ctypes.data_as(P_F), a.ctypes.data_as(P_I))
    assert rc == 0
    return m, a


def test_c2_max_matches_oracle_on_slices(c2, c2_max):
    w, F = c2
    m, a = c2_max
    S, T, P = F.shape
    tau, beta = w.time_delays, w.weights_sources
    halo = int(tau.max()) + 1
    W = 160
    rng = np.random.default_rng(22)
    starts = [0, T - W - halo] + [int(x) for x in rng.integers(1, T - W - halo, 2)]
    for t0 in starts:
        Fs = np.ascontiguousarray(F[:, t0:t0 + W + halo, :])
        mo, ao = oracle.beamform_max(Fs, tau, beta, 0, num_threads=16)
        assert bits_equal(m[t0:t0 + W], mo[:W]), t0
        assert np.array_equal(a[t0:t0 + W], ao[:W]), t0
    # past T - (least max delay) no source is valid: the caller's zeros, argmax 0
    min_hi = int(oracle.delay_bounds(tau, beta, num_threads=16)[:, 1].min())
    assert np.all(m[T - min_hi:] == 0.0) and np.all(a[T - min_hi:] == 0)
    found = sum(int(np.array_equal(tau[a[t_e]], tau[k_e])) for k_e, t_e in w.events)
    assert found >= len(w.events) - 1, (found, len(w.events))


def test_c2_flexible_max_matches_oracle_on_slices(gpu_lib, c2):
    """Flexible bounds: with delays >= 0 a halo-padded slice drops exactly
    the terms the full call drops (none before the last halo of samples)."""
    w, F = c2
    S, T, P = F.shape
    tau, beta = w.time_delays, w.weights_sources
    m = np.zeros(T, np.float32)
    a = np.zeros(T, np.int32)
    rc = gpu_lib.beamform_max(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
                              beta.ctypes.data_as(P_F), T, tau.shape[0], S, P, 1,
                              m.ctypes.data_as(P_F), a.ctypes.data_as(P_I))
    assert rc == 0
    halo = int(tau.max()) + 1
    W = 160
    for t0 in (0, T - W - halo):
        Fs = np.ascontiguousarray(F[:, t0:t0 + W + halo, :])
        mo, ao = oracle.beamform_max(Fs, tau, beta, 1, num_threads=16)
        assert bits_equal(m[t0:t0 + W], mo[:W]), t0
        assert np.array_equal(a[t0:t0 + W], ao[:W]), t0
    # the last samples: every source is a candidate, with its in-range terms
    Fs = np.ascontiguousarray(F[:, T - W:, :])
    mo, ao = oracle.beamform_max(Fs, tau, beta, 1, num_threads=16)
    assert bits_equal(m[T - W:], mo) and np.array_equal(a[T - W:], ao)


def test_c2_none_subset_matches_oracle(gpu_lib, c2, c2_max):
    w, F = c2
    m, a = c2_max
    S, T, P = F.shape
    sub = synthetic.subset_sources(w.time_delays.shape[0], 4096, seed=2)
    tau = np.ascontiguousarray(w.time_delays[sub])
    beta = np.ascontiguousarray(w.weights_sources[sub])
    K = len(sub)
    b = np.zeros((K, T), np.float32)
    rc = gpu_lib.beamform(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
                          beta.ctypes.data_as(P_F), T, K, S, P, 0, b.ctypes.data_as(P_F))
    assert rc == 0
    bo = oracle.beamform(F, tau, beta, 0, num_threads=16)
    assert bits_equal(b, bo)
    del bo
    # against the maximum over all sources
    assert np.all(b.max(axis=0) <= m)
    pos = np.searchsorted(sub, a)
    hit = (pos < K) & (sub[np.minimum(pos, K - 1)] == a)
    t_hit = np.flatnonzero(hit)
    assert t_hit.size > 0
    assert bits_equal(b[pos[t_hit], t_hit], m[t_hit])


def test_c2_prestack_on_gpu(gpu_lib, c2):
    """The CPU library's prestack signature on the GPU at this shape: the
    (50, 3, 90000) features go up and the (50, 90000, 2) traces come back
    through pinned staging buffers; bit-identical to the oracle's chain."""
    import beampower_amd as bp
    w, F = c2
    Fg = bp.prestack_traces(w.waveform_features, w.weights_phases, device="gpu")
    assert bits_equal(Fg, F)
