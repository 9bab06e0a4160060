"""Host-pointer calls on long windows, through the pinned staging copies.

beamform_max (prestacked features) and the fused call (raw features) on
windows of 140,000-150,001 samples, 33-40 MB of features, whose uploads
go through the pinned staging buffers and the host worker pool, compared
bit for bit with the oracle over the whole window (beamform.c:230-336):
strict and flexible bounds, negative delays, moveouts up to 6,000 samples,
a source tile that only the direct kernel takes, 600 feature rows of a
16,500-sample window, and the kept network (built, re-planned, reused).

Bar: bit-identical maxima and argmax.
"""
import ctypes as ct

import numpy as np
import pytest

from conftest import bits_equal
from oracle import oracle
import beampower_amd as bp

pytestmark = pytest.mark.gpu

P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)


def _case(seed, S, C, P, T, K, lo, hi, wide_tile=False):
    """Features, phase weights, moveouts in [lo, hi] smooth across sources
    (so tiles of 48 stay in the LDS classes), weights with unused stations;
    wide_tile puts sources 48-95 at spreads only the direct kernel takes."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((S, C, T)).astype(np.float32)
    alpha = rng.uniform(-1, 1, (S, C, P)).astype(np.float32)
    base = np.sort(rng.integers(lo, hi + 1, K))[:, None, None]
    jitter = rng.integers(0, max(1, (hi - lo) // 8), (K, S, P))
    tau = np.clip(base + jitter, lo, hi).astype(np.int32)
    if wide_tile:
        tau[48:96] = rng.integers(lo, hi + 5000, (48, S, P))
    beta = rng.uniform(0.2, 1.0, (K, S)).astype(np.float32)
    beta[rng.random((K, S)) < 0.1] = 0.0
    return u, alpha, tau, beta


def fused_max(lib, u, alpha, tau, beta, oob, init):
    u = np.ascontiguousarray(u, np.float32)
    alpha = np.ascontiguousarray(alpha, np.float32)
    S, C, T = u.shape
    K, _, P = tau.shape
    m = init.astype(np.float32).copy()
    a = np.full(T, -5, np.int32)
    rc = lib.bp_beamform_max_fused(u.ctypes.data_as(P_F), alpha.ctypes.data_as(P_F),
                                   tau.ctypes.data_as(P_I), beta.ctypes.data_as(P_F), T, K, S,
                                   C, P, oob, m.ctypes.data_as(P_F), a.ctypes.data_as(P_I))
    assert rc == 0, bp.core.last_error(lib)
    return m, a


def stacked_max(lib, F, tau, beta, oob, init):
    F = np.ascontiguousarray(F, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    m = init.astype(np.float32).copy()
    a = np.full(T, -5, np.int32)
    rc = lib.beamform_max(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
                          beta.ctypes.data_as(P_F), T, K, S, P, oob, m.ctypes.data_as(P_F),
                          a.ctypes.data_as(P_I))
    assert rc == 0, bp.core.last_error(lib)
    return m, a


# seed, S, C, P, T, K, delays [lo, hi], oob, a direct-kernel tile
CASES = [
    (81, 30, 2, 2, 150000, 400, 0, 1000, 0, False),
    (82, 30, 2, 2, 150001, 300, -600, 900, 1, True),    # flexible, negative delays
    (83, 30, 2, 2, 140000, 200, 0, 6000, 0, False),     # long moveouts
    (84, 100, 6, 1, 16500, 300, 0, 300, 0, False),      # many rows, a short window
]


@pytest.mark.parametrize("seed,S,C,P,T,K,lo,hi,oob,wide", CASES)
def test_long_window_calls_match_the_oracle(gpu_lib, seed, S, C, P, T, K, lo, hi, oob, wide):
    print(f"long window case seed={seed} S={S} C={C} P={P} T={T} K={K}", flush=True)
    u, alpha, tau, beta = _case(seed, S, C, P, T, K, lo, hi, wide)
    assert u.nbytes >= 32 << 20  # staged through the pinned buffers
    F = oracle.prestack(u, alpha, flavour="fma")
    init = np.full(T, 11.0, np.float32)
    mo, ao = oracle.beamform_max(F, tau, beta, oob, init=init, flavour="fma")
    gpu_lib.bp_dropin_cache_clear()
    for i in range(3):  # built and kept, re-planned, reused
        m, a = fused_max(gpu_lib, u, alpha, tau, beta, oob, init)
        assert bits_equal(m, mo) and np.array_equal(a, ao), (seed, "fused", i)
    m, a = stacked_max(gpu_lib, F, tau, beta, oob, init)
    assert bits_equal(m, mo) and np.array_equal(a, ao), (seed, "prestacked")
    gpu_lib.bp_dropin_cache_clear()
