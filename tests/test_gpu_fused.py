"""The fused host-pointer entry points (bp_beamform_max_fused,
bp_beamform_fused): raw (S, C, T) features and (S, C, P) phase weights,
prestacked on the device, against the reference's golden outputs, the
oracle, and the two calls the reference wrapper makes
(prestack_waveform_features then beamform_max / beamform,
beampower.py:103-105,156-190).

Bar: bit-identical beams, maxima and argmax.
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


def _ins(u, alpha, tau, beta):
    u = np.ascontiguousarray(u, np.float32)
    alpha = np.ascontiguousarray(alpha, np.float32)
    tau = np.ascontiguousarray(tau, np.int32)
    beta = np.ascontiguousarray(beta, np.float32)
    S, C, T = u.shape
    K, _, P = tau.shape
    keep = (u, alpha, tau, beta)
    ptrs = (u.ctypes.data_as(P_F), alpha.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
            beta.ctypes.data_as(P_F))
    return keep, ptrs, (T, K, S, C, P)


def fused_max(lib, u, alpha, tau, beta, oob, init=None):
    keep, ptrs, (T, K, S, C, P) = _ins(u, alpha, tau, beta)
    m = np.zeros(T, np.float32) if init is None else init.astype(np.float32).copy()
    a = np.full(T, -5, np.int32)
    rc = lib.bp_beamform_max_fused(*ptrs, T, K, S, C, P, oob, m.ctypes.data_as(P_F),
                                   a.ctypes.data_as(P_I))
    assert rc == 0, bp.core.last_error(lib)
    del keep
    return m, a


def fused_none(lib, u, alpha, tau, beta, oob):
    keep, ptrs, (T, K, S, C, P) = _ins(u, alpha, tau, beta)
    b = np.full((K, T), 3.0, np.float32)
    rc = lib.bp_beamform_fused(*ptrs, T, K, S, C, P, oob, b.ctypes.data_as(P_F))
    assert rc == 0, bp.core.last_error(lib)
    del keep
    return b


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_fused_golden(gpu_lib, case):
    g = load_golden(case)
    for oob in (0, 1):
        if f"max_{oob}_cmake" in g:
            m, a = fused_max(gpu_lib, g["u"], g["alpha"], g["tau"], g["beta"], oob)
            assert bits_equal(m, g[f"max_{oob}_cmake"]), (case, oob)
            assert np.array_equal(a, g[f"argmax_{oob}_cmake"]), (case, oob)
        if f"none_{oob}_cmake" in g:
            b = fused_none(gpu_lib, g["u"], g["alpha"], g["tau"], g["beta"], oob)
            assert bits_equal(b, g[f"none_{oob}_cmake"]), (case, oob)


@pytest.mark.parametrize("seed,S,C,P,T,K,tmax,oob", [
    (61, 20, 3, 2, 3000, 2500, 500, 0),   # LDS tile classes
    (62, 20, 3, 2, 3000, 2500, 500, 1),
    (63, 13, 1, 1, 2000, 1000, 300, 0),   # one channel, one phase
    (64, 9, 4, 3, 1500, 700, 200, 1),     # four channels, three phases
    (65, 6, 2, 5, 1200, 300, 150, 0),     # five phases: direct kernel
    (66, 10, 3, 2, 6000, 400, 4000, 0),   # spread beyond the LDS window: direct kernel
    (69, 5, 2, 2, 7, 50, 3, 1),           # a window shorter than one time tile
    (70, 4, 3, 2, 1, 20, 0, 0),           # one sample
])
def test_fused_against_oracle(gpu_lib, seed, S, C, P, T, K, tmax, oob):
    u, alpha, tau, beta = synthetic.random_case(seed, S=S, C=C, P=P, T=T, K=K,
                                                tau_range=(0, tmax))
    F = oracle.prestack(u, alpha, flavour="fma")
    init = np.full(T, 11.0, np.float32)
    m, a = fused_max(gpu_lib, u, alpha, tau, beta, oob, init=init)
    mo, ao = oracle.beamform_max(F, tau, beta, oob, init=init, flavour="fma")
    assert bits_equal(m, mo) and np.array_equal(a, ao)
    sub = np.random.default_rng(seed).choice(K, size=min(K, 96), replace=False)
    b = fused_none(gpu_lib, u, alpha, tau[sub], beta[sub], oob)
    bo = oracle.beamform(F, tau[sub], beta[sub], oob, flavour="fma")
    assert bits_equal(b, bo)


def test_fused_equals_the_reference_wrappers_two_calls(gpu_lib):
    """prestack_waveform_features + beamform_max (the reference wrapper's
    calls through this library) against the fused call, with the kept
    network reused in between, and with no channel (zero features)."""
    u, alpha, tau, beta = synthetic.random_case(67, S=16, C=3, P=2, T=2500, K=1800,
                                                tau_range=(0, 400))
    S, C, T = u.shape
    P = tau.shape[2]
    F = np.zeros(S * T * P, np.float32)
    uu = np.ascontiguousarray(u, np.float32)
    aa = np.ascontiguousarray(alpha, np.float32)
    assert gpu_lib.prestack_waveform_features(uu.ctypes.data_as(P_F), aa.ctypes.data_as(P_F),
                                              T, S, C, P, 1, F.ctypes.data_as(P_F)) == 0
    F = F.reshape(S, T, P)
    tau32 = np.ascontiguousarray(tau, np.int32)
    beta32 = np.ascontiguousarray(beta, np.float32)
    for oob in (0, 1):
        m2 = np.zeros(T, np.float32)
        a2 = np.zeros(T, np.int32)
        assert gpu_lib.beamform_max(F.ctypes.data_as(P_F), tau32.ctypes.data_as(P_I),
                                    beta32.ctypes.data_as(P_F), T, tau.shape[0], S, P, oob,
                                    m2.ctypes.data_as(P_F), a2.ctypes.data_as(P_I)) == 0
        for _ in range(3):  # built, re-planned, reused
            m, a = fused_max(gpu_lib, u, alpha, tau, beta, oob)
            assert bits_equal(m, m2) and np.array_equal(a, a2)
    # C = 0: the prestacked traces are zeros (beamform.c:108-152 with no channel)
    z = np.zeros((S, 0, T), np.float32)
    m, a = fused_max(gpu_lib, z, np.zeros((S, 0, P), np.float32), tau, beta, 1)
    mo, ao = oracle.beamform_max(np.zeros((S, T, P), np.float32), tau, beta, 1, flavour="fma")
    assert bits_equal(m, mo) and np.array_equal(a, ao)
    gpu_lib.bp_dropin_cache_clear()


def test_python_wrapper_inputs_in_any_layout(gpu_lib):
    """The wrapper's fused call takes float64, Fortran-ordered and int64
    inputs with the casts of the reference wrapper: same bits as float32
    C-ordered inputs, and as the oracle."""
    u, alpha, tau, beta = synthetic.random_case(68, S=12, C=3, P=2, T=2000, K=900,
                                                tau_range=(0, 300))
    m, a = bp.beamform(u, tau, alpha, beta, device="gpu", reduce="max")
    F = oracle.prestack(u, alpha, flavour="fma")
    mo, ao = oracle.beamform_max(F, tau, beta, 0, flavour="fma")
    assert bits_equal(m, mo) and np.array_equal(a, ao)
    m2, a2 = bp.beamform(np.asfortranarray(u.astype(np.float64)), tau.astype(np.int64),
                         np.asfortranarray(alpha), beta.astype(np.float64), device="gpu",
                         reduce="max")
    assert bits_equal(m2, m) and np.array_equal(a2, a)
    b = bp.beamform(u, tau[:50], alpha, beta[:50], device="gpu", reduce="none",
                    out_of_bounds="flexible")
    assert bits_equal(b, oracle.beamform(F, tau[:50], beta[:50], 1, flavour="fma"))
    gpu_lib.bp_dropin_cache_clear()
