"""The GPU engine against the reference's Makefile build (Makefile:69,
-std=c99, no fused multiply-add), whose outputs the golden fixtures hold as
the `*_make` arrays: beams within 1e-5 of the term scale and the argmax
equal wherever the top-two gap decides it (tests/contract.py).  Through the
C ABI on the Makefile build's own prestack, and through the Python API from
the raw features (prestacked on the GPU)."""
import numpy as np
import pytest

import contract
from conftest import GOLDEN_CASES, load_golden
import beampower_amd as bp
from beampower_amd import synthetic

from test_gpu_parity import abi_max, abi_none

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_golden_make_flavour_contract(gpu_lib, case):
    g = load_golden(case)
    Fm, tau, beta = g["F_make"], g["tau"], g["beta"]
    for oob in (0, 1):
        if f"max_{oob}_make" in g:
            m, a = abi_max(gpu_lib, Fm, tau, beta, oob)
            contract.check_max(m, a, Fm, tau, beta, oob)
            name = "strict" if oob == 0 else "flexible"
            m, a = bp.beamform(g["u"], tau, g["alpha"], beta, device="gpu", reduce="max",
                               out_of_bounds=name)
            contract.check_max(m, a, Fm, tau, beta, oob)
        if f"none_{oob}_make" in g:
            contract.check_beams(abi_none(gpu_lib, Fm, tau, beta, oob), Fm, tau, beta, oob)


def test_plain_workload_make_flavour_contract(gpu_lib):
    """The plain-method workload (20 stations, 4,000 sources, 6,000
    samples), strict reduce='max', against the Makefile build's golden
    maxima; the gap decides the argmax at most samples."""
    g = load_golden("plain_full")
    w = synthetic.plain_c1()
    m, a = bp.beamform(w.waveform_features, w.time_delays, w.weights_phases,
                       w.weights_sources, device="gpu", reduce="max")
    from oracle import oracle
    Fm = oracle.prestack(w.waveform_features, w.weights_phases, flavour="nofma")
    frac = contract.check_max(m, a, Fm, w.time_delays, w.weights_sources, 0)
    assert frac > 0.5
    # the checker's Makefile maxima are the golden ones
    contract.check_max(g["max_0_make"], g["argmax_0_make"], Fm, w.time_delays,
                       w.weights_sources, 0)
