"""Detection consumer on the GPU against the detection oracle
(oracle/detect_oracle.py) and scipy.signal.find_peaks.

Bar: peak indices identical (integer work); thresholds within 1e-6
relative (window statistics accumulate in a different order: the oracle in
double per window, the reference in float32 pairwise sums); peaks within
that tolerance of the threshold are excluded from the comparison of the
thresholded set.
"""
import numpy as np
import pytest
from scipy import signal

from oracle import detect_oracle as do
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def beam_like(seed, n, n_events=30):
    """A positive, noisy, smoothed series with events, plateaus and a zero
    tail (samples with no valid source)."""
    rng = np.random.default_rng(seed)
    x = np.convolve(np.abs(rng.standard_normal(n)), np.ones(9) / 9, mode="same") + 1.0
    for t in rng.integers(100, n - 100, n_events):
        x[t:t + 5] += rng.uniform(3, 20)
    for _ in range(50):
        i = int(rng.integers(2, n - 20))
        x[i:i + int(rng.integers(2, 12))] = x[i]
    x[-300:] = 0.0
    return x.astype(np.float32)


@pytest.mark.parametrize("seed,n,window,overlap", [(0, 20000, 3000, 0.66), (1, 100000, 18000, 0.66),
                                                   (2, 7000, 1000, 0.5)])
def test_threshold_matches_oracle(seed, n, window, overlap):
    torch = _torch()
    from beampower_amd import detect
    x = beam_like(seed, n)
    thr = detect.time_dependent_threshold(torch.from_numpy(x).cuda(), window, overlap, 8.0)
    ref = do.time_dependent_threshold(x, window, overlap, 8.0)
    assert np.allclose(thr.cpu().numpy(), ref, rtol=1e-6, atol=0)


@pytest.mark.parametrize("seed,n,window,overlap", [(0, 20000, 3000, 0.66), (1, 100000, 18000, 0.66),
                                                   (2, 7000, 1000, 0.5), (7, 9001, 1000, 0.66)])
def test_mad_threshold_matches_oracle(seed, n, window, overlap):
    """"mad": medians are exact selections and the arithmetic is the same
    float32 steps, so the threshold is bit-identical to the oracle's; the
    series has ties (plateaus) and a zero tail."""
    torch = _torch()
    from beampower_amd import detect
    x = beam_like(seed, n)
    if seed == 7:
        x[::2] = np.round(x[::2])  # many equal values
    thr = detect.time_dependent_threshold(torch.from_numpy(x).cuda(), window, overlap, 8.0,
                                          threshold_type="mad")
    ref = do.time_dependent_threshold(x, window, overlap, 8.0, threshold_type="mad")
    assert np.array_equal(thr.cpu().numpy(), ref)


@pytest.mark.parametrize("seed,distance", [(3, 1), (4, 7), (5, 60), (6, 1500)])
def test_peaks_match_scipy(seed, distance):
    torch = _torch()
    from beampower_amd import detect
    x = beam_like(seed, 60000)
    got = detect.find_peaks(torch.from_numpy(x).cuda(), distance)
    ref, _ = signal.find_peaks(x, distance=distance)
    assert np.array_equal(got, ref)
    assert np.array_equal(got, do.find_peaks(x, distance))


def test_thresholded_detection_matches_oracle_and_recovers_events():
    """Plain-method workload end to end: beamform on the GPU, detect on the
    GPU, compare with the oracle's detector on the same beam_max, and find
    the injected events."""
    torch = _torch()
    from beampower_amd import detect
    from beampower_amd import device as bpd
    w = synthetic.plain_c1()
    net = bpd.Network(w.time_delays, w.weights_sources)
    F = bpd.prestack_planar(torch.from_numpy(w.waveform_features).cuda(),
                            torch.from_numpy(w.weights_phases).cuda())
    bm, ba = net.beamform_max(F, 0)
    res = detect.detect(bm, ba, 25.0, window_sec=60.0, overlap=0.66, num_dev=3.0,
                        min_interval_sec=10.0)
    x = bm.cpu().numpy()
    thr_ref = do.time_dependent_threshold(x, 1500, 0.66, 3.0)
    thr = res["threshold"].cpu().numpy()
    assert np.allclose(thr, thr_ref, rtol=1e-6, atol=0)
    ref = do.detect(x, thr_ref, 250)
    close = np.abs(x - thr_ref) <= 1e-5 * np.abs(thr_ref)
    assert set(res["peaks"][~close[res["peaks"]]]) == set(ref[~close[ref]])
    assert np.array_equal(res["sources"], ba.cpu().numpy()[res["peaks"]])
    found = 0
    for k, t in w.events:
        hit = np.abs(res["peaks"] - t) <= 2
        if hit.any() and np.array_equal(w.time_delays[res["sources"][hit][0]], w.time_delays[k]):
            found += 1
    assert found >= len(w.events) - 1, (found, w.events, res["peaks"])
