"""Pins the detection oracle (oracle/detect_oracle.py) on the CPU.

* time_dependent_threshold against closed forms: constant series, an
  alternating series (every even window has mean 2 and deviation 1), a step
  (windows on either side are constant; the neighbour max and the expansion
  are checked sample by sample), and the zero rule;
* find_peaks against scipy.signal.find_peaks (the reference's dependency) on
  random series with injected plateaus, for several distances.
"""
import numpy as np
import pytest
from scipy import signal

from oracle import detect_oracle as do


def test_threshold_constant_series():
    thr = do.time_dependent_threshold(np.full(1000, 3.5, np.float32), 200, num_dev=8)
    assert thr.dtype == np.float32 and thr.shape == (1000,)
    assert np.all(thr == np.float32(3.5))


def test_threshold_alternating_series_has_mean_2_deviation_1():
    x = np.tile(np.array([1.0, 3.0], np.float32), 600)
    thr = do.time_dependent_threshold(x, 100, overlap=0.5, num_dev=8)
    assert np.allclose(thr, 2 + 8 * 1.0, rtol=1e-6)


def test_threshold_step_neighbour_max_and_expansion():
    # window 100, shift 50, half 50; 1's before 500, 5's from 500
    x = np.concatenate([np.ones(500, np.float32), np.full(500, 5.0, np.float32)])
    thr = do.time_dependent_threshold(x, 100, overlap=0.5, num_dev=2.0)
    starts = np.arange(0, 1000 - 100 + 1, 50)
    per = []
    for s in starts:
        w = x[s:s + 100].astype(np.float64)
        per.append(w.mean() + 2.0 * w.std())
    per = np.array(per, np.float32)
    assert per[0] == 1 and per[-1] == 5
    stepped = per.copy()
    stepped[1:] = np.maximum(per[:-1], per[1:])
    stepped[:-1] = np.maximum(stepped[:-1], stepped[1:])
    for t in (0, 49, 50, 120, 420, 470, 499, 500, 650, 899, 900, 999):
        j = min(t // 50, len(per) - 1) if 50 <= t < 900 else (1 if t < 50 else len(per) - 1)
        assert np.isclose(thr[t], stepped[j], rtol=1e-6), t
    # the rise starts one window before the step and the maximum holds one window after
    assert thr[0] == 1 and thr[999] == 5
    assert thr[400] > 1 and thr[650] == np.float32(stepped[13])


def test_threshold_excludes_exact_zeros():
    x = np.full(2000, 2.0, np.float32)
    x[::7] = 0.0
    thr = do.time_dependent_threshold(x, 300, num_dev=8)
    assert np.all(thr == np.float32(2.0))


def test_threshold_reference_fill_agrees_on_zero_free_series():
    rng = np.random.default_rng(0)
    x = np.abs(rng.standard_normal(5000)).astype(np.float32) + 0.1
    a = do.time_dependent_threshold(x, 700, num_dev=8)
    b = do.time_dependent_threshold(x, 700, num_dev=8, zeros=np.random.default_rng(1))
    assert np.allclose(a, b, rtol=2e-6)


def test_mad_threshold_closed_forms():
    """"mad": constant series -> deviation 0; the alternating 1, 3 series
    has median 2 (mean of the two middle values) and every |x - 2| = 1;
    1..100 repeated (window 100) has median 50.5 and median absolute
    deviation 25 (the mean of 24.5 and 25.5); exact zeros are left out."""
    thr = do.time_dependent_threshold(np.full(1000, 3.5, np.float32), 200, num_dev=8,
                                      threshold_type="mad")
    assert np.all(thr == np.float32(3.5))
    x = np.tile(np.array([1.0, 3.0], np.float32), 600)
    thr = do.time_dependent_threshold(x, 100, overlap=0.5, num_dev=8, threshold_type="MAD")
    assert np.all(thr == np.float32(2 + 8 * 1.0))
    x = np.tile(np.arange(1, 101, dtype=np.float32), 12)
    thr = do.time_dependent_threshold(x, 100, overlap=0.5, num_dev=2, threshold_type="mad")
    assert np.all(thr == np.float32(50.5 + 2 * 25.0))
    x = np.full(2000, 2.0, np.float32)
    x[::3] = 0.0
    thr = do.time_dependent_threshold(x, 300, num_dev=8, threshold_type="mad")
    assert np.all(thr == np.float32(2.0))


def test_mad_threshold_is_the_tutorials_on_zero_free_series():
    """On a series without zeros the tutorial's own steps (noise fill of no
    sample, np.median over sliding_window_view, :541-553) give the same
    bits as the oracle's per-window exclusion rule."""
    rng = np.random.default_rng(3)
    x = np.abs(rng.standard_normal(6000)).astype(np.float32) + 0.1
    a = do.time_dependent_threshold(x, 700, num_dev=8, threshold_type="mad")
    b = do.time_dependent_threshold(x, 700, num_dev=8, zeros=np.random.default_rng(1),
                                    threshold_type="mad")
    assert np.array_equal(a, b)
    # and the mad threshold differs from the rms one (a wrong branch fails)
    assert not np.allclose(a, do.time_dependent_threshold(x, 700, num_dev=8), rtol=1e-3)
    with pytest.raises(ValueError):
        do.time_dependent_threshold(x, 700, threshold_type="std")


@pytest.mark.parametrize("seed,distance", [(0, 1), (1, 5), (2, 60), (3, 250), (4, 1500)])
def test_find_peaks_matches_scipy(seed, distance):
    rng = np.random.default_rng(seed)
    x = np.convolve(rng.standard_normal(20000), np.ones(7) / 7, mode="same").astype(np.float32)
    for _ in range(40):  # plateaus of 1..9 samples, some at peaks
        i = int(rng.integers(2, len(x) - 12))
        x[i:i + int(rng.integers(1, 10))] = x[i] + (rng.random() < 0.5)
    x[-5:] = x[-5]  # plateau touching the end: never a peak
    ref, _ = signal.find_peaks(x, distance=distance)
    assert np.array_equal(do.find_peaks(x, distance), ref)


def test_find_peaks_ties_take_the_later_index():
    x = np.array([0, 2, 0, 0, 2, 0, 1, 3, 0], np.float32)
    assert list(do.find_peaks(x, 1)) == [1, 4, 7]
    assert list(do.find_peaks(x, 4)) == [1, 7]  # 7 removes 4; 1 is 6 samples from 7
    y = np.array([0, 2, 0, 0, 2, 0, 0, 0, 0], np.float32)
    assert list(do.find_peaks(y, 4)) == [4]
