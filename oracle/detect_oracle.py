"""CPU restatement of the tutorial's detection step on beam_max.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the oracle of the library's
detection consumer (bp_detect_threshold / bp_detect_peaks).

Reference: notebooks/3_beampower.ipynb
  * time_dependent_threshold (:497-578): sliding windows of `window` samples
    every int((1 - overlap) * window) samples, center + num_dev * deviation
    per window ("rms": mean and population standard deviation; "mad": median
    and median absolute deviation), one-step max with each neighbour,
    nearest-window expansion to every sample;
  * detection (:618-626): scipy.signal.find_peaks(beam_max, distance=...)
    then peaks above the threshold.

Third-party dependency of that step: SciPy (`scipy.signal.find_peaks`; this
image has scipy 1.15.3).  Its peak rules are restated in `find_peaks`
below (local maxima with plateau midpoints, then greedy selection by height
within `distance`) and pinned against scipy itself in tests/.

The reference fills exact zeros of the series with unseeded Gaussian noise of
the non-zero mean and deviation before taking window statistics
(:529-539); that cannot be reproduced.  `zeros="exclude"` drops them from the
statistics instead (the library's rule); on series without zeros the two
agree.
"""
import numpy as np


def time_dependent_threshold(series, window, overlap=0.66, num_dev=8.0, zeros="exclude",
                             threshold_type="rms"):
    """float32 threshold of len(series) (notebooks/3_beampower.ipynb:497-578).
    threshold_type: "rms" (mean and population standard deviation per
    window) or "mad" (median and median absolute deviation, numpy's float32
    median).  zeros: "exclude" (drop exact zeros from every window's
    statistics) or a numpy Generator (the reference's noise fill, seeded)."""
    x = np.asarray(series, np.float32).copy()
    kind = threshold_type.lower()
    if kind not in ("rms", "mad"):
        raise ValueError(f"threshold_type must be 'rms' or 'mad', not {threshold_type!r}")
    n = len(x)
    half = window // 2
    shift = int((1.0 - overlap) * window)
    if n <= window or shift < 1:
        raise ValueError("the series must be longer than the window, and the shift >= 1")
    z = x == 0.0
    starts = np.arange(0, n - window + 1, shift)
    if isinstance(zeros, np.random.Generator):
        fill = zeros.normal(size=int(z.sum())).astype(np.float32)
        if kind == "rms":
            x[z] = fill * np.std(x[~z]) + x[~z].mean()
        else:
            c0 = np.median(x[~z])
            x[z] = fill * np.median(np.abs(x[~z] - c0)) + c0
        win = np.lib.stride_tricks.sliding_window_view(x, window)[::shift, :]
        if kind == "rms":
            center = np.mean(win, axis=-1)
            deviation = np.std(win, axis=-1)
        else:
            center = np.median(win, axis=-1)
            deviation = np.median(np.abs(win - center[:, np.newaxis]), axis=-1)
    else:
        center = np.empty(len(starts), np.float32)
        deviation = np.empty(len(starts), np.float32)
        for j, s in enumerate(starts):
            if kind == "mad":
                w = x[s:s + window]
                w = w[w != 0.0]
                c = np.median(w) if w.size else np.float32(0.0)
                center[j] = c
                deviation[j] = np.median(np.abs(w - c)) if w.size else 0.0
                continue
            w = x[s:s + window].astype(np.float64)
            w = w[w != 0.0]
            m = w.mean() if w.size else 0.0
            center[j] = m
            deviation[j] = np.sqrt(np.mean((w - m) ** 2)) if w.size else 0.0
    thr = (center + np.float32(num_dev) * deviation).astype(np.float32)
    thr[1:] = np.maximum(thr[:-1], thr[1:])
    thr[:-1] = np.maximum(thr[:-1], thr[1:])
    t = np.arange(half, n - (window - half))
    idx = t // shift
    idx[idx >= len(thr)] = len(thr) - 1
    mid = thr[idx]
    return np.concatenate([np.full(half, mid[0], np.float32), mid,
                           np.full(window - half, mid[-1], np.float32)])


def local_maxima(x):
    """Indices of local maxima, plateaus reported at their midpoint (left
    for even lengths); the first and last samples are never maxima."""
    x = np.asarray(x)
    out = []
    i, i_max = 1, len(x) - 1
    while i < i_max:
        if x[i - 1] < x[i]:
            j = i + 1
            while j < i_max and x[j] == x[i]:
                j += 1
            if x[j] < x[i]:
                out.append((i + j - 1) // 2)
                i = j
        i += 1
    return np.asarray(out, np.int64)


def select_by_distance(peaks, heights, distance):
    """Greedy selection: the highest remaining peak is kept and removes every
    peak closer than `distance` samples.  Equal heights are taken in
    descending index order (scipy visits np.argsort(heights) from the end;
    a stable sort gives that order, scipy's default sort leaves it
    unspecified)."""
    keep = np.ones(len(peaks), bool)
    order = np.lexsort((-np.asarray(peaks), -heights.astype(np.float64)))
    for j in order:
        if not keep[j]:
            continue
        k = j - 1
        while k >= 0 and peaks[j] - peaks[k] < distance:
            keep[k] = False
            k -= 1
        k = j + 1
        while k < len(peaks) and peaks[k] - peaks[j] < distance:
            keep[k] = False
            k += 1
    return peaks[keep]


def find_peaks(x, distance):
    """Local maxima at least `distance` samples apart (scipy.signal.find_peaks
    with `distance`, notebooks/3_beampower.ipynb:621)."""
    p = local_maxima(x)
    return select_by_distance(p, np.asarray(x)[p], int(np.ceil(distance)))


def detect(beam_max, threshold, distance):
    """Peaks of beam_max above the threshold (notebooks/3_beampower.ipynb:621-626)."""
    p = find_peaks(beam_max, distance)
    return p[np.asarray(beam_max)[p] > np.asarray(threshold)[p]]
