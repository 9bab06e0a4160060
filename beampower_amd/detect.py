"""Detection consumer on the device: the tutorial's step after the
beamforming (notebooks/3_beampower.ipynb:497-626) — a time-dependent
threshold of beam_max and its peaks above it, at least a minimum interval
apart; the source of each detection is beam_argmax at the peak.

The threshold statistics run on the GPU (bp_detect_threshold); local maxima
are flagged on the GPU and selected by height on the host
(bp_detect_peaks).
"""
import ctypes as ct

import numpy as np

from .core import check, load_library


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("beampower_amd.detect needs a ROCm GPU")
    return torch


def _stream(stream):
    torch = _torch()
    return ct.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)


THRESHOLD_TYPES = {"rms": 0, "mad": 1}


def time_dependent_threshold(beam_max, window, overlap=0.66, num_dev=8.0, threshold_type="rms",
                             stream=None):
    """beam_max: (T,) float32 CUDA tensor -> (T,) float32 CUDA threshold
    (notebooks/3_beampower.ipynb:497-578).  threshold_type "rms" (mean and
    standard deviation per window) or "mad" (median and median absolute
    deviation); exact zeros are left out of the window statistics."""
    torch = _torch()
    lib = load_library("gpu")
    kind = THRESHOLD_TYPES.get(str(threshold_type).lower())
    if kind is None:
        raise ValueError(f"threshold_type must be 'rms' or 'mad', not {threshold_type!r}")
    x = beam_max.contiguous()
    out = torch.empty_like(x)
    check(lib.bp_detect_threshold_type(x.data_ptr(), x.numel(), int(window), float(overlap),
                                       float(num_dev), kind, out.data_ptr(), _stream(stream)),
          "bp_detect_threshold_type", lib)
    return out


def find_peaks(beam_max, distance, threshold=None, stream=None):
    """Peaks of beam_max (CUDA tensor) at least `distance` samples apart,
    as scipy.signal.find_peaks(distance=...), optionally only those strictly
    above `threshold` (CUDA tensor).  Returns ascending int64 indices."""
    lib = load_library("gpu")
    x = beam_max.contiguous()
    thr = None if threshold is None else threshold.contiguous()
    n = ct.c_size_t()
    cap = max(16, x.numel() // 64)
    while True:
        buf = np.empty(cap, np.int32)
        check(lib.bp_detect_peaks(x.data_ptr(), None if thr is None else thr.data_ptr(),
                                  x.numel(), int(distance), buf.ctypes.data, cap, ct.byref(n),
                                  _stream(stream)), "bp_detect_peaks", lib)
        if n.value <= cap:
            return buf[:n.value].astype(np.int64)
        cap = n.value


def detect(beam_max, beam_argmax, sampling_rate, window_sec=7200.0, overlap=0.66, num_dev=8.0,
           min_interval_sec=60.0, threshold_type="rms", stream=None):
    """The tutorial's detector (notebooks/3_beampower.ipynb:582-626): returns
    dict(peaks, sources, values, threshold) with peaks as sample indices,
    sources = beam_argmax at the peaks, values = beam_max at the peaks and
    the threshold CUDA tensor."""
    window = int(window_sec * sampling_rate)
    thr = time_dependent_threshold(beam_max, window, overlap, num_dev, threshold_type, stream)
    peaks = find_peaks(beam_max, int(min_interval_sec * sampling_rate), thr, stream)
    idx = _torch().from_numpy(peaks).to(beam_max.device)
    return dict(peaks=peaks, sources=beam_argmax[idx].cpu().numpy(),
                values=beam_max[idx].cpu().numpy(), threshold=thr)
