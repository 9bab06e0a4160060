# coding: utf-8
"""Public API: beamform and prestack_traces.

Mirrors beampower/beampower.py:10-287 of the reference -- same signatures,
defaults, dtype casts, zero-filled outputs, return shapes and error
behaviour -- with device="gpu" served by the gfx950 library
(lib/beamform_gpu.so).  Differences, all on the GPU path:

* prestack runs on the GPU for device="gpu" (the reference always prestacks
  on the CPU, beampower.py:103-105); the FMA sequence is the same, so the
  prestacked traces are bit-identical to the pip build's;
* mode="differential" runs on the GPU (the reference prints "not yet
  implemented" there, beampower.py:220-223) with the CPU branch's outputs
  (beampower.py:194-216);
* prestack_traces(device="gpu") returns the prestacked traces (the
  reference returns zeros there: its only prestack call is under
  `if device == "cpu"`, beampower.py:271-282);
* mode="direct" on the GPU makes one library call, bp_beamform_max_fused
  or bp_beamform_fused, instead of prestack_waveform_features followed by
  beamform_max / beamform: the features go to the device unstacked and are
  prestacked there, so the prestacked traces never cross PCIe (same bits);
* inputs that already have the library's dtype and C order are passed
  without the copies of `.flatten().astype(...)` (same values);
* a failing library call raises RuntimeError with the library's message
  instead of terminating the process (the reference's CUDA path calls
  exit(), beamform.cu:17-26,259-266).
"""
import ctypes as ct
from os import cpu_count

import numpy as np

from .core import check, load_library

_P_F = ct.POINTER(ct.c_float)
_P_I = ct.POINTER(ct.c_int)


def _flat(x, dtype):
    """x in C order as a 1-D array of dtype: the values of
    x.flatten().astype(dtype), copied only where x is not already so."""
    return np.ascontiguousarray(x, dtype=dtype).reshape(-1)


def _fusable(waveform_features, time_delays, weights_phases, weights_sources):
    """Shapes of the fused GPU call: (S, C, T) features, (K, S, P) delays,
    (S, C, P) phase weights and K * S source weights."""
    S, C, _ = waveform_features.shape
    K, S_d, P = time_delays.shape
    return (S_d == S and np.ndim(weights_phases) == 3
            and tuple(np.shape(weights_phases)) == (S, C, P)
            and np.size(weights_sources) == K * S)


def beamform(
    waveform_features,
    time_delays,
    weights_phases,
    weights_sources,
    device="cpu",
    reduce="max",
    mode="direct",
    out_of_bounds="strict",
    num_threads=None,
):
    """Compute the beamformed network response.

    Parameters
    ----------
    waveform_features: (n_stations, n_channels, n_samples) numpy.ndarray, float
        Characteristic function of the continuous seismograms.
    time_delays: (n_sources, n_stations, n_phases) numpy.ndarray, int
        Moveouts, in samples, from each source to each station and phase.
    weights_phases: (n_stations, n_channels, n_phases) numpy.ndarray, float
        Weight of each station and channel for each phase.
    weights_sources: (n_sources, n_stations) numpy.ndarray, float
        Source-receiver weights; a zero weight removes the station from the
        source's beam and from its out-of-bounds test.
    device: 'cpu' or 'gpu'
    reduce: 'max' returns (max over sources, argmax) per sample; 'none',
        'None' or None returns the full (n_sources, n_samples) beams.
    mode: 'direct' (or 'normal') or 'differential' (time delays relative to
        the centre sample (n_samples - 1) // 2; one beam per source).
    out_of_bounds: 'strict' -- a beam is computed only where every used
        moveout points inside the data; 'flexible' -- out-of-range terms are
        dropped.
    num_threads: CPU threads (CPU library only); None = all CPUs.

    Returns
    -------
    beam: (n_sources, n_samples) float32 for reduce='none', or
    (beam_max (n_samples,) float32, beam_argmax (n_samples,) int32);
    mode='differential': (n_sources,) float32 for reduce='none', or
    (max over sources, its first index) for reduce='max'.
    """
    if out_of_bounds not in ["strict", "flexible"]:
        print("out_of_bounds should be either of 'strict' or 'flexible',"
              f" not {out_of_bounds}")
        return
    oob = 0 if out_of_bounds == "strict" else 1

    if num_threads is None:
        num_threads = cpu_count()

    lib = load_library(device)
    on_gpu = device.lower() == "gpu"

    n_stations, _, n_samples = waveform_features.shape
    n_sources, _, n_phases = time_delays.shape

    if mode in ["normal", "direct"] and reduce not in ["none", "None", None, "max"]:
        # the reference's direct branch has no case for this reduce: after a
        # prestack whose result goes unused it falls through to its final
        # else (beampower.py:225-227), the mode message and 1
        print(f"Mode should either be 'direct' or 'differential', not {mode}.")
        return 1

    if (on_gpu and mode in ["normal", "direct"]
            and _fusable(waveform_features, time_delays, weights_phases, weights_sources)):
        wf = _flat(waveform_features, np.float32)
        wp = _flat(weights_phases, np.float32)
        td = _flat(time_delays, np.int32)
        ws = _flat(weights_sources, np.float32)
        args = (wf.ctypes.data_as(_P_F), wp.ctypes.data_as(_P_F), td.ctypes.data_as(_P_I),
                ws.ctypes.data_as(_P_F), n_samples, n_sources, n_stations,
                waveform_features.shape[1], n_phases, oob)
        if reduce in ["none", "None", None]:
            beam = np.zeros(n_sources * n_samples, dtype=np.float32)
            check(lib.bp_beamform_fused(*args, beam.ctypes.data_as(_P_F)),
                  "bp_beamform_fused", lib)
            return beam.reshape(n_sources, n_samples)
        beam_max = np.zeros(n_samples, dtype=np.float32)
        beam_argmax = np.zeros(n_samples, dtype=np.int32)
        check(lib.bp_beamform_max_fused(*args, beam_max.ctypes.data_as(_P_F),
                                        beam_argmax.ctypes.data_as(_P_I)),
              "bp_beamform_max_fused", lib)
        return beam_max, beam_argmax

    if mode in ["normal", "direct", "differential"]:
        F = prestack_traces(waveform_features, weights_phases, num_threads=num_threads,
                            device=device)
        F = _flat(F, np.float32)
        time_delays = _flat(time_delays, np.int32)
        weights_sources = _flat(weights_sources, np.float32)
        args = (F.ctypes.data_as(_P_F), time_delays.ctypes.data_as(_P_I),
                weights_sources.ctypes.data_as(_P_F), n_samples, n_sources, n_stations,
                n_phases)
        threads = () if on_gpu else (int(num_threads),)

    if mode in ["normal", "direct"]:
        if reduce in ["none", "None", None]:
            beam = np.zeros(n_sources * n_samples, dtype=np.float32)
            status = lib.beamform(*args, oob, *threads, beam.ctypes.data_as(_P_F))
            if on_gpu:
                check(status, "beamform", lib)
            return beam.reshape(n_sources, n_samples)
        elif reduce == "max":
            beam_max = np.zeros(n_samples, dtype=np.float32)
            beam_argmax = np.zeros(n_samples, dtype=np.int32)
            status = lib.beamform_max(*args, oob, *threads, beam_max.ctypes.data_as(_P_F),
                                      beam_argmax.ctypes.data_as(_P_I))
            if on_gpu:
                check(status, "beamform_max", lib)
            return beam_max, beam_argmax

    if mode == "differential":
        beam = np.zeros(n_sources, dtype=np.float32)
        status = lib.beamform_differential(*args, int(num_threads), beam.ctypes.data_as(_P_F))
        if on_gpu:
            check(status, "beamform_differential", lib)
        if reduce in ["none", "None", None]:
            return beam
        elif reduce == "max":
            return np.max(beam, axis=0), np.argmax(beam, axis=0)
        return None

    print(f"Mode should either be 'direct' or 'differential', not {mode}.")
    return 1


def prestack_traces(waveform_features, weights_phases, num_threads=None, device="cpu"):
    """Channel-wise stack of the features for each phase.

    Parameters
    ----------
    waveform_features: (n_stations, n_channels, n_samples) numpy.ndarray
    weights_phases: (n_stations, n_channels, n_phases) numpy.ndarray
    num_threads: CPU threads (CPU library only).
    device: 'cpu' or 'gpu'.

    Returns
    -------
    prestacked_traces: (n_stations, n_samples, n_phases) float32, with
    prestacked[s, t, p] = sum_c weights_phases[s, c, p] * waveform_features[s, c, t]
    """
    lib = load_library(device)
    if num_threads is None:
        num_threads = cpu_count()
    n_stations, n_channels, n_samples = waveform_features.shape
    _, _, n_phases = weights_phases.shape
    out = np.zeros(n_stations * n_samples * n_phases, dtype=np.float32)
    wf = _flat(waveform_features, np.float32)
    wp = _flat(weights_phases, np.float32)
    status = lib.prestack_waveform_features(
        wf.ctypes.data_as(_P_F), wp.ctypes.data_as(_P_F), n_samples, n_stations, n_channels,
        n_phases, int(num_threads), out.ctypes.data_as(_P_F))
    if device.lower() == "gpu":
        check(status, "prestack_waveform_features", lib)
    return out.reshape((n_stations, n_samples, n_phases))
