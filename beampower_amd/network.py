"""Delay-table builder: travel times -> integer moveouts.

The reference takes `time_delays` as an input array; its tutorial builds it
from travel times with `sec_to_samp` and (optionally) makes it relative to
each source's earliest arrival (notebooks/3_beampower.ipynb:378-396).  This
module restates that conversion and adds the closed-form homogeneous
travel times used by the synthetic workloads.
"""
import numpy as np


def sec_to_samp(t, sampling_rate, epsilon=0.2):
    """Seconds -> int32 samples, truncating |t*sr| + epsilon toward zero and
    restoring the sign (notebooks/3_beampower.ipynb:380-392)."""
    t = np.asarray(t)
    sign = np.sign(t)
    mag = np.abs(t * sampling_rate) + epsilon
    return np.int32(sign * np.int32(mag))


def homogeneous_travel_times(sources, stations, velocities):
    """Straight-ray travel times in a homogeneous medium.

    sources: (K,3) km, stations: (S,3) km, velocities: (P,) km/s
    -> (K,S,P) float64 seconds."""
    sources = np.asarray(sources, np.float64)
    stations = np.asarray(stations, np.float64)
    vel = np.asarray(velocities, np.float64)
    d = np.sqrt(((sources[:, None, :] - stations[None, :, :]) ** 2).sum(-1))
    return d[:, :, None] / vel[None, None, :]


def relative_moveouts(time_delays):
    """Subtract each source's minimum over stations and phases
    (notebooks/3_beampower.ipynb:395, the commented-out option)."""
    td = np.asarray(time_delays)
    return (td - td.min(axis=(1, 2), keepdims=True)).astype(np.int32)


def grid_sources(x_range, y_range, z_range, shape):
    """Regular (nx,ny,nz) grid flattened with z fastest -> (K,3) km."""
    xs = np.linspace(x_range[0], x_range[1], shape[0])
    ys = np.linspace(y_range[0], y_range[1], shape[1])
    zs = np.linspace(z_range[0], z_range[1], shape[2])
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def build_time_delays(sources, stations, velocities, sampling_rate, relative=True):
    """(K,S,P) int32 moveouts from homogeneous travel times."""
    tt = homogeneous_travel_times(sources, stations, velocities)
    td = sec_to_samp(tt, sampling_rate)
    return relative_moveouts(td) if relative else td.astype(np.int32)
