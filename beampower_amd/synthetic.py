"""Seeded synthetic beamforming workloads.

No real seismic data is used (there is no network access).  Each workload
follows the reference tutorial's data model (notebooks/3_beampower.ipynb):
positive envelope-like features on 3 channels, P and S moveouts from a
source grid, phase weights that route the vertical channel to P and the
horizontals to S (:423-426), and unit source weights (:445).

Features are |N(0,1)| smoothed by a 10-sample boxcar, plus injected events:
an event at source k_e and origin sample t_e adds `amplitude` at sample
t_e + tau[k_e,s,0] on channel 0 (P) and at t_e + tau[k_e,s,1] on channels
1-2 (S) of every station s.
"""
from dataclasses import dataclass, field

import numpy as np

from .network import build_time_delays, grid_sources

VP_KM_S = 6.0
VS_KM_S = 3.5
SAMPLING_RATE_HZ = 25.0


@dataclass
class Workload:
    name: str
    waveform_features: np.ndarray  # (S,C,T) float32
    time_delays: np.ndarray        # (K,S,P) int32
    weights_phases: np.ndarray     # (S,C,P) float32
    weights_sources: np.ndarray    # (K,S) float32
    events: list = field(default_factory=list)  # [(k_e, t_e)]
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        S, C, T = self.waveform_features.shape
        K, _, P = self.time_delays.shape
        return dict(S=S, C=C, T=T, K=K, P=P)


def boxcar_envelopes(rng, n_stations, n_channels, n_samples, width=10):
    """|N(0,1)| smoothed by a `width`-sample moving average (float32)."""
    x = np.abs(rng.standard_normal((n_stations, n_channels, n_samples + width - 1)))
    c = np.cumsum(x, axis=-1, dtype=np.float64)
    c = np.concatenate([np.zeros(c.shape[:-1] + (1,)), c], axis=-1)
    return ((c[..., width:] - c[..., :-width]) / width).astype(np.float32)


def tutorial_phase_weights(n_stations, n_channels=3):
    """Channel 0 -> P (1, 0); channels 1..C-1 -> S (0, 0.5)."""
    w = np.zeros((n_stations, n_channels, 2), np.float32)
    w[:, 0, 0] = 1.0
    w[:, 1:, 1] = 0.5
    return w


def inject_events(rng, features, time_delays, n_events, amplitude=5.0):
    S, C, T = features.shape
    K = time_delays.shape[0]
    span = time_delays.max(axis=(1, 2))
    events = []
    for _ in range(n_events):
        k = int(rng.integers(K))
        t = int(rng.integers(0, T - int(span[k]) - 1))
        for s in range(S):
            features[s, 0, t + time_delays[k, s, 0]] += amplitude
            features[s, 1:, t + time_delays[k, s, 1]] += amplitude
        events.append((k, t))
    return events


def make_network(seed, n_stations, box_km, grid_shape, depth_km=(0.0, 20.0)):
    rng = np.random.default_rng(seed)
    st = np.zeros((n_stations, 3))
    st[:, :2] = rng.uniform(0.0, box_km, size=(n_stations, 2))
    src = grid_sources((0.0, box_km), (0.0, box_km), depth_km, grid_shape)
    tau = build_time_delays(src, st, (VP_KM_S, VS_KM_S), SAMPLING_RATE_HZ)
    return st, src, tau


def make_workload(name, seed, n_stations, box_km, grid_shape, n_samples, n_events,
                  depth_km=(0.0, 20.0)):
    stations, sources, tau = make_network(seed, n_stations, box_km, grid_shape, depth_km)
    rng = np.random.default_rng(seed + 1)
    feats = boxcar_envelopes(rng, n_stations, 3, n_samples)
    events = inject_events(rng, feats, tau, n_events)
    return Workload(
        name=name,
        waveform_features=feats,
        time_delays=tau,
        weights_phases=tutorial_phase_weights(n_stations),
        weights_sources=np.ones(tau.shape[:2], np.float32),
        events=events,
        meta=dict(seed=seed, box_km=box_km, grid_shape=tuple(grid_shape),
                  depth_km=tuple(depth_km), sampling_rate=SAMPLING_RATE_HZ,
                  vp=VP_KM_S, vs=VS_KM_S, stations_km=stations, sources_km=sources),
    )


def plain_c1(seed=1):
    """Plain method, CPU-runnable: 20 stations in 50x50 km, 20x20x10 grid
    (4,000 sources, depth 0-20 km), 6,000 samples at 25 Hz, 5 events."""
    return make_workload("plain-20sta-4k-6000", seed, 20, 50.0, (20, 20, 10), 6000, 5)


def dropin_c2(seed=2):
    """FP32 drop-in at one-GPU shape: 50 stations in 100x100 km, 80x80x30
    grid (192,000 sources, depth 0-20 km), 1 h at 25 Hz (90,000 samples),
    50 events."""
    return make_workload("dropin-50sta-192k-1h", seed, 50, 100.0, (80, 80, 30), 90000, 50)


def precision_c3(seed=3, n_samples=2_160_000):
    """Precision form: C2 geometry, 100,000-source 50x50x40 grid, 24 h."""
    return make_workload("precision-50sta-100k-24h", seed, 50, 100.0, (50, 50, 40),
                         n_samples, 200)


def batched_c4_window(seed, minutes=(1.0, 30.0), stations=(10, 80), sources=(1_000, 200_000)):
    """One window of the batched form: length uniform in `minutes` at 25 Hz,
    station count uniform in `stations`, a source grid of about K sources (K
    uniform in `sources`; nz ~ K^(1/3)/2 depth levels, a square horizontal
    grid), plain-method geometry (50 km box, depth 0-20 km), seeded per
    window."""
    rng = np.random.default_rng(seed)
    n_samples = int(rng.uniform(*minutes) * 60 * SAMPLING_RATE_HZ)
    n_stations = int(rng.integers(stations[0], stations[1] + 1))
    k = float(rng.uniform(*sources))
    nz = max(2, int(round(k ** (1.0 / 3.0) / 2)))
    nxy = max(2, int(round((k / nz) ** 0.5)))
    return make_workload(f"batch-{seed}", seed, n_stations, 50.0, (nxy, nxy, nz), n_samples,
                         max(1, n_samples // 3000))


SERVING_NETWORKS = (  # (stations, box km, source grid): 10k, 50k, 192k, 300k sources
    (20, 50.0, (25, 25, 16)),
    (50, 100.0, (50, 50, 20)),
    (80, 150.0, (80, 80, 30)),
    (120, 200.0, (100, 100, 30)),
)


def serving_networks(seed=5, shapes=SERVING_NETWORKS):
    """The serving form's resident networks: (moveouts, source weights,
    phase weights) per network, plain-method geometry in boxes growing with
    the station count."""
    out = []
    for i, (n_stations, box, grid) in enumerate(shapes):
        _, _, tau = make_network(seed + i, n_stations, box, grid)
        out.append((tau, np.ones(tau.shape[:2], np.float32), tutorial_phase_weights(n_stations)))
    return out


def serving_requests(networks, n, seed, seconds=(30.0, 600.0), pool_per_network=6):
    """n window requests (network index, features (S,3,T)) with networks
    uniform and lengths uniform in `seconds` at 25 Hz; features come from a
    pool of pool_per_network envelope windows per network (reused)."""
    rng = np.random.default_rng(seed)
    pool = []
    for tau, _, _ in networks:
        S = tau.shape[1]
        pool.append([boxcar_envelopes(rng, S, 3, int(rng.uniform(*seconds) * SAMPLING_RATE_HZ))
                     for _ in range(pool_per_network)])
    reqs = []
    for _ in range(n):
        k = int(rng.integers(len(networks)))
        reqs.append((k, pool[k][int(rng.integers(pool_per_network))]))
    return reqs


def differential_workload(name, seed, n_stations, box_km, grid_shape, margin=10, n_events=1,
                          depth_km=(0.0, 20.0)):
    """mode='differential': the "stations" are the n_stations*(n_stations-1)/2
    station pairs (i < j) and the features their cross-correlation
    envelopes, one channel, zero lag at the centre sample (T - 1) / 2.
    time_delays[k, (i, j), p] = tau[k, j, p] - tau[k, i, p] for absolute
    P and S moveouts tau; T = 2 * max|delay| + 1 + 2 * margin.  Features are
    smoothed |N(0,1)| with, for each of n_events true sources, a peak of 5
    at the P and S differential times of every pair; phase weights 1; source
    weights 1."""
    stations_km, sources_km, _ = make_network(seed, n_stations, box_km, grid_shape, depth_km)
    tt = build_time_delays(sources_km, stations_km, (VP_KM_S, VS_KM_S), SAMPLING_RATE_HZ,
                           relative=False)
    i, j = np.triu_indices(n_stations, k=1)
    diff = np.ascontiguousarray((tt[:, j, :] - tt[:, i, :]).astype(np.int32))
    half = int(np.abs(diff).max()) + margin
    T = 2 * half + 1
    rng = np.random.default_rng(seed + 1)
    n_pairs = len(i)
    feats = boxcar_envelopes(rng, n_pairs, 1, T, width=3)
    events = []
    for _ in range(n_events):
        k = int(rng.integers(len(sources_km)))
        for q in range(n_pairs):
            for ph in range(2):
                feats[q, 0, half + diff[k, q, ph]] += 5.0
        events.append((k, half))
    return Workload(
        name=name, waveform_features=feats, time_delays=diff,
        weights_phases=np.ones((n_pairs, 1, 2), np.float32),
        weights_sources=np.ones((len(sources_km), n_pairs), np.float32), events=events,
        meta=dict(seed=seed, box_km=box_km, grid_shape=tuple(grid_shape), n_stations=n_stations,
                  sampling_rate=SAMPLING_RATE_HZ, margin=margin))


def differential_c5(seed=7):
    """Differential mode, CPU-runnable: 20 stations (190 pairs) in 50x50 km,
    20x20x10 grid (4,000 sources)."""
    return differential_workload("differential-20sta-190pairs-4k", seed, 20, 50.0, (20, 20, 10))


def differential_c6(seed=8):
    """Differential mode at one-GPU shape: 50 stations (1,225 pairs) in
    100x100 km, 80x80x30 grid (192,000 sources)."""
    return differential_workload("differential-50sta-1225pairs-192k", seed, 50, 100.0,
                                 (80, 80, 30))


def subset_sources(n_sources, n_subset, seed):
    """Seeded sorted subset of source indices (the reduce='none' case)."""
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n_sources, size=n_subset, replace=False))


def random_case(seed, S, C, P, T, K, tau_range=(0, 40), beta_mode="mixed",
                negative_delays=False):
    """Small random case with zero, negative and all-zero weight rows."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((S, C, T)).astype(np.float32)
    alpha = rng.uniform(-1, 1, (S, C, P)).astype(np.float32)
    lo, hi = tau_range
    if negative_delays:
        lo = -hi // 2
    tau = rng.integers(lo, hi + 1, (K, S, P)).astype(np.int32)
    if beta_mode == "ones":
        beta = np.ones((K, S), np.float32)
    else:
        beta = rng.uniform(-1.0, 2.0, (K, S)).astype(np.float32)
        beta[rng.random((K, S)) < 0.2] = 0.0
        if K > 2:
            beta[1, :] = 0.0            # a source with no used station
            beta[2, 0] = -0.0           # negative zero counts as unused
    return u, alpha, tau, beta
