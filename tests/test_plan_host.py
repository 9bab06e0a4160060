"""bp_plan, the host-only planner: no device is touched, so these run on the
CPU.  Closed forms of the kernel fit (largest LDS window per phase count,
LDS per work-group), of the sizes and the admission rule, and of the tile
plan on networks built so that their classes and windows are known.  The
GPU test test_gpu_plan checks the plan against networks the engine
actually builds."""
import numpy as np
import pytest

from beampower_amd import plan as bpp
from beampower_amd import synthetic


def test_lds_window_limit_per_phase_count():
    lims = [bpp.plan(100, 5, P, 1000)["lds_window_limit"] for P in (1, 2, 3, 4, 5)]
    assert lims == [2048, 2048, 1600, 1216, 0]


def test_sizes_and_admission_of_the_dropin_shape():
    K, S, P, T = 192_000, 50, 2, 90_000
    p = bpp.plan(K, S, P, T)
    tile = p["source_tile"]
    n_tiles = -(-K // tile)
    sched = n_tiles * S * tile * (P + 1) * 4
    assert p["network_bytes"] == (K * S * P * 4 + K * S * 4 + 2 * K * 4 + n_tiles * S * P * 4
                                  + n_tiles * 10 * 4 + sched + S * 4)
    assert p["feature_bytes"] == S * P * T * 4  # no padding without moveouts
    assert p["output_bytes"] == T * 16
    assert p["call_bytes"] == p["network_bytes"] + p["feature_bytes"] + S * T * P * 4 + T * 16
    assert p["fits"]
    assert bpp.plan(K, S, P, T, hbm_budget=p["call_bytes"])["fits"]
    assert not bpp.plan(K, S, P, T, hbm_budget=p["call_bytes"] - 1)["fits"]
    none = bpp.plan(4096, S, P, T, reduce="none")
    assert none["output_bytes"] == 4096 * T * 4
    b = bpp.plan(K, S, P, T, fmt="bf16")
    assert b["feature_bytes"] == 2 * S * P * (T + (T & 1)) * 2
    assert p["tiles_per_class"] == [-1, -1, -1] and p["similarity_order"] == -1


def _network_with_spreads(spreads, S=6, P=2, tile=48, seed=0):
    """One source tile per entry: every (s,p) of the tile spans exactly
    `spread` samples (delays 0 and spread on two of its sources)."""
    rng = np.random.default_rng(seed)
    K = tile * len(spreads)
    tau = np.zeros((K, S, P), np.int32)
    for j, d in enumerate(spreads):
        blk = rng.integers(0, d + 1, (tile, S, P))
        blk[0], blk[1] = 0, d
        tau[j * tile:(j + 1) * tile] = blk
    return tau, np.ones((K, S), np.float32)


def test_tile_plan_of_known_spreads():
    tile = bpp.plan(48, 2, 2, 100)["source_tile"]
    TT = bpp.plan(48, 2, 2, 100)["time_tile"]
    # every tile spreads 40 samples: one class, window TT + 40 + 1 rounded up to 8
    tau, beta = _network_with_spreads([40] * 10, tile=tile)
    p = bpp.plan(tau.shape[0], 6, 2, 5000, time_delays=tau, weights_sources=beta,
                 order="natural")
    lw = -(-(TT + 41) // 8) * 8
    assert p["tiles_per_class"] == [10, 0, 0] and p["window_per_class"] == [lw, 0]
    assert p["similarity_order"] == 0 and p["pad_before"] == 0
    # the least delay of every (s, p) is 0: staging reaches lw + 7 samples past t0
    assert p["pad_after"] == -(-(0 + lw + 6) // 8) * 8
    assert p["lds_bytes"][0] > 0 and p["workgroups_per_cu"][0] >= 1
    assert p["feature_bytes"] == 6 * 2 * (5000 + p["pad_after"]) * 4
    # a wide tile beyond the small window and one beyond the largest window.
    # The small window covers 97 % of the tiles (spread of the tile at that
    # quantile, at most TT + 512 samples); the large one the widest LDS tile,
    # capped at the LDS limit (2,048 samples for two phases)
    spreads = [40] * 40 + [900, 3000]
    tau, beta = _network_with_spreads(spreads, tile=tile)
    p = bpp.plan(tau.shape[0], 6, 2, 9000, time_delays=tau, weights_sources=beta,
                 order="natural")
    q = sorted(spreads)[int(0.97 * len(spreads))]
    lw0 = min(TT + 512, 2048, -(-(TT + q + 1) // 8) * 8)
    assert p["window_per_class"] == [lw0, 2048]
    assert p["tiles_per_class"] == [40 + (q <= lw0 - TT - 1), 1 - (900 <= lw0 - TT - 1), 1]
    # a non-finite weight sends its tile to the direct kernel
    tau, beta = _network_with_spreads([40] * 5, tile=tile)
    beta[2 * tile + 3, 1] = np.inf
    p = bpp.plan(tau.shape[0], 6, 2, 5000, time_delays=tau, weights_sources=beta,
                 order="natural")
    assert p["tiles_per_class"] == [4, 0, 1]
    # five phases: no LDS kernel, every tile direct
    tau5 = np.zeros((tile * 3, 4, 5), np.int32)
    p = bpp.plan(tile * 3, 4, 5, 1000, time_delays=tau5,
                 weights_sources=np.ones((tile * 3, 4), np.float32), order="natural")
    assert p["tiles_per_class"] == [0, 0, 3] and p["lds_bytes"] == [0, 0]


def test_source_order_decision():
    """A shuffled grid: the similarity order narrows its tiles, so "auto"
    keeps it; the same grid in grid order keeps the caller's order."""
    _, _, tau = synthetic.make_network(3, 16, 90.0, (24, 20, 10))
    beta = np.ones(tau.shape[:2], np.float32)
    K, S, P = tau.shape
    grid = bpp.plan(K, S, P, 20000, time_delays=tau, weights_sources=beta, order="auto")
    shuffled = np.ascontiguousarray(tau[np.random.default_rng(1).permutation(K)])
    shuf = bpp.plan(K, S, P, 20000, time_delays=shuffled, weights_sources=beta, order="auto")
    nat = bpp.plan(K, S, P, 20000, time_delays=shuffled, weights_sources=beta, order="natural")
    assert shuf["similarity_order"] == 1 and nat["similarity_order"] == 0
    assert shuf["window_per_class"][0] < nat["window_per_class"][0]
    assert shuf["network_bytes"] == nat["network_bytes"] + K * 4  # the order table
    assert grid["tiles_per_class"][2] == 0


def test_plan_refuses_bad_requests():
    from beampower_amd.core import load_library
    lib = load_library("gpu")
    for kw in (dict(fmt=7), dict(reduce=3), dict(order=9)):
        with pytest.raises(RuntimeError):
            bpp.plan(10, 2, 2, 100, **kw)
    with pytest.raises(ValueError):
        bpp.plan(10, 2, 2, 100, time_delays=np.zeros((9, 2, 2)), weights_sources=np.ones((9, 2)))
    assert lib.bp_plan(None, None) == -22
