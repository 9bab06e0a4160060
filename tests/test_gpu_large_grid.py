"""Launches beyond one dispatch: a dispatch's grid holds fewer than 2^32
work-items, so a window class of more work-groups than that (a long record
of a large network) runs as several launches of consecutive block ranges.

Bar: the maxima of the whole record, bit for bit, equal those of short
halo-padded slices (which fit one launch) at its start, middle and end, and
one slice equals the oracle.
"""
import numpy as np
import pytest

from conftest import bits_equal
from oracle import oracle

pytestmark = pytest.mark.gpu

DISPATCH_BLOCKS = (2 ** 32 - 1) // 512  # work-groups of 8 waves in one dispatch


def test_class_split_over_several_launches():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beampower_amd import device as bpd
    rng = np.random.default_rng(77)
    import ctypes as ct
    from beampower_amd import load_library
    tile_c, tt_c = ct.c_int(), ct.c_int()
    assert load_library("gpu").bp_kernel_shape(ct.byref(tile_c), ct.byref(tt_c), None) == 0
    tile, time_tile = tile_c.value, tt_c.value
    S, P, T, W = 2, 1, 1_100_000, 2000
    n_tt = (T + time_tile - 1) // time_tile
    n_tiles = DISPATCH_BLOCKS // n_tt + 100
    K = n_tiles * tile
    tau = rng.integers(0, 50, (K, S, P)).astype(np.int32)
    beta = rng.uniform(0.5, 1.5, (K, S)).astype(np.float32)
    halo = int(tau.max()) + 1
    net = bpd.Network(tau, beta, order="natural")
    info = net.info()
    assert info["tiles_per_class"][0] * n_tt > DISPATCH_BLOCKS, info
    F = rng.uniform(0.0, 1.0, (S, P, T)).astype(np.float32)
    Fd = torch.from_numpy(F).cuda()
    m, a = net.beamform_max(Fd, 0)
    m, a = m.cpu().numpy(), a.cpu().numpy()
    for t0 in (0, T // 2, T - W - halo):
        sub = Fd[:, :, t0:t0 + W + halo].contiguous()
        ms, as_ = net.beamform_max(sub, 0)
        assert bits_equal(m[t0:t0 + W], ms.cpu().numpy()[:W]), t0
        assert np.array_equal(a[t0:t0 + W], as_.cpu().numpy()[:W]), t0
    t0 = T - W - halo
    F_stp = np.ascontiguousarray(F[:, :, t0:t0 + W + halo].transpose(0, 2, 1))
    mo, ao = oracle.beamform_max(F_stp, tau, beta, 0, flavour="fma")
    assert bits_equal(m[t0:t0 + W], mo[:W])
    assert np.array_equal(a[t0:t0 + W], ao[:W])
