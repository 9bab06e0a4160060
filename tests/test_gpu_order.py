"""Source order of a network's plan (bp_network_create_ordered) on the GPU.

The similarity order changes which sources share a tile, never a result:
beams are written, and max ties broken, by the caller's source index.

Bar: bit-identical maxima, argmax (lowest caller index among equal
values), full beams and differential beams to the oracle in every order,
and natural-order results equal bit for bit to similarity-order results in
the 16-bit storage formats and in the batched form.
"""
import ctypes as ct

import numpy as np
import pytest

from conftest import bits_equal
from oracle import oracle
import beampower_amd as bp
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu

P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def shuffled_network(seed, S=12, grid=(16, 10, 10), box_km=80.0, T=4000, n_dup=40):
    """Plain-method geometry with the sources in a seeded random order, and
    n_dup rows copied onto other rows far away in that order, so that equal
    beams (ties) sit in different tiles of the natural order and may be
    reordered by the similarity order."""
    rng = np.random.default_rng(seed)
    _, _, tau = synthetic.make_network(seed, S, box_km, grid)
    tau = np.ascontiguousarray(tau[rng.permutation(tau.shape[0])])
    K = tau.shape[0]
    beta = np.ones((K, S), np.float32)
    src = rng.choice(K, n_dup, replace=False)
    dst = rng.choice(np.setdiff1d(np.arange(K), src), n_dup, replace=False)
    tau[dst] = tau[src]
    beta[dst] = beta[src]
    u = synthetic.boxcar_envelopes(rng, S, 3, T)
    alpha = synthetic.tutorial_phase_weights(S)
    return u, alpha, tau, beta


def planar(torch, F_stp):
    return torch.from_numpy(np.ascontiguousarray(F_stp.transpose(0, 2, 1))).cuda()


def test_similarity_order_is_a_permutation_with_narrower_tiles():
    torch = _torch()
    from beampower_amd import device as bpd
    _, _, tau, beta = shuffled_network(1)
    nat = bpd.Network(tau, beta, order="natural")
    sim = bpd.Network(tau, beta, order="similarity")
    auto = bpd.Network(tau, beta)
    p_nat, o_nat = nat.source_order()
    p_sim, o_sim = sim.source_order()
    assert not p_nat and np.array_equal(o_nat, np.arange(tau.shape[0]))
    assert p_sim and np.array_equal(np.sort(o_sim), np.arange(tau.shape[0]))
    # shuffled sources: every natural tile spans the volume, the similarity
    # order groups neighbours (and AUTO takes it)
    assert sim.info()["window_per_class"][0] < nat.info()["window_per_class"][0]
    assert auto.source_order()[0]
    torch.cuda.synchronize()


@pytest.mark.parametrize("oob", [0, 1])
@pytest.mark.parametrize("order", ["natural", "similarity", "auto"])
def test_max_in_every_order_matches_oracle(order, oob):
    torch = _torch()
    from beampower_amd import device as bpd
    u, alpha, tau, beta = shuffled_network(2)
    F = oracle.prestack(u, alpha, flavour="fma")
    net = bpd.Network(tau, beta, order=order)
    m, a = net.beamform_max(planar(torch, F), oob)
    mo, ao = oracle.beamform_max(F, tau, beta, oob, flavour="fma")
    assert bits_equal(m.cpu().numpy(), mo)
    assert np.array_equal(a.cpu().numpy(), ao)


@pytest.mark.parametrize("oob", [0, 1])
def test_max_with_direct_tiles_and_mixed_weights(oob):
    """Random delays wide enough for the large-window and direct kernels,
    zero, -0.0 and all-zero weight rows, duplicated rows; forced order."""
    torch = _torch()
    from beampower_amd import device as bpd
    u, alpha, tau, beta = synthetic.random_case(3, S=7, C=3, P=2, T=4500, K=900,
                                                tau_range=(0, 1700))
    tau[:450] %= 600  # half the sources with narrower delays: large-window tiles
    tau[700:760] = tau[100:160]
    beta[700:760] = beta[100:160]
    F = oracle.prestack(u, alpha, flavour="fma")
    net = bpd.Network(tau, beta, order="similarity")
    cls = net.info()["tiles_per_class"]
    assert cls[1] > 0 and cls[2] > 0, cls
    m, a = net.beamform_max(planar(torch, F), oob)
    mo, ao = oracle.beamform_max(F, tau, beta, oob, flavour="fma")
    assert bits_equal(m.cpu().numpy(), mo)
    assert np.array_equal(a.cpu().numpy(), ao)


@pytest.mark.parametrize("oob", [0, 1])
def test_full_beams_full_and_partial_ranges(oob):
    torch = _torch()
    from beampower_amd import device as bpd
    u, alpha, tau, beta = shuffled_network(4, grid=(10, 8, 6), T=2500)
    F = oracle.prestack(u, alpha, flavour="fma")
    ref = oracle.beamform(F, tau, beta, oob, flavour="fma")
    net = bpd.Network(tau, beta, order="similarity")
    Fd = planar(torch, F)
    b = net.beamform(Fd, oob)
    assert bits_equal(b.cpu().numpy(), ref)
    for lo, hi in ((0, 1), (37, 301), (300, 480)):
        part = net.beamform(Fd, oob, source_begin=lo, source_end=hi)
        assert bits_equal(part.cpu().numpy(), ref[lo:hi]), (lo, hi)


@pytest.mark.parametrize("fmt_name", ["bf16", "f16"])
def test_16bit_formats_identical_in_both_orders(fmt_name):
    torch = _torch()
    from beampower_amd import device as bpd
    fmt = {"bf16": bpd.BP_BF16, "f16": bpd.BP_F16}[fmt_name]
    u, alpha, tau, beta = shuffled_network(5)
    F = oracle.prestack(u, alpha, flavour="fma")
    q = bpd.quantize(planar(torch, F), fmt)
    outs = []
    for order in ("natural", "similarity"):
        net = bpd.Network(tau, beta, order=order)
        m, a = net.beamform_max(q, 0)
        b = net.beamform(q, 1, source_begin=11, source_end=250)
        outs.append((m.cpu().numpy(), a.cpu().numpy(), b.cpu().numpy()))
    assert bits_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert bits_equal(outs[0][2], outs[1][2])


def test_differential_in_similarity_order():
    torch = _torch()
    from beampower_amd import device as bpd
    u, alpha, tau, beta = synthetic.random_case(6, S=9, C=3, P=2, T=121, K=777,
                                                tau_range=(-60, 60))
    F = oracle.prestack(u, alpha, flavour="fma")
    net = bpd.Network(tau, beta, order="similarity")
    assert net.source_order()[0]
    d = net.beamform_differential(planar(torch, F))
    ref = oracle.beamform_differential(F, tau, beta, flavour="fma")
    assert bits_equal(d.cpu().numpy(), ref)


def test_batch_of_ordered_networks_equals_natural_windows():
    torch = _torch()
    from beampower_amd import device as bpd
    wins_sim, wins_nat = [], []
    for i, T in enumerate((3000, 1700, 2600)):
        u, alpha, tau, beta = shuffled_network(20 + i, grid=(12, 10, 8), T=T)
        Fd = planar(torch, oracle.prestack(u, alpha, flavour="fma"))
        wins_sim.append((bpd.Network(tau, beta, order="similarity"), Fd, i % 2))
        wins_nat.append((bpd.Network(tau, beta, order="natural"), Fd, i % 2))
    batch = bpd.Batch(wins_sim)
    batch.launch()
    torch.cuda.synchronize()
    for i, (net, Fd, oob) in enumerate(wins_nat):
        m, a = net.beamform_max(Fd, oob)
        assert bits_equal(batch.beam_max[i].cpu().numpy(), m.cpu().numpy()), i
        assert np.array_equal(batch.beam_argmax[i].cpu().numpy(), a.cpu().numpy()), i


@pytest.mark.parametrize("oob", [0, 1])
def test_dropin_on_shuffled_sources(gpu_lib, oob):
    """The per-call drop-in on shuffled sources with ties far apart.  At this
    size its plan keeps the caller's order (an order would cost the call more
    than it saves); the similarity-order plans of larger calls are those the
    Network tests above check in every order."""
    u, alpha, tau, beta = shuffled_network(7, grid=(20, 12, 10), box_km=100.0, T=3500)
    F = oracle.prestack(u, alpha, flavour="fma")
    S, T, P = F.shape
    K = tau.shape[0]
    m = np.zeros(T, np.float32)
    a = np.full(T, -5, np.int32)
    rc = gpu_lib.beamform_max(F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I),
                              beta.ctypes.data_as(P_F), T, K, S, P, oob, m.ctypes.data_as(P_F),
                              a.ctypes.data_as(P_I))
    assert rc == 0, bp.core.last_error(gpu_lib)
    mo, ao = oracle.beamform_max(F, tau, beta, oob, flavour="fma")
    assert bits_equal(m, mo)
    assert np.array_equal(a, ao)
    sub = slice(500, 1012)
    t_, b_ = np.ascontiguousarray(tau[sub]), np.ascontiguousarray(beta[sub])
    b = np.full((t_.shape[0], T), 3.0, np.float32)
    rc = gpu_lib.beamform(F.ctypes.data_as(P_F), t_.ctypes.data_as(P_I), b_.ctypes.data_as(P_F),
                          T, t_.shape[0], S, P, oob, b.ctypes.data_as(P_F))
    assert rc == 0, bp.core.last_error(gpu_lib)
    assert bits_equal(b, oracle.beamform(F, t_, b_, oob, flavour="fma"))
