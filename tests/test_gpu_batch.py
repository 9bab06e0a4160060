"""Batched form on the GPU: one persistent launch over windows of several
resident networks (bp_batch_*), against each window computed alone with
bp_network_beamform_max.

Bar: bit-identical beam_max and beam_argmax per window (the batch runs the
same tile code on the same tiles; the max is an order-independent packed-key
atomicMax), for the plain launch, every hipGraph replay and every tuned
configuration.
"""
import numpy as np
import pytest

from conftest import bits_equal
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _planar(torch, bpd, u, alpha):
    return bpd.prestack_planar(torch.from_numpy(np.ascontiguousarray(u)).cuda(),
                               torch.from_numpy(np.ascontiguousarray(alpha)).cuda())


def make_windows(fmt, seed0=100):
    """Windows of plain-method networks with uneven stations, sources and
    lengths, plus random networks whose tiles need the large LDS window and
    the direct kernel, plus a second window on the first network."""
    torch = _torch()
    from beampower_amd import device as bpd
    wins = []
    for i in range(6):
        w = synthetic.batched_c4_window(seed0 + i, minutes=(1.0, 4.0), stations=(10, 30),
                                        sources=(500, 6000))
        F = _planar(torch, bpd, w.waveform_features, w.weights_phases)
        wins.append([bpd.Network(w.time_delays, w.weights_sources), F, i % 2])
    for seed, spread, T in ((7, 1200, 5000), (8, 3000, 7000)):
        u, alpha, tau, beta = synthetic.random_case(seed, S=9, C=3, P=2, T=T, K=300,
                                                    tau_range=(0, spread))
        wins.append([bpd.Network(tau, beta), _planar(torch, bpd, u, alpha), 1])
    u0 = synthetic.boxcar_envelopes(np.random.default_rng(5), wins[0][0].n_stations, 3, 2100)
    F0 = _planar(torch, bpd, u0, synthetic.tutorial_phase_weights(wins[0][0].n_stations))
    wins.append([wins[0][0], F0, 0])
    if fmt != bpd.BP_F32:
        for w in wins:
            w[1] = bpd.quantize(w[1], fmt)
    return [tuple(w) for w in wins]


def alone(wins):
    out = []
    for net, feats, oob in wins:
        m, a = net.beamform_max(feats, oob)
        out.append((m.cpu().numpy(), a.cpu().numpy()))
    return out


def assert_batch_equals(batch, ref, what):
    for i, (m, a) in enumerate(ref):
        assert bits_equal(batch.beam_max[i].cpu().numpy(), m), (what, i)
        assert np.array_equal(batch.beam_argmax[i].cpu().numpy(), a), (what, i)


@pytest.mark.parametrize("fmt_name", ["f32", "bf16"])
def test_batch_matches_windows_alone(fmt_name):
    torch = _torch()
    from beampower_amd import device as bpd
    fmt = {"f32": bpd.BP_F32, "bf16": bpd.BP_BF16}[fmt_name]
    wins = make_windows(fmt)
    classes = [net.info()["tiles_per_class"] for net, _, _ in wins]
    assert any(c[1] for c in classes) and any(c[2] for c in classes), classes
    ref = alone(wins)
    batch = bpd.Batch(wins)
    batch.launch()
    torch.cuda.synchronize()
    assert_batch_equals(batch, ref, "launch")
    for rep in range(3):
        for m in batch.beam_max:
            m.zero_()
        batch.graph_launch()
        torch.cuda.synchronize()
        assert_batch_equals(batch, ref, f"graph replay {rep}")
    for bpc in (-1, 1, 0):  # each window alone, persistent, one work-group per item
        batch.set_config(blocks_per_cu=bpc)
        for m in batch.beam_max:
            m.zero_()
        batch.graph_launch()
        torch.cuda.synchronize()
        assert_batch_equals(batch, ref, f"blocks_per_cu {bpc}")
    best, tried = batch.tune(reps=1)
    assert len(tried) == 40 and all(t["ms"] > 0 for t in tried)
    assert best["ms"] == min(t["ms"] for t in tried)
    assert batch.config()["time_group"] == best["time_group"]
    for m in batch.beam_max:
        m.zero_()
    batch.graph_launch()
    torch.cuda.synchronize()
    assert_batch_equals(batch, ref, f"tuned {best}")


def test_batch_f16_windows_with_different_scales_on_one_network():
    """Two FP16 windows of one network with their own station scales in one
    batch: each window uses its own schedule, so each equals its result
    alone."""
    torch = _torch()
    from beampower_amd import device as bpd
    wins = make_windows(bpd.BP_F16)
    shared = [w for w in wins if w[0] is wins[0][0]]
    assert len(shared) == 2
    ref = alone(wins)
    batch = bpd.Batch(wins)
    batch.launch()
    torch.cuda.synchronize()
    assert_batch_equals(batch, ref, "f16, one network with two scale sets")


def test_batch_f16_without_shared_networks():
    torch = _torch()
    from beampower_amd import device as bpd
    wins = make_windows(bpd.BP_F16)[:-1]
    ref = alone(wins)
    batch = bpd.Batch(wins)
    batch.launch()
    torch.cuda.synchronize()
    assert_batch_equals(batch, ref, "f16")


@pytest.mark.parametrize("fmt_name", ["f32", "bf16", "f16"])
def test_padded_views_match_the_plain_layout(fmt_name):
    """Rows padded by bp_network_padding put every LDS tile on the staged
    (interior) path; the sums are the same FMA chains, so the results are
    bit-identical to the unpadded layout, alone and batched."""
    torch = _torch()
    from beampower_amd import device as bpd
    fmt = {"f32": bpd.BP_F32, "bf16": bpd.BP_BF16, "f16": bpd.BP_F16}[fmt_name]
    wins = make_windows(bpd.BP_F32)
    if fmt == bpd.BP_F16:
        wins = wins[:-1]  # one F16 window per network (scales differ per window)
    padded, ref = [], []
    for net, F, oob in wins:
        plain = F if fmt == bpd.BP_F32 else bpd.quantize(F, fmt)
        m0, a0 = net.beamform_max(plain, oob)
        pb, pa = net.padding()
        lds_tiles = sum(net.info()["tiles_per_class"][:2])
        assert pb % 8 == 0 and (pa >= 512 or lds_tiles == 0), (pb, pa)
        v = bpd.pack(F, fmt, pb, pa)
        if fmt != bpd.BP_F32:
            assert np.array_equal(v.weight_scale, plain.weight_scale)
        m1, a1 = net.beamform_max(v, oob)
        m0, a0, m1, a1 = (x.cpu().numpy() for x in (m0, a0, m1, a1))
        assert bits_equal(m1, m0) and np.array_equal(a1, a0), net.info()
        padded.append((net, v, oob))
        ref.append((m0, a0))
    batch = bpd.Batch(padded)
    batch.graph_launch()
    torch.cuda.synchronize()
    assert_batch_equals(batch, ref, f"padded {fmt_name}")


def test_fp16_station_scales_per_batch_and_call():
    """FP16 features of one network quantized with different per-station
    scales (the weights then differ by those scales): a batch keeps its own
    schedule, so relaunching and replaying it after batches and calls with
    other scales -- more scale sets than the network keeps -- still gives
    each window's result alone; a batch may mix windows of one network with
    different scales."""
    torch = _torch()
    from beampower_amd import device as bpd
    w = synthetic.batched_c4_window(300, minutes=(1.0, 1.5), stations=(12, 16),
                                    sources=(2000, 3000))
    net = bpd.Network(w.time_delays, w.weights_sources)
    F = _planar(torch, bpd, w.waveform_features, w.weights_phases)
    factors = (1.0, 37.0, 1e-3, 5e3, 0.3, 1e4, 3e-5)
    Q = [bpd.quantize(F * f, bpd.BP_F16) for f in factors]
    scales = {tuple(np.asarray(q.weight_scale).tolist()) for q in Q}
    assert len(scales) == len(factors), "every factor must give its own scale set"

    def alone_of(q, oob):
        m, a = net.beamform_max(q, oob)
        return m.cpu().numpy(), a.cpu().numpy()

    ref_a = [alone_of(Q[0], 0), alone_of(Q[0], 1)]
    batch_a = bpd.Batch([(net, Q[0], 0), (net, Q[0], 1)])
    batch_a.launch()
    assert_batch_equals(batch_a, ref_a, "first launch")
    ref_b = [alone_of(Q[1], 0)]
    batch_b = bpd.Batch([(net, Q[1], 0)])
    batch_b.launch()
    assert_batch_equals(batch_b, ref_b, "second batch")
    for q in Q[2:]:  # more scale sets than the network keeps
        net.beamform_max(q, 0)
    torch.cuda.synchronize()
    for what in ("relaunch", "graph", "graph again"):
        for b in (batch_a.beam_max + batch_b.beam_max):
            b.zero_()
        if what == "relaunch":
            batch_a.launch()
            batch_b.launch()
        else:
            batch_a.graph_launch()
            batch_b.graph_launch()
        torch.cuda.synchronize()
        assert_batch_equals(batch_a, ref_a, what)
        assert_batch_equals(batch_b, ref_b, what)
    mixed = bpd.Batch([(net, Q[0], 0), (net, Q[1], 0), (net, Q[3], 1)])
    mixed.launch()
    assert_batch_equals(mixed, [ref_a[0], ref_b[0], alone_of(Q[3], 1)], "mixed scales")
    torch.cuda.synchronize()
