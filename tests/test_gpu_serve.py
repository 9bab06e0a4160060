"""Serving form on the GPU: resident networks, budget admission, continuous
batching (bp_server_*).

Bar: every request's beam_max / beam_argmax is bit-identical whether it is
served alone (submitted and waited for before the next) or in a mix of
concurrent requests on several networks, and equals the same window computed
by bp_network_beamform_max on features packed the same way (storage format,
rows padded for the network).
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


def networks():
    """Three plain-method networks of different size (moveouts, source
    weights, phase weights)."""
    out = []
    for seed, S, grid in ((11, 10, (12, 12, 8)), (12, 24, (20, 20, 10)), (13, 40, (24, 24, 12))):
        st, src, tau = synthetic.make_network(seed, S, 50.0, grid)
        out.append((tau, np.ones(tau.shape[:2], np.float32), synthetic.tutorial_phase_weights(S)))
    return out


def requests(nets, n, seed, lengths=(750, 6000)):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        k = int(rng.integers(len(nets)))
        T = int(rng.integers(*lengths))
        S = nets[k][0].shape[1]
        reqs.append((k, synthetic.boxcar_envelopes(rng, S, 3, T)))
    return reqs


def alone_reference(nets, reqs, fmt):
    """Each window by the device API: prestack, pack in `fmt` with the
    network's padding, bp_network_beamform_max_view."""
    torch = _torch()
    from beampower_amd import device as bpd
    dnets = [bpd.Network(tau, beta) for tau, beta, _ in nets]
    out = []
    for k, u in reqs:
        F = bpd.prestack_planar(torch.from_numpy(u).cuda(),
                                torch.from_numpy(nets[k][2]).cuda())
        v = bpd.pack(F, fmt, *dnets[k].padding())
        m, a = dnets[k].beamform_max(v, 0)
        out.append((m.cpu().numpy(), a.cpu().numpy()))
    return out


@pytest.mark.parametrize("fmt_name", ["bf16", "f32"])
def test_served_results_are_the_same_alone_and_in_the_mix(fmt_name):
    _torch()
    from beampower_amd import device as bpd
    from beampower_amd.serve import Server, poisson_run
    fmt = {"bf16": bpd.BP_BF16, "f32": bpd.BP_F32}[fmt_name]
    nets = networks()
    reqs = requests(nets, 36, seed=5)
    ref = alone_reference(nets, reqs, fmt)
    # one batch slot and a burst of arrivals: while a batch runs, the next
    # requests are admitted and wait together
    srv = Server(hbm_budget=256 << 20, max_batch=8, fmt=fmt, in_flight=1)
    try:
        ids = [srv.add_network(*n) for n in nets]
        assert ids == [0, 1, 2]
        # alone: one request on the device at a time
        alone = []
        for k, u in reqs[:12]:
            t = srv.submit(k, u)
            m, a = srv.wait(t)
            assert t.info["batch_windows"] == 1
            alone.append((m.copy(), a.copy()))
        # in the mix: Poisson arrivals much faster than service
        tickets, summary = poisson_run(srv, reqs, rate=1e6, seed=3)
        assert max(t.info["batch_windows"] for t in tickets) > 1, summary
        for i, t in enumerate(tickets):
            assert bits_equal(t.beam_max, ref[i][0]), ("mix vs device API", i)
            assert np.array_equal(t.beam_argmax, ref[i][1]), ("mix vs device API", i)
            if i < len(alone):
                assert bits_equal(t.beam_max, alone[i][0]), ("mix vs alone", i)
                assert np.array_equal(t.beam_argmax, alone[i][1]), ("mix vs alone", i)
        st = srv.stats()
        assert st["completed"] == len(reqs) + 12 and st["arena_used"] == 0
        assert st["arena_peak"] <= st["arena_bytes"]
    finally:
        srv.close()


def test_budget_admission_serialises_and_rejects_oversized_windows():
    _torch()
    from beampower_amd import device as bpd
    from beampower_amd.serve import Server, poisson_run
    nets = networks()
    reqs = requests(nets, 10, seed=9, lengths=(3000, 4000))
    ref = alone_reference(nets, reqs, bpd.BP_BF16)
    # room for about two windows of the largest network at once
    srv = Server(hbm_budget=12 << 20, max_batch=8, fmt=bpd.BP_BF16, in_flight=2)
    try:
        for n in nets:
            srv.add_network(*n)
        tickets, summary = poisson_run(srv, reqs, rate=5000.0, seed=4)
        for i, t in enumerate(tickets):
            assert bits_equal(t.beam_max, ref[i][0]) and np.array_equal(t.beam_argmax, ref[i][1])
        st = srv.stats()
        assert st["arena_peak"] <= 12 << 20
        big = srv.submit(2, np.zeros((nets[2][0].shape[1], 3, 200_000), np.float32))
        with pytest.raises(RuntimeError, match="budget"):
            srv.wait(big)
    finally:
        srv.close()
