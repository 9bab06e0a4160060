"""bp_plan (host-only) predicts the plan the engine builds: for networks of
several shapes and source orders, the tile classes, LDS windows, row
padding and source-order decision of bp_plan equal those of the network
bp_network_create_ordered makes on the device."""
import numpy as np
import pytest

from beampower_amd import plan as bpp
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu


def networks():
    yield "plain", synthetic.plain_c1().time_delays, None
    _, _, tau = synthetic.make_network(3, 16, 90.0, (24, 20, 10))
    yield "grid", tau, None
    yield "shuffled", np.ascontiguousarray(tau[np.random.default_rng(1).permutation(len(tau))]), None
    u, alpha, tau, beta = synthetic.random_case(7, S=9, C=3, P=3, T=100, K=48 * 30,
                                                tau_range=(0, 700))
    beta[100, 2] = np.inf  # a direct tile
    yield "wide_p3", tau, beta


@pytest.mark.parametrize("order", ["natural", "similarity", "auto"])
def test_plan_matches_the_built_network(gpu_lib, order):
    from beampower_amd import device as bpd
    for name, tau, beta in networks():
        if beta is None:
            beta = np.ones(tau.shape[:2], np.float32)
        K, S, P = tau.shape
        p = bpp.plan(K, S, P, 9000, time_delays=tau, weights_sources=beta, order=order)
        net = bpd.Network(tau, beta, order=order)
        info = net.info()
        assert p["tiles_per_class"] == list(info["tiles_per_class"]), (name, order)
        assert p["window_per_class"] == list(info["window_per_class"][:2]), (name, order)
        assert (p["pad_before"], p["pad_after"]) == net.padding(), (name, order)
        assert p["similarity_order"] == int(net.source_order()[0]), (name, order)
        net.close()
