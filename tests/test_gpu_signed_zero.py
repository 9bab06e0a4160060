"""Beams that end at a signed zero, on the GPU, bit for bit.

Tiny features and weights (about 1e-30) make products underflow to a zero
that carries the product's sign, so a beam can end at +0.0 or -0.0.  The
reference skips unused stations and, in flexible mode, out-of-range terms
(beamform.c:61-62,94-99), which keeps a zero's sign, and its '>' ties -0.0
with +0.0 (beamform.c:326), so the lower index wins with its own sign.
Checked against the reference's golden fixture (through the C ABI, in
test_gpu_parity) and here against the oracle on larger inputs that put tiles
in every kernel class, through the drop-in entry points and resident
networks with padded (staged by copies) and unpadded (guarded staging) rows.
"""
import numpy as np
import pytest

from conftest import bits_equal
from make_golden import signed_zero_case
from oracle import oracle

from test_gpu_parity import abi_max, abi_none

pytestmark = pytest.mark.gpu


def _zero_signs(x):
    return int(np.sum((x == 0) & np.signbit(x))), int(np.sum((x == 0) & ~np.signbit(x)))


@pytest.mark.parametrize("T,K,spread", [
    (1500, 48 * 7, "small"),      # several source and time tiles, small LDS window
    (4600, 48 * 24, "classes"),   # small and large LDS windows and direct gathers
])
def test_signed_zero_vs_oracle(gpu_lib, T, K, spread):
    u, alpha, tau, beta = signed_zero_case(31, S=7, C=2, P=2, T=T, K=K, tau_max=20)
    if spread == "classes":
        rng = np.random.default_rng(32)
        tau[48 * 5:48 * 6] = rng.integers(0, 1400, (48, 7, 2))    # large window
        tau[48 * 12:48 * 13] = rng.integers(0, 4000, (48, 7, 2))  # direct gathers
    F = oracle.prestack(u, alpha)
    for oob in (0, 1):
        mo, ao = oracle.beamform_max(F, tau, beta, oob)
        assert sum(_zero_signs(mo)) > 0
        m, a = abi_max(gpu_lib, F, tau, beta, oob)
        assert bits_equal(m, mo), (oob, _zero_signs(m), _zero_signs(mo))
        assert np.array_equal(a, ao), oob
        bo = oracle.beamform(F, tau, beta, oob)
        b = abi_none(gpu_lib, F, tau, beta, oob)
        assert bits_equal(b, bo), (oob, _zero_signs(b), _zero_signs(bo))


def test_signed_zero_resident_views(gpu_lib):
    """Resident network, padded and unpadded float32 rows (the flexible edge
    tiles then stage by copies and by the guarded path); the source order of
    the plan may differ from the caller's."""
    import torch
    from beampower_amd import device as bpd
    u, alpha, tau, beta = signed_zero_case(33, S=6, C=2, P=2, T=1300, K=48 * 5, tau_max=30)
    F = oracle.prestack(u, alpha)
    Fd = torch.from_numpy(np.ascontiguousarray(F.transpose(0, 2, 1))).cuda()
    for order in ("natural", "similarity"):
        net = bpd.Network(tau, beta, order=order)
        for pad in (net.padding(), (0, 0)):
            V = bpd.pack(Fd, bpd.BP_F32, *pad)
            for oob in (0, 1):
                mo, ao = oracle.beamform_max(F, tau, beta, oob)
                m, a = net.beamform_max(V, oob)
                assert bits_equal(m.cpu().numpy(), mo), (order, pad, oob)
                assert np.array_equal(a.cpu().numpy(), ao), (order, pad, oob)
                bo = oracle.beamform(F, tau, beta, oob)
                b = net.beamform(V, oob).cpu().numpy()
                assert bits_equal(b, bo), (order, pad, oob)
    torch.cuda.synchronize()
