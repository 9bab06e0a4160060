"""bp_network_beamform_max_noalloc: reduce='max' whose key scratch is the
library's own device array instead of a caller workspace, through the C
ABI.  Its results must equal bp_network_beamform_max_view's (the allocating
entry point with a workspace) bit for bit, on consecutive calls that reuse
the array -- the second call shorter than the first, so that stale keys of
the first would show -- in every storage format; oversized windows and
missing schedules are refused before anything is launched."""
import ctypes as ct

import numpy as np
import pytest

from conftest import bits_equal
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu


def _noalloc(lib, net, view, oob, T):
    import torch
    m = torch.zeros(T, dtype=torch.float32, device="cuda")
    a = torch.full((T,), -7, dtype=torch.int32, device="cuda")
    cv = view.c_view()
    rc = lib.bp_network_beamform_max_noalloc(net._h, ct.byref(cv), int(oob), ct.c_void_p(
        m.data_ptr()), ct.c_void_p(a.data_ptr()), None)
    return rc, m, a


def test_noalloc_matches_the_workspace_entry_point(gpu_lib):
    import torch
    from beampower_amd import device as bpd
    u, alpha, tau, beta = synthetic.random_case(401, S=10, C=3, P=2, T=20000, K=3000,
                                                tau_range=(0, 300))
    net = bpd.Network(tau, beta)
    F = bpd.prestack_planar(torch.from_numpy(u).cuda(), torch.from_numpy(alpha).cuda())
    pad = net.padding()
    cases = []
    for fmt in (bpd.BP_F32, bpd.BP_BF16, bpd.BP_F16):
        long_view = bpd.pack(F, fmt, *pad)
        short_view = bpd.pack(F[:, :, :7000].contiguous(), fmt, *pad)
        cases += [(long_view, 0), (short_view, 1)]
    for i, (view, oob) in enumerate(cases):
        T = view.n_samples
        ref_m, ref_a = net.beamform_max(view, oob)  # builds the format's schedule
        torch.cuda.synchronize()
        rc, m, a = _noalloc(gpu_lib, net, view, oob, T)
        assert rc == 0, gpu_lib.bp_last_error().decode()
        torch.cuda.synchronize()
        assert bits_equal(m.cpu().numpy(), ref_m.cpu().numpy()), (i, view.fmt, T)
        assert np.array_equal(a.cpu().numpy(), ref_a.cpu().numpy()), (i, view.fmt, T)
    # two calls in a row, the second shorter, with no other call between them
    rc1, m1, a1 = _noalloc(gpu_lib, net, cases[0][0], 0, cases[0][0].n_samples)
    rc2, m2, a2 = _noalloc(gpu_lib, net, cases[1][0], 1, cases[1][0].n_samples)
    assert rc1 == 0 and rc2 == 0
    ref1 = net.beamform_max(cases[0][0], 0)
    ref2 = net.beamform_max(cases[1][0], 1)
    torch.cuda.synchronize()
    assert bits_equal(m1.cpu().numpy(), ref1[0].cpu().numpy())
    assert np.array_equal(a1.cpu().numpy(), ref1[1].cpu().numpy())
    assert bits_equal(m2.cpu().numpy(), ref2[0].cpu().numpy())
    assert np.array_equal(a2.cpu().numpy(), ref2[1].cpu().numpy())


def test_noalloc_refuses_before_launching(gpu_lib):
    import torch
    from beampower_amd import device as bpd
    from beampower_amd.core import FeatureViewC
    u, alpha, tau, beta = synthetic.random_case(402, S=6, C=3, P=2, T=3000, K=500,
                                                tau_range=(0, 100))
    net = bpd.Network(tau, beta)
    F = bpd.prestack_planar(torch.from_numpy(u).cuda(), torch.from_numpy(alpha).cuda())
    V = bpd.pack(F, bpd.BP_F32)
    m = torch.zeros(8, dtype=torch.float32, device="cuda")
    a = torch.zeros(8, dtype=torch.int32, device="cuda")
    limit = 4194304
    big = FeatureViewC(V.data.data_ptr(), bpd.BP_F32, limit + 1, limit + 1, 0, None)
    rc = gpu_lib.bp_network_beamform_max_noalloc(net._h, ct.byref(big), 0,
                                                 ct.c_void_p(m.data_ptr()),
                                                 ct.c_void_p(a.data_ptr()), None)
    assert rc == -22 and "exceeds" in gpu_lib.bp_last_error().decode()
    # FP16 scales never used with this network: no schedule, refused
    Q = bpd.pack(F * 3.0, bpd.BP_F16)
    cv = Q.c_view()
    rc = gpu_lib.bp_network_beamform_max_noalloc(net._h, ct.byref(cv), 0,
                                                 ct.c_void_p(m.data_ptr()),
                                                 ct.c_void_p(a.data_ptr()), None)
    assert rc == -22 and "no schedule" in gpu_lib.bp_last_error().decode()
    torch.cuda.synchronize()
    assert not m.any().item() and not a.any().item()
