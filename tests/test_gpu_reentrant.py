"""Re-entrancy of the drop-in C ABI: ctypes releases the GIL during a call,
so Python threads can be inside the library together.  Every call uses its
own stream and buffers, and calls share the kept network only when their
moveouts and weights are equal; concurrent calls must give the same bits as
the same calls made one at a time."""
import ctypes as ct
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import bits_equal
from oracle import oracle
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu

P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)


def call(lib, kind, F, tau, beta, oob):
    S, T, P = F.shape
    K = tau.shape[0]
    args = (F.ctypes.data_as(P_F), tau.ctypes.data_as(P_I), beta.ctypes.data_as(P_F), T, K, S, P)
    if kind == "max":
        m = np.zeros(T, np.float32)
        a = np.zeros(T, np.int32)
        rc = lib.beamform_max(*args, oob, m.ctypes.data_as(P_F), a.ctypes.data_as(P_I))
        out = (m, a)
    elif kind == "none":
        b = np.zeros((K, T), np.float32)
        rc = lib.beamform(*args, oob, b.ctypes.data_as(P_F))
        out = (b,)
    else:
        b = np.zeros(K, np.float32)
        rc = lib.beamform_differential(*args, 1, b.ctypes.data_as(P_F))
        out = (b,)
    assert rc == 0
    return out


def test_concurrent_calls_match_sequential(gpu_lib):
    jobs = []
    for i in range(12):
        kind = ("max", "none", "diff")[i % 3]
        T = 2000 + 137 * i
        if kind == "diff":
            half = T // 2
            u, alpha, tau, beta = synthetic.random_case(70 + i, S=9, C=3, P=2, T=T, K=900,
                                                        tau_range=(-half, half))
        else:
            u, alpha, tau, beta = synthetic.random_case(70 + i, S=9, C=3, P=2, T=T, K=900,
                                                        tau_range=(0, 250))
        F = oracle.prestack(u, alpha)
        jobs.append((kind, F, tau, beta, i % 2))
    seq = [call(gpu_lib, *j) for j in jobs]
    with ThreadPoolExecutor(max_workers=6) as ex:
        for rep in range(2):
            par = list(ex.map(lambda j: call(gpu_lib, *j), jobs))
            for i, (x, y) in enumerate(zip(seq, par)):
                for u, v in zip(x, y):
                    assert (bits_equal(u, v) if u.dtype == np.float32
                            else np.array_equal(u, v)), (rep, i, jobs[i][0])
    # and the sequential results are the oracle's
    for (kind, F, tau, beta, oob), out in zip(jobs[:3], seq[:3]):
        if kind == "max":
            mo, ao = oracle.beamform_max(F, tau, beta, oob)
            assert bits_equal(out[0], mo) and np.array_equal(out[1], ao)
        elif kind == "none":
            assert bits_equal(out[0], oracle.beamform(F, tau, beta, oob))
        else:
            assert bits_equal(out[0], oracle.beamform_differential(F, tau, beta))


def test_device_api_on_a_side_stream(gpu_lib):
    """Device-API calls on a stream other than the current one, with the
    wrapper's default outputs and temporaries (zero-filled on that stream,
    marked as used there), give the current-stream results while other work
    allocates and fills memory on the current stream."""
    import torch
    from beampower_amd import device as bpd
    u, alpha, tau, beta = synthetic.random_case(90, S=10, C=3, P=2, T=6000, K=2000,
                                                tau_range=(0, 300))
    net = bpd.Network(tau, beta)
    F = bpd.prestack_planar(torch.from_numpy(u).cuda(), torch.from_numpy(alpha).cuda())
    m0, a0 = net.beamform_max(F, 1)
    b0 = net.beamform(F, 1, 0, 300)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    outs, junk = [], []
    for rep in range(3):
        Ft = F.transpose(1, 2).contiguous().transpose(1, 2)  # a non-contiguous view
        m, a = net.beamform_max(Ft, 1, stream=side)
        b = net.beamform(Ft, 1, 0, 300, stream=side)
        del Ft
        junk += [torch.full((1 << 20,), 7.0, device="cuda") for _ in range(4)]
        outs.append((m, a, b))
    side.synchronize()
    torch.cuda.synchronize()
    for m, a, b in outs:
        assert bits_equal(m.cpu().numpy(), m0.cpu().numpy())
        assert np.array_equal(a.cpu().numpy(), a0.cpu().numpy())
        assert bits_equal(b.cpu().numpy(), b0.cpu().numpy())


def test_concurrent_large_transfers(gpu_lib):
    """Calls whose inputs or outputs reach 32 MiB cross PCIe through the
    two pinned staging stages; with four such calls at once two of them find
    both stages taken and copy from pageable memory instead.  Every result
    must equal the same call made alone."""
    jobs = []
    for i in range(4):
        # reduce='none' with a 43 MB output (1,200 sources x 9,000 samples)
        u, alpha, tau, beta = synthetic.random_case(90 + i, S=9, C=3, P=2, T=9000, K=1200,
                                                    tau_range=(0, 300))
        jobs.append(("none", oracle.prestack(u, alpha), tau, beta, i % 2))
        # reduce='max' with 35 MB of features (40 stations x 110,000 samples)
        u, alpha, tau, beta = synthetic.random_case(95 + i, S=40, C=3, P=2, T=110_000, K=300,
                                                    tau_range=(0, 400))
        jobs.append(("max", oracle.prestack(u, alpha), tau, beta, i % 2))
    seq = [call(gpu_lib, *j) for j in jobs]
    with ThreadPoolExecutor(max_workers=8) as ex:
        par = list(ex.map(lambda j: call(gpu_lib, *j), jobs))
    for i, (x, y) in enumerate(zip(seq, par)):
        for u, v in zip(x, y):
            assert (bits_equal(u, v) if u.dtype == np.float32 else np.array_equal(u, v)), \
                (i, jobs[i][0])
    kind, F, tau, beta, oob = jobs[0]
    assert bits_equal(seq[0][0], oracle.beamform(F, tau, beta, oob))
    kind, F, tau, beta, oob = jobs[1]
    mo, ao = oracle.beamform_max(F, tau, beta, oob)
    assert bits_equal(seq[1][0], mo) and np.array_equal(seq[1][1], ao)
    gpu_lib.bp_dropin_cache_clear()
