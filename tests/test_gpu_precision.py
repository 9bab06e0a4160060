"""GPU checks of the reduced-precision feature forms (BF16 and FP16 storage,
FP32 accumulation).

* The quantizer matches a NumPy rounding of F * scale (round to nearest
  even), with per-station power-of-two scales.
* The 16-bit kernels are checked bit for bit against the FP32 oracle run on
  the dequantized features and the folded weights: storage is the only
  approximation, the FMA chain is the reference's.
* Against the FP64 oracle on the plain-method workload: beam error and
  argmax agreement bounds.
"""
import numpy as np
import pytest

from conftest import bits_equal
from oracle import oracle
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu

BP_BF16, BP_F16 = 1, 2


def bf16_round(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def dequant(bits, fmt):
    if fmt == BP_F16:
        return bits.view(np.float16).astype(np.float32)
    return (bits.astype(np.uint32) << 16).view(np.float32)


def quantize_ref(F_spt, fmt):
    """NumPy restatement of the device quantizer: (planes, weight scales)."""
    S, P, T = F_spt.shape
    if fmt == BP_F16:
        m = np.abs(F_spt).reshape(S, -1).max(1)
        e = np.frexp(m)[1]
        sigma = np.where(m > 0, np.ldexp(np.float32(1), 15 - e), 1).astype(np.float32)
    else:
        sigma = np.ones(S, np.float32)
    x = F_spt * sigma[:, None, None]
    q = x.astype(np.float16).view(np.uint16) if fmt == BP_F16 else bf16_round(x)
    return q, (1.0 / sigma).astype(np.float32)


def run_precision(F_spt, tau, beta, fmt, oob):
    import torch
    from beampower_amd import device as bpd
    net = bpd.Network(tau, beta)
    Fd = torch.from_numpy(np.ascontiguousarray(F_spt)).cuda()
    Q = bpd.quantize(Fd, fmt)
    m, a = net.beamform_max(Q, oob)
    torch.cuda.synchronize()
    return Q, m.cpu().numpy(), a.cpu().numpy()


@pytest.mark.parametrize("fmt", [BP_F16, BP_BF16])
def test_quantizer_matches_numpy(gpu_lib, fmt):
    import torch
    from beampower_amd import device as bpd
    rng = np.random.default_rng(3)
    F = (rng.standard_normal((5, 2, 1003)) * rng.uniform(1e-3, 1e4, (5, 1, 1))).astype(np.float32)
    Q = bpd.quantize(torch.from_numpy(F).cuda(), fmt)
    q = Q.data.cpu().numpy().view(np.uint16)
    q_ref, w_ref = quantize_ref(F, fmt)
    assert np.array_equal(Q.weight_scale, w_ref)
    T = F.shape[2]
    assert np.array_equal(q[:, :, 0, :T], q_ref)
    assert np.array_equal(q[:, :, 1, :T - 1], q_ref[:, :, 1:])
    assert np.all(q[:, :, 1, T - 1] == 0)


@pytest.mark.parametrize("fmt", [BP_F16, BP_BF16])
@pytest.mark.parametrize("oob", [0, 1])
def test_16bit_kernel_is_the_fp32_chain_on_dequantized_data(gpu_lib, fmt, oob):
    """Against the oracle on the dequantized features and folded weights (for
    BF16 the weights are bf16-rounded as the kernel uses them).  FP16
    (v_fma_mix_f32) is the same FMA chain: bit-exact.  BF16 accumulates with
    v_dot2_f32_bf16, whose rounding is not that of a fused FMA in every case:
    there the bound is a few float32 ulps of the sum of |terms|."""
    u, alpha, tau, beta = synthetic.random_case(61, S=9, C=3, P=2, T=2500, K=700,
                                                tau_range=(0, 400), negative_delays=bool(oob))
    F_stp = oracle.prestack(u, alpha)
    F_spt = np.ascontiguousarray(F_stp.transpose(0, 2, 1))
    Q, m, a = run_precision(F_spt, tau, beta, fmt, oob)
    q_ref, w = quantize_ref(F_spt, fmt)
    Fdq = np.ascontiguousarray(dequant(q_ref, fmt).transpose(0, 2, 1))
    bw = (beta * w[None, :]).astype(np.float32)
    if fmt == BP_BF16:
        bw = dequant(bf16_round(bw), BP_BF16)
    mo, ao = oracle.beamform_max(Fdq, tau, bw, oob)
    if fmt == BP_F16:
        assert bits_equal(m, mo)
        assert np.array_equal(a, ao)
        return
    # sum of |terms| per source and sample bounds the accumulated rounding
    mag = oracle.beamform(np.abs(Fdq), tau, np.abs(bw), 1).max(axis=0)
    tol = 64 * np.finfo(np.float32).eps * (mag + 1e-30)
    assert np.all(np.abs(m - mo) <= tol)
    b = oracle.beamform(Fdq, tau, bw, oob)
    T = b.shape[1]
    if oob == 0:
        lohi = oracle.delay_bounds(tau, beta).astype(np.int64)
        t = np.arange(T)[None, :]
        ok = (lohi[:, :1] <= lohi[:, 1:]) & (t + lohi[:, :1] >= 0) & (t + lohi[:, 1:] < T)
        b = np.where(ok, b, -np.inf)
    srt = np.sort(b, axis=0)
    with np.errstate(invalid="ignore"):
        clear = np.isfinite(srt[-1]) & ((srt[-1] - srt[-2]) > 2 * tol)
    assert np.array_equal(a[clear], ao[clear])


@pytest.mark.parametrize("fmt", [BP_F16, BP_BF16])
def test_16bit_against_fp64_on_plain_workload(gpu_lib, fmt):
    """Plain-method workload: errors of the 16-bit storage forms relative to
    the FP64 oracle, and argmax agreement wherever the FP64 top two differ
    by more than the storage's error bound."""
    w = synthetic.plain_c1()
    F_stp = oracle.prestack(w.waveform_features, w.weights_phases)
    F_spt = np.ascontiguousarray(F_stp.transpose(0, 2, 1))
    _, m, a = run_precision(F_spt, w.time_delays, w.weights_sources, fmt, 0)
    F64 = oracle.prestack_f64(w.waveform_features, w.weights_phases)
    mx, sec, arg, valid = oracle.beamform_max_f64(F64, w.time_delays,
                                                  w.weights_sources.astype(np.float64), 0)
    rel = np.abs(m[valid] - mx[valid]) / np.abs(mx[valid])
    unit = 2.0 ** -11 if fmt == BP_F16 else 2.0 ** -8  # half-ulp bound of the storage
    assert rel.max() < 4 * unit, rel.max()
    gap = (mx - sec) / np.abs(mx)
    clear = valid & (gap > 4 * unit)
    agree = np.mean(a[clear] == arg[clear])
    assert agree == 1.0, agree
    for k_e, t_e in w.events:
        assert np.array_equal(w.time_delays[a[t_e]], w.time_delays[k_e])
