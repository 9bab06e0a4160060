"""The batched form on windows of the batched workload's ranges (10-80
stations, 1,000-200,000 sources, 1-30 min at 25 Hz, 3 channels, 2 phases),
against the oracle window by window.

Windows: the four corners (10 and 80 stations x 1,000 and 200,000 sources,
1 and 30 min) and eight windows drawn from the full ranges, strict and
flexible bounds alternating, all in one batch (hipGraph replay).

* FP32 features: every window's beam_max and beam_argmax bit-identical to
  the oracle (beamform.c:230-336 restated) at three ranges of samples
  (start, interior, end) -- the oracle evaluates those samples of the whole
  window -- and to the window run alone.
* BF16 features: error against the FP64 oracle on the same samples (max and
  RMS relative error of beam_max, argmax agreement with FP64, overall and
  where the FP64 top-two gap exceeds 1e-4 + 4 units of bfloat16, which
  storage rounding cannot bridge).  Asserted: relative error below 4 units
  (2^-6) and agreement 1 where the gap decides.  BP_BATCH_REPORT=<file>
  writes the numbers as JSON.
"""
import json
import os

import numpy as np
import pytest

from conftest import bits_equal
from oracle import oracle
from beampower_amd import synthetic

pytestmark = pytest.mark.gpu

W = 96  # samples per checked range
BF16_UNIT = 2.0 ** -8
CORNERS = ((10, 1_000, 1.0), (80, 200_000, 30.0), (10, 200_000, 30.0), (80, 1_000, 1.0))


def c4_windows():
    wins = [synthetic.batched_c4_window(5000 + i, minutes=(m, m), stations=(s, s),
                                        sources=(k, k))
            for i, (s, k, m) in enumerate(CORNERS)]
    wins += [synthetic.batched_c4_window(5100 + i) for i in range(8)]
    return wins


def ranges(T, rng):
    mid = int(rng.integers(W, T - 2 * W))
    return ((0, W), (mid, mid + W), (T - W, T))


def fp64_report(errs, agree, decides):
    err = np.concatenate(errs)
    ag = np.concatenate(agree)
    dec = np.concatenate(decides)
    return {"beam_max_rel_err_max": float(err.max()),
            "beam_max_rel_err_rms": float(np.sqrt(np.mean(err ** 2))),
            "argmax_agreement_fp64": float(ag.mean()),
            "argmax_agreement_fp64_where_gap_decides": float(ag[dec].mean()),
            "decidable_fraction": float(dec.mean()),
            "samples": int(err.size)}


def test_batched_windows_against_the_oracle():
    import torch
    from beampower_amd import device as bpd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    wins = c4_windows()
    shapes = [w.shape for w in wins]
    assert min(s["S"] for s in shapes) == 10 and max(s["S"] for s in shapes) == 80
    assert min(s["K"] for s in shapes) <= 1_100 and max(s["K"] for s in shapes) >= 190_000
    assert max(s["T"] for s in shapes) == 45_000 and min(s["T"] for s in shapes) == 1_500
    nets, v32, v16 = [], [], []
    for w in wins:
        net = bpd.Network(w.time_delays, w.weights_sources)
        F = bpd.prestack_planar(torch.from_numpy(w.waveform_features).cuda(),
                                torch.from_numpy(w.weights_phases).cuda())
        pad = net.padding()
        nets.append(net)
        v32.append(bpd.pack(F, bpd.BP_F32, *pad))
        v16.append(bpd.pack(F, bpd.BP_BF16, *pad))
        del F
    oobs = [i % 2 for i in range(len(wins))]
    b32 = bpd.Batch(list(zip(nets, v32, oobs)))
    b16 = bpd.Batch(list(zip(nets, v16, oobs)))
    b32.graph_launch()
    b16.graph_launch()
    torch.cuda.synchronize()

    rng = np.random.default_rng(44)
    acc = {"f32": ([], [], []), "bf16": ([], [], [])}
    checked = 0
    for i, (w, net, oob) in enumerate(zip(wins, nets, oobs)):
        T = w.shape["T"]
        tau, beta = w.time_delays, w.weights_sources
        m32 = b32.beam_max[i].cpu().numpy()
        a32 = b32.beam_argmax[i].cpu().numpy()
        alone = net.beamform_max(v32[i], oob)
        assert bits_equal(m32, alone[0].cpu().numpy()), i
        assert np.array_equal(a32, alone[1].cpu().numpy()), i
        m16 = b16.beam_max[i].cpu().numpy()
        a16 = b16.beam_argmax[i].cpu().numpy()
        F = oracle.prestack(w.waveform_features, w.weights_phases, num_threads=16)
        F64 = oracle.prestack_f64(w.waveform_features, w.weights_phases, num_threads=16)
        for t0, t1 in ranges(T, rng):
            mo, ao = oracle.beamform_max(F, tau, beta, oob, num_threads=16, t_begin=t0,
                                         t_end=t1)
            assert bits_equal(m32[t0:t1], mo), (i, w.shape, oob, t0)
            assert np.array_equal(a32[t0:t1], ao), (i, w.shape, oob, t0)
            checked += t1 - t0
            mx, sec, ax, val = oracle.beamform_max_f64(F64, tau, beta.astype(np.float64), oob,
                                                       t0, t1, num_threads=16)
            with np.errstate(invalid="ignore", divide="ignore"):
                gap = (mx - sec) / np.abs(mx)
            for name, (m, a) in (("f32", (m32, a32)), ("bf16", (m16, a16))):
                e, ag, dec = acc[name]
                e.append(np.abs(m[t0:t1][val] - mx[val]) / np.abs(mx[val]))
                ag.append(a[t0:t1][val] == ax[val])
                tol = 1e-4 + (4 * BF16_UNIT if name == "bf16" else 0.0)
                dec.append(~(gap[val] <= tol))
    report = {"windows": len(wins), "samples_checked_per_form": checked,
              "shapes": [(s["S"], s["K"], s["T"]) for s in shapes],
              "f32": fp64_report(*acc["f32"]), "bf16": fp64_report(*acc["bf16"])}
    path = os.environ.get("BP_BATCH_REPORT")
    if path:
        with open(path, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps(report), flush=True)
    assert report["f32"]["argmax_agreement_fp64_where_gap_decides"] == 1.0
    assert report["bf16"]["beam_max_rel_err_max"] < 4 * BF16_UNIT
    assert report["bf16"]["argmax_agreement_fp64_where_gap_decides"] == 1.0
