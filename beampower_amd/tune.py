"""Tuner of the stack kernel's compile-time shape on a workload of one's own.

    python -m beampower_amd.tune [--stations 50] [--grid 80 80 30] [--box-km 100]
                                 [--samples 90000] [--fmt f32|bf16|f16] [--reps 6] [--apply]
                                 [--set shapes|tiles|schedule] [--build-dir DIR]
                                 [--build-only | --prebuilt]

The shift-and-stack kernel keeps a tile's accumulators in registers, so its
source tile (sources per wave x waves per work-group), its time tile (64 x
samples per lane), the waves per SIMD it is compiled for (and with them how
many work-groups share a CU), the gather pipeline depth, the staging
buffers, the block-order time group and the window classes (the small LDS
window's share of the tiles) are fixed at build time (macros of
bp_kernels.h / beamform_gpu.hip).  This tool builds
the library once per candidate shape into a scratch directory, loads every
build into this process, plans the same network with each (the plan depends
on the source tile), and times reduce='max' on the same device-resident,
row-padded features with HIP events, the builds interleaved over `reps`
rounds.  All builds must give bit-identical maxima and argmax.  --apply
writes the fastest shape to beampower_amd/lib/build_flags.json, which
`python -m beampower_amd.build` (and __graft_entry__.build()) then use.
--build-only builds the candidates into --build-dir on any host (hipcc
cross-compiles), and --prebuilt times those builds on a GPU host.

The run-time configuration of a batch (block order, staging buffers, grid)
is tuned separately by bp_batch_tune.
"""
import argparse
import concurrent.futures
import ctypes as ct
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

from . import build as bp_build
from .core import FeatureViewC

# Candidate shapes; the first is the sources' default.
CANDIDATES = (
    {"BP_KW": 6, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 1},
    {"BP_KW": 6, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 2},
    {"BP_KW": 5, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 1},
    {"BP_KW": 4, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 2},
    {"BP_KW": 8, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 1},
    {"BP_KW": 8, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 2},
    {"BP_KW": 6, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 1, "BP_TIME_GROUP": 4},
    {"BP_KW": 6, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 1, "BP_TIME_GROUP": 16},
    # 1,024-sample time tiles (16 samples per lane): half the address adds per FMA
    {"BP_KW": 4, "BP_NACC": 16, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 1},
    {"BP_KW": 3, "BP_NACC": 16, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 1},
    # window classes: the small window's share of the tiles
    {"BP_KW": 6, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 1, "BP_SMALL_SHARE_PCT": 90},
    {"BP_KW": 6, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 1, "BP_SMALL_SHARE_PCT": 100},
)
# Source tiles (sources per wave x waves per work-group) and time tiles
# (64 x samples per lane) toward the batched workload's ranges, as far as
# they stay register-resident and fit a work-group: a lane holds
# KW x NACC <= 64 accumulators (48-64 VGPRs of the 80 or 128 a wave has at 6
# or 4 waves per SIMD), a work-group holds at most 16 waves (1,024 threads),
# and the LDS window of one station (two staging buffers) fits a CU's
# 160 KiB for windows up to 2,048 samples.  So source tiles of 32-128 and
# time tiles of 512 or 1,024 samples; wider tiles would spill accumulators
# or exceed a work-group.
TILE_CANDIDATES = (
    {"BP_KW": 6, "BP_W": 8, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 1},      # 48 x 512
    {"BP_KW": 6, "BP_W": 12, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 1},     # 72 x 512
    {"BP_KW": 8, "BP_W": 8, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 1},      # 64 x 512
    {"BP_KW": 8, "BP_W": 12, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 1},     # 96 x 512
    {"BP_KW": 8, "BP_W": 16, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 1},     # 128 x 512
    {"BP_KW": 4, "BP_W": 16, "BP_WAVES_PER_EU": 6, "BP_PIPE_DEPTH": 1},     # 64 x 512
    {"BP_KW": 4, "BP_W": 8, "BP_NACC": 16, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 1},   # 32 x 1024
    {"BP_KW": 3, "BP_W": 16, "BP_NACC": 16, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 1},  # 48 x 1024
    {"BP_KW": 4, "BP_W": 16, "BP_NACC": 16, "BP_WAVES_PER_EU": 4, "BP_PIPE_DEPTH": 1},  # 64 x 1024
)
# Where each wave issues its LDS reads and its global->LDS copies among its
# products (the order of one instruction stream, results unchanged): the
# FP32 reads of the next step before product words a and b of a step's 8,
# and a later station's copies at step c of the station's gather steps.
SCHEDULE_CANDIDATES = tuple(
    {"BP_GATHER_AT_F32_0": a, "BP_GATHER_AT_F32_1": b, "BP_DMA_AT": c}
    for a, b, c in ((3, 5, 4), (0, 0, 0), (4, 4, 4), (3, 5, 0), (3, 5, 6), (2, 6, 4), (4, 6, 4),
                    (3, 4, 4)))
CANDIDATE_SETS = {"shapes": CANDIDATES, "tiles": TILE_CANDIDATES,
                  "schedule": SCHEDULE_CANDIDATES}
FORMATS = {"f32": 0, "bf16": 1, "f16": 2}


def name_of(macros):
    return ",".join(f"{k[3:].lower()}={v}" for k, v in sorted(macros.items()))


def build_all(candidates, out_dir, jobs=4, log=print, prebuilt=False):
    """Build one library per candidate into out_dir (or, with prebuilt, use
    the libraries already there); returns the paths."""
    def one(i):
        path = os.path.join(out_dir, f"beamform_gpu_{name_of(candidates[i])}.so")
        if prebuilt:
            if not os.path.exists(path):
                raise RuntimeError(f"{path} is not built (run with --build-only first)")
            return path
        res = subprocess.run(bp_build.command(path, bp_build.shape_flags(candidates[i])),
                             capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"build of {name_of(candidates[i])} failed:\n{res.stderr[-2000:]}")
        return path
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        paths = list(ex.map(one, range(len(candidates))))
    log(f"built {len(paths)} candidates")
    return paths


class Candidate:
    """One build loaded into this process, its network plan and buffers."""

    def __init__(self, path, tau, beta, F, fmt):
        import torch
        vp, sz = ct.c_void_p, ct.c_size_t
        lib = ct.CDLL(path)
        lib.bp_last_error.restype = ct.c_char_p
        lib.bp_network_create.argtypes = [vp, vp, sz, sz, sz, ct.POINTER(vp)]
        lib.bp_network_destroy.argtypes = [vp]
        lib.bp_network_padding.argtypes = [vp, ct.POINTER(sz), ct.POINTER(sz)]
        lib.bp_pack_features.argtypes = [vp, sz, sz, sz, ct.c_int, sz, sz, vp, vp, vp]
        lib.bp_network_beamform_max_view.argtypes = [vp, ct.POINTER(FeatureViewC), ct.c_int,
                                                     vp, vp, vp, vp]
        self.lib = lib
        K, S, P = tau.shape
        T = F.shape[2]
        self.h = vp()
        self._check(lib.bp_network_create(tau.ctypes.data, beta.ctypes.data, K, S, P,
                                          ct.byref(self.h)))
        pb, pa = sz(), sz()
        self._check(lib.bp_network_padding(self.h, ct.byref(pb), ct.byref(pa)))
        before = pb.value + (pb.value & 1 if fmt else 0)
        stride = before + T + pa.value
        stride += stride & 1 if fmt else 0
        self.scale = np.ones(S, np.float32)
        if fmt == 0:
            self.buf = torch.zeros(S * P * stride, dtype=torch.float32, device=F.device)
        else:
            self.buf = torch.zeros(S * P * 2 * stride, dtype=torch.int16, device=F.device)
        self._check(lib.bp_pack_features(F.data_ptr(), S, P, T, fmt, before, stride,
                                         self.scale.ctypes.data if fmt else None,
                                         self.buf.data_ptr(), None))
        v = FeatureViewC()
        v.data, v.fmt, v.n_samples = self.buf.data_ptr(), fmt, T
        v.row_stride, v.pad_before = stride, before
        v.station_scale = self.scale.ctypes.data if fmt else None
        self.view = v
        self.bmax = torch.zeros(T, dtype=torch.float32, device=F.device)
        self.barg = torch.empty(T, dtype=torch.int32, device=F.device)
        self.keys = torch.empty(T, dtype=torch.int64, device=F.device)
        torch.cuda.synchronize()

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.bp_last_error().decode())

    def run(self):
        self._check(self.lib.bp_network_beamform_max_view(
            self.h, ct.byref(self.view), 0, self.bmax.data_ptr(), self.barg.data_ptr(),
            self.keys.data_ptr(), None))

    def close(self):
        if self.h:
            self.lib.bp_network_destroy(self.h)
            self.h = None


def time_all(cands, reps, calls=3):
    """Median milliseconds per call of each candidate, interleaved rounds."""
    import torch
    for c in cands:
        c.run()
    torch.cuda.synchronize()
    times = [[] for _ in cands]
    for _ in range(reps):
        for i, c in enumerate(cands):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(calls):
                c.run()
            e1.record()
            torch.cuda.synchronize()
            times[i].append(e0.elapsed_time(e1) / calls)
    return [float(np.median(t)) for t in times]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--stations", type=int, default=50)
    ap.add_argument("--grid", type=int, nargs=3, default=(80, 80, 30))
    ap.add_argument("--box-km", type=float, default=100.0)
    ap.add_argument("--samples", type=int, default=90_000)
    ap.add_argument("--fmt", choices=sorted(FORMATS), default="f32")
    ap.add_argument("--reps", type=int, default=6)
    ap.add_argument("--jobs", type=int, default=4, help="parallel builds")
    ap.add_argument("--set", choices=sorted(CANDIDATE_SETS), default="shapes",
                    help="candidate set: kernel shapes around the default, source/time tiles, "
                         "or where the reads and copies go among the products")
    ap.add_argument("--build-dir", default=None,
                    help="keep the candidate builds in this directory")
    ap.add_argument("--build-only", action="store_true",
                    help="build the candidates into --build-dir and stop (no GPU needed)")
    ap.add_argument("--prebuilt", action="store_true",
                    help="time the builds already in --build-dir instead of building")
    ap.add_argument("--apply", action="store_true", help="record the fastest shape for build()")
    args = ap.parse_args(argv)
    cands_m = CANDIDATE_SETS[args.set]
    log = lambda m: print(f"[tune] {m}", flush=True)
    if args.build_only or args.prebuilt:
        if not args.build_dir:
            sys.exit("--build-only and --prebuilt need --build-dir")
        os.makedirs(args.build_dir, exist_ok=True)
    if args.build_only:
        build_all(cands_m, args.build_dir, args.jobs, log)
        return None
    import torch
    from . import synthetic
    if not torch.cuda.is_available():
        sys.exit("the tuner times builds on the GPU: no GPU is visible")
    _, _, tau = synthetic.make_network(2, args.stations, args.box_km, tuple(args.grid))
    beta = np.ones(tau.shape[:2], np.float32)
    rng = np.random.default_rng(2)
    S, P = args.stations, tau.shape[2]
    u = torch.from_numpy(synthetic.boxcar_envelopes(rng, S, 3, args.samples)).cuda()
    alpha = torch.from_numpy(synthetic.tutorial_phase_weights(S)).cuda()
    F = (u[:, :, None, :] * alpha[:, :, :, None]).sum(1).contiguous()  # (S, P, T)
    del u
    log(f"workload: {tau.shape[0]} sources, {S} stations, {args.samples} samples, {args.fmt}")
    with tempfile.TemporaryDirectory(prefix="bp_tune_") as tmp:
        d = args.build_dir or tmp
        paths = build_all(cands_m, d, args.jobs, log, prebuilt=args.prebuilt)
        cands = []
        for p, m in zip(paths, cands_m):
            cands.append(Candidate(p, tau, beta, F, FORMATS[args.fmt]))
            log(f"planned {name_of(m)}")
        ms = time_all(cands, args.reps)
        ref = (cands[0].bmax.cpu().numpy().view(np.uint32), cands[0].barg.cpu().numpy())
        same = [np.array_equal(c.bmax.cpu().numpy().view(np.uint32), ref[0])
                and np.array_equal(c.barg.cpu().numpy(), ref[1]) for c in cands]
        for c in cands:
            c.close()
    work = tau.shape[0] * args.samples
    rows = sorted(zip(ms, cands_m, same), key=lambda r: r[0])
    for t, m, ok in rows:
        print(f"{name_of(m):48s} {t:9.3f} ms  {work / (t * 1e-3):.3e} source·samples/s"
              f"{'' if ok else '  RESULTS DIFFER'}", flush=True)
    if not all(same):
        sys.exit("builds disagree: not applying")
    best = rows[0][1]
    if args.apply:
        os.makedirs(bp_build.LIB_DIR, exist_ok=True)
        with open(bp_build.FLAGS_PATH, "w") as f:
            json.dump({"macros": best, "ms": rows[0][0],
                       "workload": {"stations": S, "grid": list(args.grid),
                                    "samples": args.samples, "fmt": args.fmt}}, f, indent=1)
        log(f"wrote {bp_build.FLAGS_PATH}: rebuild with `python -m beampower_amd.build`")
    return rows


if __name__ == "__main__":
    main()
