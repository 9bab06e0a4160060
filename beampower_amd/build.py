"""Build recipe of the gfx950 library beampower_amd/lib/beamform_gpu.so.

    python -m beampower_amd.build

Compiles the HIP sources under beampower_amd/csrc with hipcc for gfx950
only.  Packed-FP32 instruction selection is disabled (-packed-fp32-ops).
The stack kernel's compile-time shape is the sources' defaults unless
beampower_amd/lib/build_flags.json (written by `python -m beampower_amd.tune
--apply`) lists macro definitions to use instead.
"""
import json
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "beamform_gpu.so")
SOURCES = [os.path.join(CSRC, "beamform_gpu.hip"), os.path.join(CSRC, "bp_server.hip")]
HEADERS = [os.path.join(CSRC, f) for f in ("bp_device.h", "bp_internal.h", "bp_kernels.h",
                                            "bp_order.h", "bp_host_pool.h")] + [
    os.path.join(REPO_DIR, "include", "beamform_gpu.h")]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    return path


FLAGS_PATH = os.path.join(LIB_DIR, "build_flags.json")
# macros of the kernel shape the tuner may set (bp_kernels.h, beamform_gpu.hip)
SHAPE_MACROS = ("BP_KW", "BP_W", "BP_WAVES_PER_EU", "BP_PIPE_DEPTH", "BP_STAGE_BUFFERS",
                "BP_TIME_GROUP", "BP_NACC", "BP_SMALL_WINDOW", "BP_SMALL_SHARE_PCT",
                "BP_GATHER_AT_F32_0", "BP_GATHER_AT_F32_1", "BP_GATHER_AT_16",
                "BP_PIPE_DEPTH_16", "BP_DMA_AT", "BP_DMA_GAP", "BP_PRIO", "BP_PRIO_PCT")


def shape_flags(macros):
    """-D flags of a {macro: int} kernel shape (unknown names are refused)."""
    out = []
    for name, value in sorted(macros.items()):
        if name not in SHAPE_MACROS:
            raise ValueError(f"{name} is not a kernel shape macro ({', '.join(SHAPE_MACROS)})")
        out.append(f"-D{name}={int(value)}")
    return out


def tuned_flags():
    """-D flags recorded by the tuner, or none."""
    if not os.path.exists(FLAGS_PATH):
        return []
    with open(FLAGS_PATH) as f:
        return shape_flags(json.load(f)["macros"])


def command(out=LIB_PATH, extra=()):
    return [
        hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        "-I", os.path.join(REPO_DIR, "include"), "-I", CSRC,
        *extra, *SOURCES, "-o", out,
    ]


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = command(tmp, tuned_flags())
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building beamform_gpu.so")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_probe(verbose=False):
    """Peak probe of the LDS and VALU pipes (measurement only, not part of
    the library): beampower_amd/lib/probe_peaks."""
    out = os.path.join(LIB_DIR, "probe_peaks")
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", os.path.join(CSRC, "probe_peaks.hip"), "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building probe_peaks")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    if "--probe" in sys.argv:
        print(build_probe(verbose=True))
