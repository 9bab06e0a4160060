"""Throughput of the FP32 drop-in beamform(reduce='max') on one MI355X.

Workload (default): 50 stations in a 100x100 km box, 80x80x30 source grid
(192,000 sources, depth 0-20 km), homogeneous Vp=6.0 / Vs=3.5 km/s at 25 Hz,
1 h of features (90,000 samples) with 50 injected events, strict bounds.
One step is one beamform_max pass over all sources and samples with the
moveouts, weights and prestacked features already resident in HBM.

Prints one JSON line.  `value` is source*samples per second of the whole
step (keys init + shift-and-stack + finalize); `roofline` prices the
dominant shift-and-stack kernel by its algorithmic LDS gather bytes;
`cpu_baseline` times the reference C engine (oracle/_ref, its own sources
and flags) on this host on a bounded time slice of the same workload.
"""
import argparse
import ctypes as ct
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("source·samples/s, FP32 drop-in vs BF16 features; "
          "p50/p99 served-window latency")

# gfx950 (MI355X) ceilings.  The LDS pipe serves 16 lanes per clock at any
# width up to 16 bytes, so the gather roofline is the ds_read_b128 rate:
# 256 B/clk/CU (profiles/peaks_probe.md measures 138 TB/s of the 157 TB/s).
N_CU = 256
CLOCK_HZ = 2.4e9
LDS_BYTES_PER_CLK_CU = 256
LDS_PEAK = N_CU * LDS_BYTES_PER_CLK_CU * CLOCK_HZ  # 157.3 TB/s
HBM_PEAK = 8.0e12
VALU_FP32_PEAK_FLOPS = 157.3e12          # published, counts packed FMA
VALU_FP32_UNPACKED_FLOPS = 78.6e12       # v_fma_f32: 64 FMA/clk/CU
VALU_FMA_MEASURED = 34.3e12              # v_fma_f32 probe (profiles/peaks_probe.md)
HBM_READ_MEASURED = 6.1e12               # streaming-read probe (profiles/peaks_probe.md)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpus():
    """CPUs this process may run on (its affinity set)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_threads():
    """Threads of the CPU baseline: every CPU of the affinity set, or the
    host's CPU share when OMP_NUM_THREADS states it (a shared GPU host
    grants each GPU a share of its cores)."""
    n = host_cpus()
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, budget_s):
    """Time the reference engine (or the oracle port) on a halo-padded time
    slice with all sources.  A slice of width W + halo (halo = max moveout +
    1) computes every source at the first W samples exactly as the full call
    would; the rate counts the (source, sample) pairs the call evaluated."""
    import numpy as np
    from oracle import oracle

    kind = "reference" if oracle.ref_available("cmake") else "port"
    threads = cpu_threads()
    tau, beta = w.time_delays, w.weights_sources
    F = oracle.prestack(w.waveform_features, w.weights_phases, num_threads=threads)
    lohi = oracle.delay_bounds(tau, beta, num_threads=threads)
    halo = int(lohi[:, 1].max()) + 1
    start = 1000

    def run(W):
        Tp = W + halo
        Fs = np.ascontiguousarray(F[:, start:start + Tp, :])
        t0 = time.perf_counter()
        if kind == "reference":
            oracle.ref_beamform_max(Fs, tau, beta, 0, num_threads=threads, flavour="cmake")
        else:
            oracle.beamform_max(Fs, tau, beta, 0, num_threads=threads, flavour="fma")
        dt = time.perf_counter() - t0
        lo, hi = lohi[:, 0].astype(np.int64), lohi[:, 1].astype(np.int64)
        n_valid = np.clip(np.minimum(Tp, Tp - hi) - np.maximum(0, -lo), 0, None)
        return dt, float(n_valid.sum()), Tp

    # size the slice for about budget_s: a 16-sample probe, then up to two
    # re-sizings from the last run's rate (a short slice under-estimates it:
    # its halo costs as much as its samples)
    mean_hi = float(np.clip(lohi[:, 1], 0, None).mean())
    W_max = w.waveform_features.shape[2] - start - halo
    W = 16
    dt, work, Tp = run(W)
    for _ in range(2):
        if dt >= 0.5 * budget_s or W >= W_max:
            break
        W = max(W + 1, min(W_max, int(budget_s * (work / dt) / tau.shape[0] + mean_hi - halo)))
        dt, work, Tp = run(W)
    return {
        "value": work / dt,
        "unit": "source·samples/s",
        "cores": threads,
        "host_cpus": host_cpus(),
        "cpu_model": cpu_model(),
        "kind": kind,
        "sample": (f"beamform_max strict, all {tau.shape[0]} sources, {Tp}-sample slice "
                   f"({W} + {halo} halo) of the same features; {work:.3e} valid "
                   f"(source, sample) pairs in {dt:.2f} s; "
                   + ("reference beamform.c built by oracle/Makefile (CMake flags)"
                      if kind == "reference" else "oracle port (fma flavour)")),
    }


def profile_traffic():
    """HBM bytes per launch of the stack kernel from the committed rocprofv3
    counter summary, if present (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    path = os.path.join(REPO, "profiles", "stack_lds_dropin_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def bench_batched(args, log):
    """Batched form: uneven windows (1-30 min, 10-80 stations, 1k-200k
    sources, plain-method geometry seeded per window) of as many networks in
    one persistent launch per window class; FP32 and BF16 features; hipGraph
    replay; optional autotuning; compared with the windows launched one by
    one (bit equality and time)."""
    import numpy as np
    import torch
    from beampower_amd import device as bpd
    from beampower_amd import synthetic

    import concurrent.futures
    n = args.batch_windows
    log(f"batched form: generating {n} windows")
    t_gen = time.perf_counter()
    wins, work, fmas = [], 0, 0
    with concurrent.futures.ThreadPoolExecutor(max_workers=cpu_threads()) as ex:
        gen = ex.map(lambda i: synthetic.batched_c4_window(1000 + i), range(n))
        for w in gen:
            F = bpd.prestack_planar(torch.from_numpy(w.waveform_features).cuda(),
                                    torch.from_numpy(w.weights_phases).cuda())
            net = bpd.Network(w.time_delays, w.weights_sources)
            K, S, P = w.time_delays.shape
            work += K * F.shape[2]
            fmas += K * S * P * F.shape[2]
            wins.append((net, F, 0))
            del w
    log(f"batched form: {n} windows ready in {time.perf_counter() - t_gen:.1f} s")
    planar = wins
    wins = [(net, bpd.pack(F, bpd.BP_F32, *net.padding()), oob) for net, F, oob in planar]
    torch.cuda.synchronize()
    out = {"windows": n, "source_samples": work,
           "stations": [wi[0].n_stations for wi in wins][:8] + ["..."]}

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    def alone():
        for net, F, oob in wins:
            net.beamform_max(F, oob)

    reps = 2  # a batch of the full workload runs for seconds
    t_alone = timed(alone, reps)
    ref = []
    for net, F, oob in wins:
        m, a = net.beamform_max(F, oob)
        ref.append((m.cpu().numpy(), a.cpu().numpy()))
    b32 = bpd.Batch(wins)
    if args.tune:
        log("batched form: tuning")
        best, tried = b32.tune(reps=1)
        out["tuned"] = best
        out["tried"] = tried
    out["config"] = b32.config()
    t_launch = timed(b32.launch, reps)
    t_graph = timed(b32.graph_launch, reps)
    same = all(np.array_equal(b32.beam_max[i].cpu().numpy().view(np.uint32),
                              ref[i][0].view(np.uint32))
               and np.array_equal(b32.beam_argmax[i].cpu().numpy(), ref[i][1])
               for i in range(n))
    out["f32"] = {"ms": t_launch * 1e3, "graph_ms": t_graph * 1e3, "value": work / t_graph,
                  "gather_fma_per_s": fmas / t_graph,
                  "windows_one_by_one_ms": t_alone * 1e3,
                  "identical_to_windows_alone": bool(same)}
    log(f"batched f32: {t_graph * 1e3:.1f} ms (graph), {t_alone * 1e3:.1f} ms one by one")
    # every grid policy on the same batch (one timed replay each)
    if not args.tune:
        cfg = b32.config()
        out["grids"] = []
        for bpc in (-1, 0, 2, 3):
            b32.set_config(cfg["time_group"], cfg["stage_buffers"], bpc)
            t = timed(b32.graph_launch, 1)
            ok = all(np.array_equal(b32.beam_max[i].cpu().numpy().view(np.uint32),
                                    ref[i][0].view(np.uint32))
                     and np.array_equal(b32.beam_argmax[i].cpu().numpy(), ref[i][1])
                     for i in range(n))
            out["grids"].append({"blocks_per_cu": bpc, "graph_ms": t * 1e3,
                                 "identical_to_windows_alone": bool(ok)})
            log(f"batched f32, blocks_per_cu {bpc}: {t * 1e3:.1f} ms")
        b32.set_config(cfg["time_group"], cfg["stage_buffers"], cfg["blocks_per_cu"])
    del b32
    q = [(net, bpd.pack(F, bpd.BP_BF16, *net.padding()), oob) for net, F, oob in planar]
    b16 = bpd.Batch(q)
    t16 = timed(b16.graph_launch, reps)
    agree = float(np.mean(np.concatenate(
        [b16.beam_argmax[i].cpu().numpy() == ref[i][1] for i in range(n)])))
    out["bf16"] = {"graph_ms": t16 * 1e3, "value": work / t16,
                   "speedup_vs_f32": t_graph / t16, "argmax_agreement_vs_f32": agree}
    log(f"batched bf16: {t16 * 1e3:.1f} ms (graph)")
    del b16, q, wins, planar
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def bench_batched_mix(args, log):
    """The batched form on a serving-like mix: requests of the four serving
    networks (20/50/80/120 stations, 10k-300k sources, 30 s - 10 min
    windows, networks shared by several windows), FP32 features, in one
    batch under each grid policy -- large windows in launches of their own,
    everything fused with one work-group per work item, and one persistent
    launch of 1, 2 or 3 (the default) work-groups per CU over a work queue --
    against the windows launched one by one; all bit-identical."""
    import numpy as np
    import torch
    from beampower_amd import device as bpd
    from beampower_amd import synthetic

    nets_h = synthetic.serving_networks()
    reqs = synthetic.serving_requests(nets_h, args.mix_windows, seed=77)
    nets = [bpd.Network(t, b) for t, b, _ in nets_h]
    alphas = [torch.from_numpy(a).cuda() for _, _, a in nets_h]
    packed = {}
    wins, work = [], 0
    for k, u in reqs:
        key = (k, id(u))
        if key not in packed:
            F = bpd.prestack_planar(torch.from_numpy(u).cuda(), alphas[k])
            packed[key] = bpd.pack(F, bpd.BP_F32, *nets[k].padding())
            del F
        wins.append((nets[k], packed[key], 0))
        work += nets_h[k][0].shape[0] * u.shape[2]
    torch.cuda.synchronize()

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    def alone():
        for net, F, oob in wins:
            net.beamform_max(F, oob)

    ref = [tuple(x.cpu().numpy() for x in net.beamform_max(F, oob)) for net, F, oob in wins]
    out = {"windows": len(wins), "source_samples": work,
           "windows_one_by_one_ms": timed(alone) * 1e3, "grids": []}
    b = bpd.Batch(wins)
    for bpc in (-1, 0, 1, 2, 3):
        b.set_config(blocks_per_cu=bpc)
        t = timed(b.graph_launch)
        b.graph_launch()
        torch.cuda.synchronize()
        same = all(np.array_equal(b.beam_max[i].cpu().numpy().view(np.uint32),
                                  ref[i][0].view(np.uint32))
                   and np.array_equal(b.beam_argmax[i].cpu().numpy(), ref[i][1])
                   for i in range(len(wins)))
        out["grids"].append({"blocks_per_cu": bpc, "graph_ms": t * 1e3, "value": work / t,
                             "identical_to_windows_alone": bool(same)})
        log(f"batched mix, blocks_per_cu {bpc}: {t * 1e3:.1f} ms")
    b.set_config()
    del b, wins, packed
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def bench_serving(args, log):
    """Serving form: four resident networks (20/50/80/120 stations,
    10k/50k/192k/300k sources), window requests of 30 s - 10 min at 25 Hz on
    random networks arriving as a Poisson process at each offered rate,
    admitted by an HBM budget and continuously batched; BF16 window storage
    (the serving default) and FP32 storage beside it, on the same request
    streams.  Latency is submit -> results in host memory."""
    import numpy as np
    import torch
    from beampower_amd import device as bpd
    from beampower_amd import synthetic
    from beampower_amd.serve import Server, poisson_run

    log("serving form: building networks")
    nets = synthetic.serving_networks()
    out = {"networks": [{"stations": t.shape[1], "sources": t.shape[0]} for t, _, _ in nets],
           "hbm_budget_bytes": args.serve_budget_gb << 30, "max_batch": 64,
           "in_flight": args.serve_in_flight}
    served = {}
    for name, fmt in (("bf16", bpd.BP_BF16), ("f32", bpd.BP_F32)):
        srv = Server(hbm_budget=args.serve_budget_gb << 30, max_batch=64, fmt=fmt,
                     in_flight=args.serve_in_flight)
        for n in nets:
            srv.add_network(*n)
        warm = synthetic.serving_requests(nets, 8, seed=99)
        poisson_run(srv, warm, rate=1000.0, seed=1)
        rates = []
        for rate in args.serve_rates:
            n = int(max(16, min(600, rate * args.serve_seconds)))
            reqs = synthetic.serving_requests(nets, n, seed=int(rate))
            tickets, summary = poisson_run(srv, reqs, rate=float(rate), seed=int(rate))
            rates.append(summary)
            log(f"serving {name} {rate} req/s: p50 {summary['p50_ms']:.1f} ms, "
                f"p99 {summary['p99_ms']:.1f} ms, {summary['throughput_req_s']:.1f} req/s, "
                f"batch {summary['mean_batch_windows']:.1f}")
        served[name] = tickets  # the last rate's requests
        out[name] = {"rates": rates, "stats": srv.stats()}
        srv.close()
    if args.serve_rates:
        top16, top32 = out["bf16"]["rates"][-1], out["f32"]["rates"][-1]
        out["bf16_over_f32_source_samples_per_s_at_top_rate"] = (
            top16["source_samples_per_s"] / top32["source_samples_per_s"])
        # argmax agreement of the BF16-served windows with the FP32 drop-in
        # (every sample of 24 windows), and FP32 serving against the same
        agree, total, same32 = 0, 0, 0
        dnets = [bpd.Network(t, b) for t, b, _ in nets]
        for t16, t32 in list(zip(served["bf16"], served["f32"]))[:24]:
            k = t16.network
            F = bpd.prestack_planar(torch.from_numpy(t16.features).cuda(),
                                    torch.from_numpy(nets[k][2]).cuda())
            m, a = dnets[k].beamform_max(bpd.pack(F, bpd.BP_F32, *dnets[k].padding()), 0)
            a = a.cpu().numpy()
            agree += int((a == t16.beam_argmax).sum())
            same32 += int((a == t32.beam_argmax).sum())
            total += a.size
        out["bf16_argmax_agreement_vs_f32"] = agree / max(1, total)
        out["f32_served_argmax_equal_to_dropin"] = same32 / max(1, total)
        del dnets
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def bench_none_subset(args, log, w, F_plain):
    """reduce='none' on a seeded 4,096-source subset of the drop-in workload:
    full (4096, 90000) beams from a resident network of the subset's rows,
    features resident and padded for it.  The reference C engine's
    reduce='none' (16 threads) is timed beside it on a prefix of the
    subset."""
    import numpy as np
    import torch
    from beampower_amd import device as bpd
    from beampower_amd import synthetic

    K_all, S, P = w.time_delays.shape
    T = w.waveform_features.shape[2]
    sub = synthetic.subset_sources(K_all, 4096, seed=2)
    tau = np.ascontiguousarray(w.time_delays[sub])
    beta = np.ascontiguousarray(w.weights_sources[sub])
    K = len(sub)
    net = bpd.Network(tau, beta)
    V = bpd.pack(F_plain, bpd.BP_F32, *net.padding())
    beam = torch.zeros((K, T), dtype=torch.float32, device="cuda")
    for _ in range(max(1, args.warmup)):
        net.beamform(V, 0, beam=beam)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n = max(5, args.steps)
    e0.record()
    for _ in range(n):
        net.beamform(V, 0, beam=beam)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    out = {"sources": K, "samples": T, "value": K * T / (ms * 1e-3),
           "unit": "source·samples/s", "ms_per_step": ms,
           "output_bytes": K * T * 4}
    log(f"reduce='none' subset: {ms:.2f} ms per step")
    del net, V, beam
    torch.cuda.empty_cache()
    # the Python call with host arrays: uploads, prestack, kernels, and the
    # (K, T) beams copied back into a numpy array
    import beampower_amd as bp
    call_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        b = bp.beamform(w.waveform_features, tau, w.weights_phases, beta, device="gpu",
                        reduce="none")
        call_ms.append((time.perf_counter() - t0) * 1e3)
        del b
    out["python_call"] = {"ms": min(call_ms), "calls_ms": call_ms,
                          "note": "beampower_amd.beamform(..., reduce='none') with numpy "
                                  "inputs; the 1.47 GB of beams come back through pinned "
                                  "staging buffers"}
    log(f"reduce='none' subset, python call: {min(call_ms):.1f} ms")
    bp.load_library("gpu").bp_dropin_cache_clear()
    if not args.no_cpu_baseline:
        from oracle import oracle
        kind = "reference" if oracle.ref_available("cmake") else "port"
        threads = cpu_threads()
        F = oracle.prestack(w.waveform_features, w.weights_phases, num_threads=threads)

        def run(kk):
            t0 = time.perf_counter()
            if kind == "reference":
                oracle.ref_beamform(F, tau[:kk], beta[:kk], 0, num_threads=threads,
                                    flavour="cmake")
            else:
                oracle.beamform(F, tau[:kk], beta[:kk], 0, num_threads=threads, flavour="fma")
            return time.perf_counter() - t0

        kk = 64
        dt = run(kk)
        if dt < 0.5 * args.cpu_seconds:
            kk = max(kk + 1, min(K, int(kk * 0.5 * args.cpu_seconds / dt)))
            dt = run(kk)
        out["cpu_baseline"] = {
            "value": kk * T / dt, "unit": "source·samples/s", "cores": threads, "kind": kind,
            "sample": (f"beamform reduce='none' strict, the first {kk} sources of the subset, "
                       f"all {T} samples, {dt:.2f} s")}
    return out


def bench_differential(args, log):
    """mode='differential' at one-GPU shape: 50 stations (1,225 station
    pairs), 192,000 sources, a 1,581-sample cross-correlation window.  One
    step is one beam per source from resident tables (the wave-transposed
    tables are built on the first call, outside the timed region).  The
    kernel streams the delay and weight tables once, so it is priced by
    those HBM bytes."""
    import numpy as np
    import torch
    from beampower_amd import device as bpd
    from beampower_amd import synthetic

    log("differential form: generating workload")
    w = synthetic.differential_c6()
    S, C, T = w.waveform_features.shape
    K, _, P = w.time_delays.shape
    # the differential kernel does not use the stack plan's source order
    # (its tables are in the caller's order), so none is computed
    net = bpd.Network(w.time_delays, w.weights_sources, order="natural")
    F = bpd.prestack_planar(torch.from_numpy(w.waveform_features).cuda(),
                            torch.from_numpy(w.weights_phases).cuda())
    beam = torch.zeros(K, dtype=torch.float32, device="cuda")
    for _ in range(max(1, args.warmup)):
        net.beamform_differential(F, beam=beam)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n = max(20, args.steps)
    e0.record()
    for _ in range(n):
        net.beamform_differential(F, beam=beam)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    recovered = int(int(beam.argmax().item()) == w.events[0][0])
    # table bytes the kernel streams: two-phase delays of windows up to 32768
    # samples as one word of two int16 per (source, station pair), else int32;
    # weights as float32, or one used bit each when every nonzero weight has
    # the same value (the form the library chose is read back)
    tables = net.diff_tables()
    pairs = P == 2 and T <= 32768
    weight_bytes = K * S / 8 if tables["weight_form"] == "used_bits" else K * S * 4
    table_bytes = K * S * (4 if pairs else P * 4) + weight_bytes + K * 4
    feature_bytes = S * P * T * 4
    achieved = (table_bytes + feature_bytes) / (ms * 1e-3)
    out = {
        "workload": w.name, "stations": w.meta["n_stations"], "station_pairs": S, "phases": P,
        "sources": K, "samples": T, "value": K / (ms * 1e-3), "unit": "sources/s",
        "ms_per_step": ms, "event_recovered": recovered,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "frac_of_measured_stream": achieved / HBM_READ_MEASURED,
                     "kernel": "differential_kernel",
                     "delay_table": "int16 pairs" if pairs else "int32",
                     "weight_table": tables["weight_form"],
                     "resident_table_bytes": tables["delay_bytes"] + tables["weight_bytes"],
                     "algorithmic_bytes_per_launch": table_bytes + feature_bytes},
    }
    log(f"differential: {ms:.3f} ms per step, {achieved / 1e9:.0f} GB/s")
    del net
    # general weights: two values (1 and 0.5 on alternate station pairs), so
    # the weight table stays float32 (4 bytes per source and pair)
    beta2 = np.ascontiguousarray(w.weights_sources * np.where(np.arange(S) % 2, 0.5, 1.0)
                                 .astype(np.float32)[None, :])
    net2 = bpd.Network(w.time_delays, beta2, order="natural")
    for _ in range(max(1, args.warmup)):
        net2.beamform_differential(F, beam=beam)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(n):
        net2.beamform_differential(F, beam=beam)
    e1.record()
    torch.cuda.synchronize()
    ms2 = e0.elapsed_time(e1) / n
    tables2 = net2.diff_tables()
    table2 = K * S * (4 if pairs else P * 4) + K * S * 4 + K * 4
    achieved2 = (table2 + feature_bytes) / (ms2 * 1e-3)
    out["float32_weights"] = {
        "value": K / (ms2 * 1e-3), "unit": "sources/s", "ms_per_step": ms2,
        "weight_table": tables2["weight_form"],
        "roofline": {"bound": "hbm", "achieved": achieved2 / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved2 / HBM_PEAK,
                     "algorithmic_bytes_per_launch": table2 + feature_bytes}}
    log(f"differential, float32 weights: {ms2:.3f} ms per step, {achieved2 / 1e9:.0f} GB/s")
    del net2, F, beam
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    # the Python call with host arrays: every table crosses PCIe on each call
    import beampower_amd as bp
    call_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        bp.beamform(w.waveform_features, w.time_delays, w.weights_phases, w.weights_sources,
                    device="gpu", mode="differential", reduce="none")
        call_ms.append((time.perf_counter() - t0) * 1e3)
    out["python_call"] = {"ms": min(call_ms), "calls_ms": call_ms,
                          "note": "beampower_amd.beamform(..., mode='differential') with numpy "
                                  "inputs: delays packed into int16 pairs and uniform weights "
                                  "into used bits by the staging host threads, 0.97 GB over PCIe"}
    log(f"differential python call: {min(call_ms):.1f} ms")
    if not args.no_cpu_baseline:
        # the reference engine single-threaded (its per-source bounds are
        # shared between OpenMP threads), whole calls repeated for about
        # cpu_seconds
        from oracle import oracle
        kind = "reference" if oracle.ref_available("cmake") else "port"
        Fh = oracle.prestack(w.waveform_features, w.weights_phases, num_threads=cpu_threads())
        tau, beta = w.time_delays, w.weights_sources

        def call():
            if kind == "reference":
                oracle.ref_beamform_differential(Fh, tau, beta, num_threads=1, flavour="cmake")
            else:
                oracle.beamform_differential(Fh, tau, beta, num_threads=1, flavour="fma")

        t0 = time.perf_counter()
        call()
        first = time.perf_counter() - t0
        reps = max(1, int(args.cpu_seconds / max(first, 1e-3)))
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        dt = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {
            "value": K / dt, "unit": "sources/s", "cores": 1, "kind": kind,
            "sample": (f"beamform_differential, the whole workload ({K} sources, {S} pairs), "
                       f"{reps} calls of {dt:.2f} s, one thread; "
                       + ("reference beamform.c built by oracle/Makefile (CMake flags)"
                          if kind == "reference" else "oracle port (fma flavour)"))}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--oob", choices=["strict", "flexible"], default="strict")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-windows", type=int, default=256,
                    help="windows of the batched form (256 = the full batched workload; 0 = skip)")
    ap.add_argument("--tune", action="store_true", help="autotune the batched form first")
    ap.add_argument("--mix-windows", type=int, default=48,
                    help="windows of the batched serving-like mix (0 = skip)")
    ap.add_argument("--serve-rates", type=float, nargs="*",
                    default=[5.0, 50.0, 100.0, 125.0, 200.0],
                    help="offered request rates of the serving form (none = skip)")
    ap.add_argument("--serve-in-flight", type=int, default=2,
                    help="batches of the serving form on the device at once (1-4)")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="arrival span per offered rate")
    ap.add_argument("--serve-budget-gb", type=int, default=16)
    ap.add_argument("--no-differential", action="store_true",
                    help="skip the differential-mode form")
    ap.add_argument("--differential-only", action="store_true",
                    help="run only the differential-mode form and print its object")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from beampower_amd import device as bpd
    from beampower_amd import synthetic

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    torch.cuda.set_device(0)
    oob = 0 if args.oob == "strict" else 1
    if args.differential_only:
        print(json.dumps(bench_differential(args, log)), flush=True)
        return

    log("generating workload")
    w = synthetic.dropin_c2()
    S, C, T = w.waveform_features.shape
    K, _, P = w.time_delays.shape
    u = torch.from_numpy(w.waveform_features).cuda()
    alpha = torch.from_numpy(w.weights_phases).cuda()
    F_plain = bpd.prestack_planar(u, alpha)
    del u
    net = bpd.Network(w.time_delays, w.weights_sources)
    info = net.info()
    info["source_order"] = "similarity" if net.source_order()[0] else "natural"
    shape = [ct.c_int() for _ in range(3)]
    net.lib.bp_kernel_shape(*[ct.byref(x) for x in shape])
    info["time_tile"], info["waves_per_simd"] = shape[1].value, shape[2].value
    # rows zero-padded for this network (as the drop-in entry points stage
    # them): every LDS tile takes the staged path
    pad = net.padding()
    F = bpd.pack(F_plain, bpd.BP_F32, *pad)
    bmax = torch.zeros(T, dtype=torch.float32, device="cuda")
    barg = torch.empty(T, dtype=torch.int32, device="cuda")
    keys = torch.empty(T, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        net.beamform_max(F, oob, bmax, barg, keys)
    torch.cuda.synchronize()

    log(f"timing {args.steps} steps")
    # the stack kernels are bracketed with HIP events on the launch stream
    # inside the same timed loop, so kernel_ms <= ms_per_step
    with bpd.KernelTimer() as kt:
        t0 = time.perf_counter()
        for _ in range(args.steps):
            net.beamform_max(F, oob, bmax, barg, keys)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    value = K * T / (dt / args.steps)
    kernel_ms = kt.total_ms / max(1, kt.launches)
    log(f"step {ms_per_step:.2f} ms, stack kernel {kernel_ms:.2f} ms")

    # reduced-precision feature storage on the same workload
    precision = {}
    a32 = barg.cpu().numpy().copy()
    # top-two gap of the FP32 beams on four seeded 1-minute slices (full
    # beams of every source): where the two leading sources differ by more
    # than 4 units of a storage format's precision, its rounding cannot
    # reorder them
    W_gap, halo = 1500, int(w.time_delays.max()) + 1
    gap_starts = np.sort(np.random.default_rng(7).choice(
        np.arange(0, T - W_gap - halo, W_gap), size=4, replace=False))
    gaps = []
    for t0 in gap_starts:
        sub = bpd.FeatureView(F.data, F.fmt, S, P, W_gap + halo, F.row_stride,
                              F.pad_before + int(t0), F.weight_scale)
        b32 = net.beamform(sub, oob)[:, :W_gap]
        top2 = torch.topk(b32, 2, dim=0).values
        gaps.append(((top2[0] - top2[1]) / top2[0].abs()).cpu().numpy())
        del b32, top2
    gaps = np.concatenate(gaps)
    gap_idx = np.concatenate([np.arange(t0, t0 + W_gap) for t0 in gap_starts])
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    unit = {"bf16": 2.0 ** -8, "f16": 2.0 ** -11}
    for name, fmt in (("bf16", bpd.BP_BF16), ("f16", bpd.BP_F16)):
        Q = bpd.pack(F_plain, fmt, *pad)
        for _ in range(args.warmup):
            net.beamform_max(Q, oob, bmax, barg, keys)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            net.beamform_max(Q, oob, bmax, barg, keys)
        torch.cuda.synchronize()
        dtq = (time.perf_counter() - t1) / args.steps
        aq = barg.cpu().numpy()
        decides = gaps > 4 * unit[name] + 1e-6
        precision[name] = {"value": K * T / dtq, "ms_per_step": dtq * 1e3,
                           "speedup_vs_f32": (dt / args.steps) / dtq,
                           "argmax_agreement_vs_f32": float(np.mean(aq == a32)),
                           "argmax_agreement_vs_f32_where_gap_decides": float(
                               np.mean(aq[gap_idx][decides] == a32[gap_idx][decides])),
                           "decidable_fraction": float(np.mean(decides)),
                           "gap_sample": f"{len(gaps)} samples in 4 seeded 1-minute slices"}
        log(f"{name}: step {dtq * 1e3:.2f} ms")
        del Q
    # restore the FP32 result for the checks below
    net.beamform_max(F, oob, bmax, barg, keys)
    torch.cuda.synchronize()

    # the host-pointer drop-in entry point on the same data: network plan,
    # host->device copies of F, tau and beta, the kernels and the copy back
    from beampower_amd import load_library
    lib = load_library("gpu")
    F_host = np.ascontiguousarray(F_plain.permute(0, 2, 1).cpu().numpy())  # (S, T, P)
    tau_h = np.ascontiguousarray(w.time_delays, np.int32)
    beta_h = np.ascontiguousarray(w.weights_sources, np.float32)
    m_h = np.zeros(T, np.float32)
    a_h = np.zeros(T, np.int32)
    P_F, P_I = ct.POINTER(ct.c_float), ct.POINTER(ct.c_int)
    lib.bp_dropin_cache_clear()

    def host_call():
        t1 = time.perf_counter()
        rc = lib.beamform_max(F_host.ctypes.data_as(P_F), tau_h.ctypes.data_as(P_I),
                              beta_h.ctypes.data_as(P_F), T, K, S, P, oob,
                              m_h.ctypes.data_as(P_F), a_h.ctypes.data_as(P_I))
        assert rc == 0
        return (time.perf_counter() - t1) * 1e3

    # with the plan rebuilt on every call (reuse off), then the same arrays
    # passed again (the first call builds and keeps the network, the second
    # re-plans it once, the others reuse it)
    os.environ["BP_DROPIN_CACHE"] = "0"
    cold_ms = min(host_call() for _ in range(3))
    del os.environ["BP_DROPIN_CACHE"]
    first_ms = [host_call() for _ in range(2)]
    warm_ms = min(host_call() for _ in range(3))
    lib.bp_dropin_cache_clear()
    dropin_host = {"ms": cold_ms, "value": K * T / (cold_ms * 1e-3),
                   "unit": "source·samples/s",
                   "identical_to_resident": bool(np.array_equal(a_h, barg.cpu().numpy())),
                   "note": "beamform_max with host pointers: uploads and plan on every call",
                   "repeated_inputs": {
                       "ms": warm_ms, "value": K * T / (warm_ms * 1e-3),
                       "first_two_calls_ms": first_ms,
                       "note": "same time_delays and weights_sources on every call (a detection "
                               "loop over windows): the kept network is reused, features are "
                               "uploaded per call"}}
    log(f"host-pointer drop-in: {cold_ms:.1f} ms planned per call, {warm_ms:.1f} ms reused "
        f"(first calls {first_ms[0]:.1f}, {first_ms[1]:.1f} ms)")
    del F_host

    # the Python call itself, beamform(waveform_features, time_delays,
    # weights_phases, weights_sources, device="gpu"): raw features in, one
    # fused library call (prestack on the device)
    import beampower_amd as bp

    def py_call():
        t1 = time.perf_counter()
        m, a = bp.beamform(w.waveform_features, w.time_delays, w.weights_phases,
                           w.weights_sources, device="gpu", reduce="max")
        return (time.perf_counter() - t1) * 1e3, a

    os.environ["BP_DROPIN_CACHE"] = "0"
    py_cold = [py_call() for _ in range(3)]
    del os.environ["BP_DROPIN_CACHE"]
    py_first = [py_call()[0] for _ in range(2)]
    py_warm = [py_call() for _ in range(3)]
    lib.bp_dropin_cache_clear()
    py_cold_ms = min(t for t, _ in py_cold)
    py_warm_ms = min(t for t, _ in py_warm)
    a_ref = barg.cpu().numpy()
    dropin_host["python_call"] = {
        "ms": py_cold_ms, "value": K * T / (py_cold_ms * 1e-3),
        "identical_to_resident": all(bool(np.array_equal(a, a_ref)) for _, a in py_cold + py_warm),
        "repeated_inputs_ms": py_warm_ms, "first_two_calls_ms": py_first,
        "note": "beampower_amd.beamform(..., device='gpu', reduce='max') on the (50, 3, 90000) "
                "features: prestack on the device in the same library call, inputs as numpy "
                "arrays, outputs returned as numpy arrays"}
    log(f"python drop-in call: {py_cold_ms:.1f} ms planned per call, {py_warm_ms:.1f} ms reused")

    # known answer on the benchmark output itself
    a = barg.cpu().numpy()
    recovered = sum(int(np.array_equal(w.time_delays[a[t_e]], w.time_delays[k_e]))
                    for k_e, t_e in w.events)

    used_terms = float((w.weights_sources != 0).sum()) * P  # (k, s, p) terms per sample
    lds_bytes = used_terms * T * 4.0
    fmas = used_terms * T
    achieved = lds_bytes / (kernel_ms * 1e-3)
    roofline = {
        "bound": "lds",
        "achieved": achieved / 1e9,
        "peak": LDS_PEAK / 1e9,
        "unit": "GB/s",
        "frac": achieved / LDS_PEAK,
        "traffic": profile_traffic(),
        "kernel": "stack_lds_kernel",
        "kernel_ms": kernel_ms,
        "algorithmic_bytes_per_launch": lds_bytes,
        "valu_frac": (2.0 * fmas / (kernel_ms * 1e-3)) / VALU_FP32_PEAK_FLOPS,
        "valu_frac_unpacked": (2.0 * fmas / (kernel_ms * 1e-3)) / VALU_FP32_UNPACKED_FLOPS,
        # against the v_fma_f32 rate the pipe probe measures under load
        # (34.3 TFMA/s, profiles/peaks_probe.md)
        "valu_frac_of_measured_peak": (fmas / (kernel_ms * 1e-3)) / VALU_FMA_MEASURED,
        "hbm_compulsory_frac": ((S * P * T * 4 + K * S * P * 4 + K * S * 4 + 8 * T)
                                / (kernel_ms * 1e-3)) / HBM_PEAK,
    }

    del net, F, bmax, barg, keys
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    none_subset = bench_none_subset(args, log, w, F_plain)
    del F_plain
    torch.cuda.empty_cache()
    batched = bench_batched(args, log) if args.batch_windows > 0 else None
    if batched is not None and args.mix_windows > 0:
        batched["serving_mix"] = bench_batched_mix(args, log)
    serving = bench_serving(args, log) if args.serve_rates else None
    differential = None if args.no_differential else bench_differential(args, log)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(w, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "source·samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded homogeneous-velocity moveouts, envelope-like features)",
        "config": {
            "workload": w.name, "stations": S, "channels": C, "phases": P,
            "sources": K, "samples": T, "reduce": "max", "out_of_bounds": args.oob,
            "source_tile": info["source_tile"], "max_tile_spread": info["max_tile_spread"],
            "tiles_per_class": info["tiles_per_class"], "window_per_class": info["window_per_class"],
            "row_padding": list(pad),
            "source_order": info["source_order"], "time_tile": info["time_tile"],
            "waves_per_simd": info["waves_per_simd"],
        },
        "gather_fma_per_s": value * S * P,
        "events_recovered": f"{recovered}/{len(w.events)}",
        "roofline": roofline,
        "dropin_host_pointers": dropin_host,
        "reduce_none_subset": none_subset,
        "precision_forms": precision,
        "batched_form": batched,
        "serving_form": serving,
        "differential_form": differential,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
