"""Serving form: resident networks, HBM-budget admission and continuous
batching of window requests (bp_server_* of include/beamform_gpu.h; no
counterpart in the reference, which serves one window per beamform() call).

The request loop is native (two threads inside the library); this module
holds the host buffers of each request and drives Poisson arrivals for
measurements.
"""
import ctypes as ct
import threading
import time

import numpy as np

from .core import RequestInfo, ServerConfig, ServerCounters, load_library
from .device import BP_BF16, BP_F32


class Ticket:
    """One submitted window: its id, host features and host outputs.  The
    server's native threads read the features and write the outputs, so the
    Server holds every ticket until its request has finished (wait() got
    its final status) or the server is closed."""

    def __init__(self, request_id, network, features, beam_max, beam_argmax):
        self.id, self.network = request_id, network
        self.features, self.beam_max, self.beam_argmax = features, beam_max, beam_argmax
        self.info = None


class Server:
    """Resident networks and a continuous-batching request loop on the
    current device.  fmt: BP_BF16 (default) or BP_F32 window storage."""

    def __init__(self, hbm_budget=8 << 30, max_batch=64, fmt=BP_BF16, in_flight=2):
        self.lib = load_library("gpu")
        cfg = ServerConfig(int(hbm_budget), int(max_batch), int(fmt), int(in_flight))
        h = ct.c_void_p()
        self._check(self.lib.bp_server_create(ct.byref(cfg), ct.byref(h)), "bp_server_create")
        self._h = h
        self.fmt = int(fmt)
        self.networks = []
        self._pending = {}  # request id -> Ticket, until its final status

    def _check(self, status, what):
        if status != 0:
            msg = self.lib.bp_server_last_error().decode(errors="replace")
            raise RuntimeError(f"{what} failed with status {status}: {msg}")

    def close(self):
        """Stop the server: bp_server_destroy joins the native threads (the
        batches in flight complete, requests not yet loaded are dropped), and
        only then are the pending requests' buffers released."""
        if getattr(self, "_h", None):
            self.lib.bp_server_destroy(self._h)
            self._h = None
        self._pending = {}

    def __del__(self):
        self.close()

    def add_network(self, time_delays, weights_sources, weights_phases):
        """Make a network resident: moveouts (K,S,P), source weights (K,S),
        phase weights (S,C,P).  Returns its id."""
        tau = np.ascontiguousarray(time_delays, dtype=np.int32)
        beta = np.ascontiguousarray(weights_sources, dtype=np.float32)
        alpha = np.ascontiguousarray(weights_phases, dtype=np.float32)
        K, S, P = tau.shape
        C = alpha.shape[1]
        if beta.shape != (K, S) or alpha.shape != (S, C, P):
            raise ValueError("shapes do not match (K,S,P) / (K,S) / (S,C,P)")
        nid = ct.c_int()
        self._check(self.lib.bp_server_add_network(self._h, tau.ctypes.data, beta.ctypes.data,
                                                   alpha.ctypes.data, K, S, C, P,
                                                   ct.byref(nid)), "bp_server_add_network")
        self.networks.append((S, C, P, K))
        return nid.value

    def submit(self, network, waveform_features, out_of_bounds=0):
        """Queue one window (S,C,T) float32 of `network`; returns a Ticket."""
        u = np.ascontiguousarray(waveform_features, dtype=np.float32)
        S, C, P, _ = self.networks[network]
        if u.ndim != 3 or u.shape[:2] != (S, C):
            raise ValueError(f"features {u.shape} do not match network ({S}, {C}, T)")
        T = u.shape[2]
        bmax = np.zeros(T, np.float32)
        barg = np.zeros(T, np.int32)
        rid = ct.c_longlong()
        self._check(self.lib.bp_server_submit(self._h, int(network), u.ctypes.data, T,
                                              int(out_of_bounds), bmax.ctypes.data,
                                              barg.ctypes.data, ct.byref(rid)),
                    "bp_server_submit")
        ticket = Ticket(rid.value, network, u, bmax, barg)
        self._pending[ticket.id] = ticket
        return ticket

    def wait(self, ticket, timeout=None):
        """Block until the request is done; returns (beam_max, beam_argmax),
        with ticket.info holding its timings."""
        info = RequestInfo()
        st = self.lib.bp_server_wait(self._h, ticket.id, -1.0 if timeout is None else timeout,
                                     ct.byref(info))
        if st == 1:
            raise TimeoutError(f"request {ticket.id} not done after {timeout} s")
        self._pending.pop(ticket.id, None)  # final status: the threads are done with it
        self._check(st, "request")
        ticket.info = {f: getattr(info, f) for f, _ in RequestInfo._fields_}
        return ticket.beam_max, ticket.beam_argmax

    def stats(self):
        c = ServerCounters()
        self._check(self.lib.bp_server_stats(self._h, ct.byref(c)), "bp_server_stats")
        return {f: getattr(c, f) for f, _ in ServerCounters._fields_}


def poisson_run(server, requests, rate, seed=0):
    """Submit `requests` (a list of (network, features)) at Poisson arrival
    times of `rate` per second from a submitting thread, wait for all, and
    return (tickets, summary) with latency percentiles (submit to results in
    host memory) and throughput."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(requests))
    tickets = [None] * len(requests)
    start = time.perf_counter()

    def feed():
        t_next = start
        for i, (net, u) in enumerate(requests):
            t_next += gaps[i]
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = server.submit(net, u)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    for t in tickets:
        server.wait(t)
    lat = np.array([t.info["done_s"] - t.info["submit_s"] for t in tickets])
    span = max(t.info["done_s"] for t in tickets) - min(t.info["submit_s"] for t in tickets)
    samples = sum(t.features.shape[2] for t in tickets)
    work = sum(server.networks[t.network][3] * t.features.shape[2] for t in tickets)
    # gather-FMAs (source * station * phase * sample): work comparable across networks
    # of different station counts
    fmas = sum(np.prod(np.array(server.networks[t.network])[[0, 2, 3]], dtype=np.float64)
               * t.features.shape[2] for t in tickets)
    batch = np.array([t.info["batch_windows"] for t in tickets])
    # where a request's time goes (medians, ms): waiting for HBM budget,
    # upload and prestack, waiting for a launch, on the GPU, results copied
    stages = (("admit", "submit_s", "admit_s"), ("stage", "admit_s", "ready_s"),
              ("queue", "ready_s", "launch_s"), ("gpu", "launch_s", "gpu_done_s"),
              ("download", "gpu_done_s", "done_s"))
    parts = {name: float(np.median([t.info[b] - t.info[a] for t in tickets]) * 1e3)
             for name, a, b in stages}
    summary = {
        "rate_offered": rate, "requests": len(requests),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "throughput_req_s": len(requests) / span,
        "source_samples_per_s": work / span,
        "gather_fma_per_s": fmas / span,
        "samples": samples,
        "mean_batch_windows": float(batch.mean()),
        "median_stage_ms": parts,
    }
    return tickets, summary
