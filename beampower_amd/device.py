"""Device-resident front-end of the gfx950 library (no reference counterpart).

The reference copies its inputs to the GPU on every call
(beamform.cu:314-340).  Here a `Network` uploads moveouts and source weights
once (`bp_network_create`) and then beamforms device-resident features.
PyTorch provides device memory and streams only; every computation runs in
the library's HIP kernels.
"""
import ctypes as ct

import numpy as np

from .core import BatchConfig, FeatureViewC, check, load_library

BP_F32, BP_BF16, BP_F16 = 0, 1, 2


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("beampower_amd.device needs a ROCm GPU")
    return torch


def _stream_ptr(stream, *tensors):
    """HIP stream handle of a library call on `stream` (None = the current
    stream).  When the call runs on another stream than the current one:
    that stream first waits for the work already on the current stream
    (the wrapper's contiguous copies and zero fills, and the caller's own
    preparation of the inputs), and the tensors the call reads or writes are
    marked as used on it (record_stream), so that torch's caching allocator
    does not hand their memory to other work while the call's kernels may
    still run, even if the caller drops them on return."""
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    elif stream != torch.cuda.current_stream():
        stream.wait_stream(torch.cuda.current_stream())
        for t in tensors:
            if t is not None:
                t.record_stream(stream)
    return ct.c_void_p(stream.cuda_stream)


def _zeros(shape, dtype, device, stream):
    """A zero-filled default output, filled on the stream the library call
    runs on (a fill on the current stream would not be ordered before the
    call's kernels on another stream)."""
    torch = _torch()
    if stream is None:
        return torch.zeros(shape, dtype=dtype, device=device)
    with torch.cuda.stream(stream):
        return torch.zeros(shape, dtype=dtype, device=device)


# Source order of a network's plan (bp_network_create_ordered).
ORDERS = {"natural": 0, "similarity": 1, "auto": 2}


class Network:
    """Moveouts (K,S,P) int32 and source weights (K,S) float32 resident on the
    current device, with per-source and per-source-tile delay bounds.

    order: "auto" (the default: sources grouped into tiles by delay
    similarity when that plan is cheaper), "similarity" (always) or
    "natural" (tiles of consecutive source indices).  Results do not depend
    on it: beams are indexed, and max ties broken, by the caller's index."""

    def __init__(self, time_delays, weights_sources, order="auto"):
        if order not in ORDERS:
            raise ValueError(f"order must be one of {sorted(ORDERS)}, not {order!r}")
        self.lib = load_library("gpu")
        tau = np.ascontiguousarray(time_delays, dtype=np.int32)
        beta = np.ascontiguousarray(weights_sources, dtype=np.float32)
        if tau.ndim != 3 or beta.shape != tau.shape[:2]:
            raise ValueError(f"shapes {tau.shape} / {beta.shape} do not match (K,S,P) / (K,S)")
        self.n_sources, self.n_stations, self.n_phases = tau.shape
        handle = ct.c_void_p()
        check(self.lib.bp_network_create_ordered(tau.ctypes.data, beta.ctypes.data,
                                                 self.n_sources, self.n_stations, self.n_phases,
                                                 ORDERS[order], ct.byref(handle)),
              "bp_network_create_ordered", self.lib)
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            self.lib.bp_network_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def info(self):
        tile, spread, n = ct.c_int(), ct.c_int(), ct.c_int()
        check(self.lib.bp_network_info(self._h, ct.byref(tile), ct.byref(spread),
                                       ct.byref(n)), "bp_network_info", self.lib)
        counts, windows = (ct.c_int * 3)(), (ct.c_int * 3)()
        check(self.lib.bp_network_classes(self._h, counts, windows), "bp_network_classes",
              self.lib)
        return dict(source_tile=tile.value, max_tile_spread=spread.value, n_tiles=n.value,
                    tiles_per_class=list(counts), window_per_class=list(windows))

    def source_order(self):
        """(permuted, order): whether the plan reorders the sources, and the
        caller's source index at each plan position (int32, n_sources)."""
        permuted = ct.c_int()
        order = np.empty(self.n_sources, dtype=np.int32)
        check(self.lib.bp_network_source_order(self._h, ct.byref(permuted), order.ctypes.data),
              "bp_network_source_order", self.lib)
        return bool(permuted.value), order

    def diff_tables(self):
        """Device bytes of the differential tables built so far:
        dict(delay_bytes, weight_bytes, weight_form), weight_form "none"
        (not built), "float32", or "used_bits" (every nonzero weight has the
        same value: one bit per source and station)."""
        d, w, f = ct.c_size_t(), ct.c_size_t(), ct.c_int()
        check(self.lib.bp_network_diff_table_bytes(self._h, ct.byref(d), ct.byref(w),
                                                   ct.byref(f)),
              "bp_network_diff_table_bytes", self.lib)
        return dict(delay_bytes=d.value, weight_bytes=w.value,
                    weight_form=("none", "float32", "used_bits")[f.value])

    def padding(self):
        """(pad_before, pad_after): zero padding of feature rows under which
        no LDS tile of this network takes the guarded edge path."""
        pb, pa = ct.c_size_t(), ct.c_size_t()
        check(self.lib.bp_network_padding(self._h, ct.byref(pb), ct.byref(pa)),
              "bp_network_padding", self.lib)
        return pb.value, pa.value

    def beamform_max(self, features, out_of_bounds=0, beam_max=None, beam_argmax=None,
                     keys=None, stream=None):
        """features: planar (S,P,T) float32 CUDA tensor, a QuantizedFeatures
        (16-bit storage) or a FeatureView (padded rows, any format).  Returns
        device tensors (beam_max (T,), beam_argmax (T,)); asynchronous."""
        torch = _torch()
        if isinstance(features, FeatureView):
            if features.n_stations_phases != (self.n_stations, self.n_phases):
                raise ValueError("features do not match the network")
            T, dev = features.n_samples, features.data.device
            if beam_max is None:
                beam_max = _zeros(T, torch.float32, dev, stream)
            if beam_argmax is None:
                beam_argmax = torch.empty(T, dtype=torch.int32, device=dev)
            if keys is None:
                keys = torch.empty(T, dtype=torch.int64, device=dev)
            view = features.c_view()
            check(self.lib.bp_network_beamform_max_view(
                self._h, ct.byref(view), int(out_of_bounds), beam_max.data_ptr(),
                beam_argmax.data_ptr(), keys.data_ptr(),
                _stream_ptr(stream, features.data, beam_max, beam_argmax, keys)),
                "bp_network_beamform_max_view", self.lib)
            return beam_max, beam_argmax
        fmt, scale = BP_F32, None
        if isinstance(features, QuantizedFeatures):
            fmt, T, dev = features.fmt, features.n_samples, features.data.device
            S, P = features.data.shape[:2]
            scale = features.weight_scale.ctypes.data
            data = features.data
            ptr = data.data_ptr()
        else:
            S, P, T = features.shape
            if features.dtype != torch.float32:
                raise ValueError("features must be float32 (n_stations, n_phases, n_samples)")
            data = features.contiguous()
            dev, ptr = data.device, data.data_ptr()
        if (S, P) != (self.n_stations, self.n_phases):
            raise ValueError(f"features are for {(S, P)} stations/phases, network for "
                             f"{(self.n_stations, self.n_phases)}")
        if beam_max is None:
            beam_max = _zeros(T, torch.float32, dev, stream)
        if beam_argmax is None:
            beam_argmax = torch.empty(T, dtype=torch.int32, device=dev)
        if keys is None:
            keys = torch.empty(T, dtype=torch.int64, device=dev)
        check(self.lib.bp_network_beamform_max(
            self._h, ptr, fmt, scale, T, int(out_of_bounds),
            beam_max.data_ptr(), beam_argmax.data_ptr(), keys.data_ptr(),
            _stream_ptr(stream, data, beam_max, beam_argmax, keys)), "bp_network_beamform_max",
            self.lib)
        return beam_max, beam_argmax

    def beamform(self, features, out_of_bounds=0, source_begin=0, source_end=None,
                 beam=None, stream=None):
        """Full beams of sources [source_begin, source_end) as a (n, T)
        float32 CUDA tensor (cells invalid in strict mode stay 0).  features:
        planar (S,P,T) float32 tensor, QuantizedFeatures or FeatureView."""
        torch = _torch()
        if isinstance(features, (FeatureView, QuantizedFeatures)):
            v = FeatureView.of(features)
            source_end = self.n_sources if source_end is None else int(source_end)
            n = source_end - int(source_begin)
            if beam is None:
                beam = _zeros((n, v.n_samples), torch.float32, v.data.device, stream)
            cv = v.c_view()
            check(self.lib.bp_network_beamform_view(
                self._h, ct.byref(cv), int(out_of_bounds), int(source_begin), source_end,
                beam.data_ptr(), _stream_ptr(stream, v.data, beam)), "bp_network_beamform_view",
                self.lib)
            return beam
        S, P, T = features.shape
        source_end = self.n_sources if source_end is None else int(source_end)
        n = source_end - int(source_begin)
        if beam is None:
            beam = _zeros((n, T), torch.float32, features.device, stream)
        data = features.contiguous()
        check(self.lib.bp_network_beamform(
            self._h, data.data_ptr(), T, int(out_of_bounds),
            int(source_begin), source_end, beam.data_ptr(), _stream_ptr(stream, data, beam)),
            "bp_network_beamform", self.lib)
        return beam

    def beamform_differential(self, features, beam=None, stream=None):
        """mode='differential' (beamform.c:338-394): one beam per source at
        the centre sample (T - 1) // 2, as a (n_sources,) float32 CUDA
        tensor; sources with a used delay outside [0, T) keep `beam`'s value
        (zeros by default).  features: planar (S,P,T) float32 tensor,
        QuantizedFeatures or FeatureView."""
        torch = _torch()
        v = FeatureView.of(features)
        if beam is None:
            beam = _zeros(self.n_sources, torch.float32, v.data.device, stream)
        cv = v.c_view()
        check(self.lib.bp_network_beamform_differential(self._h, ct.byref(cv), beam.data_ptr(),
                                                        _stream_ptr(stream, v.data, beam)),
              "bp_network_beamform_differential", self.lib)
        return beam


class QuantizedFeatures:
    """16-bit feature storage: parity planes (S, P, 2, ldq) as uint16 on the
    device (plane 0 x[t]*scale, plane 1 x[t+1]*scale) and the host weight
    multipliers 1/scale per station."""

    def __init__(self, data, weight_scale, fmt, n_samples):
        self.data, self.weight_scale, self.fmt, self.n_samples = data, weight_scale, fmt, n_samples


class FeatureView:
    """Device features with zero-padded rows (bp_feature_view): `data` holds
    S*P rows (F32) or S*P*2 parity-plane rows (16-bit) of `row_stride`
    elements, sample 0 at element `pad_before`."""

    def __init__(self, data, fmt, n_stations, n_phases, n_samples, row_stride, pad_before,
                 weight_scale=None):
        self.data, self.fmt = data, int(fmt)
        self.n_stations_phases = (int(n_stations), int(n_phases))
        self.n_samples, self.row_stride, self.pad_before = int(n_samples), int(row_stride), int(pad_before)
        self.weight_scale = weight_scale

    @classmethod
    def of(cls, features):
        """A view of a planar float32 tensor (S,P,T) or a QuantizedFeatures."""
        if isinstance(features, FeatureView):
            return features
        if isinstance(features, QuantizedFeatures):
            S, P, _, ldq = features.data.shape
            return cls(features.data, features.fmt, S, P, features.n_samples, ldq, 0,
                       features.weight_scale)
        S, P, T = features.shape
        return cls(features.contiguous(), BP_F32, S, P, T, T, 0)

    def c_view(self):
        scale = None if self.weight_scale is None else self.weight_scale.ctypes.data
        return FeatureViewC(self.data.data_ptr(), self.fmt, self.n_samples, self.row_stride,
                            self.pad_before, scale)


def pack(F_spt, fmt=BP_F32, pad_before=0, pad_after=0, stream=None):
    """Planar (S,P,T) float32 CUDA tensor -> FeatureView in `fmt` with
    `pad_before` / `pad_after` zero samples around every row (16-bit rows are
    padded to even lengths)."""
    torch = _torch()
    lib = load_library("gpu")
    S, P, T = F_spt.shape
    pad_before, pad_after = int(pad_before), int(pad_after)
    if fmt != BP_F32:
        pad_before += pad_before & 1
    row_stride = pad_before + T + pad_after
    if fmt != BP_F32:
        row_stride += row_stride & 1
    rows = S * P * (1 if fmt == BP_F32 else 2)
    dtype = torch.float32 if fmt == BP_F32 else torch.int16
    data = torch.empty(rows * row_stride, dtype=dtype, device=F_spt.device)
    scale = None if fmt == BP_F32 else np.empty(S, np.float32)
    src = F_spt.contiguous()
    check(lib.bp_pack_features(src.data_ptr(), S, P, T, int(fmt), pad_before,
                               row_stride, None if scale is None else scale.ctypes.data,
                               data.data_ptr(), _stream_ptr(stream, src, data)),
          "bp_pack_features", lib)
    return FeatureView(data, fmt, S, P, T, row_stride, pad_before, scale)


def quantize(F_spt, fmt=BP_BF16, stream=None):
    """Planar (S,P,T) float32 CUDA tensor -> QuantizedFeatures in BP_F16 or
    BP_BF16 (per-station power-of-two scales)."""
    torch = _torch()
    lib = load_library("gpu")
    S, P, T = F_spt.shape
    ldq = lib.bp_parity_stride(T)
    out = torch.empty((S, P, 2, ldq), dtype=torch.int16, device=F_spt.device)
    wscale = np.empty(S, np.float32)
    src = F_spt.contiguous()
    check(lib.bp_quantize_planar(src.data_ptr(), S, P, T, int(fmt),
                                 wscale.ctypes.data, out.data_ptr(), _stream_ptr(stream, src, out)),
          "bp_quantize_planar", lib)
    return QuantizedFeatures(out, wscale, int(fmt), T)


def interleaved_to_planar(F_stp, stream=None):
    """(S,T,P) float32 CUDA tensor -> (S,P,T)."""
    torch = _torch()
    lib = load_library("gpu")
    S, T, P = F_stp.shape
    out = torch.empty((S, P, T), dtype=torch.float32, device=F_stp.device)
    src = F_stp.contiguous()
    check(lib.bp_interleaved_to_planar(src.data_ptr(), S, T, P,
                                       out.data_ptr(), _stream_ptr(stream, src, out)),
          "bp_interleaved_to_planar", lib)
    return out


def prestack_planar(u, weights_phases, stream=None):
    """(S,C,T) features and (S,C,P) weights (CUDA float32) -> (S,P,T)."""
    torch = _torch()
    lib = load_library("gpu")
    S, C, T = u.shape
    P = weights_phases.shape[2]
    out = torch.empty((S, P, T), dtype=torch.float32, device=u.device)
    uc, wc = u.contiguous(), weights_phases.contiguous()
    check(lib.bp_prestack_planar(uc.data_ptr(), wc.data_ptr(),
                                 S, C, T, P, out.data_ptr(), _stream_ptr(stream, uc, wc, out)),
          "bp_prestack_planar", lib)
    return out


class Batch:
    """Batched reduce='max' over windows of resident networks
    (bp_batch_*): one persistent launch per LDS window class for all windows.

    windows: sequence of (network, features, out_of_bounds), features being
    a planar float32 CUDA tensor or QuantizedFeatures (one format for the
    whole batch).  Outputs are allocated here: `beam_max[i]` (zeros) and
    `beam_argmax[i]` CUDA tensors of the window's length."""

    def __init__(self, windows):
        torch = _torch()
        self.lib = load_library("gpu")
        n = len(windows)
        if n == 0:
            raise ValueError("a batch needs at least one window")
        self.windows = list(windows)  # keeps networks and features alive
        views = []
        self.beam_max, self.beam_argmax, self._keys = [], [], []
        for net, feats, _ in self.windows:
            if not isinstance(feats, (FeatureView, QuantizedFeatures)) and (
                    feats.dtype != torch.float32 or not feats.is_contiguous()):
                raise ValueError("features must be contiguous float32 (S, P, T)")
            v = FeatureView.of(feats)
            if v.n_stations_phases != (net.n_stations, net.n_phases):
                raise ValueError("features do not match their network")
            views.append(v)
            T, dev = v.n_samples, v.data.device
            self.beam_max.append(torch.zeros(T, dtype=torch.float32, device=dev))
            self.beam_argmax.append(torch.empty(T, dtype=torch.int32, device=dev))
            self._keys.append(torch.empty(max(T, 1), dtype=torch.int64, device=dev))
        if len({v.fmt for v in views}) != 1:
            raise ValueError("one feature format per batch")
        # the outputs' zero fill is on the current stream; launches may use any
        torch.cuda.current_stream().synchronize()
        self._views = views
        vp = ct.c_void_p
        arr = lambda xs: (vp * n)(*xs)
        self._args = dict(
            nets=arr([net._h.value for net, _, _ in self.windows]),
            views=(FeatureViewC * n)(*[v.c_view() for v in views]),
            oob=(ct.c_int * n)(*[int(o) for _, _, o in self.windows]),
            bmax=arr([t.data_ptr() for t in self.beam_max]),
            barg=arr([t.data_ptr() for t in self.beam_argmax]),
            keys=arr([t.data_ptr() for t in self._keys]))
        h = vp()
        a = self._args
        check(self.lib.bp_batch_create(n, a["nets"], a["views"], a["oob"], a["bmax"], a["barg"],
                                       a["keys"], ct.byref(h)), "bp_batch_create", self.lib)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.bp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _tensors(self):
        return ([v.data for v in self._views] + self.beam_max + self.beam_argmax +
                self._keys)

    def launch(self, stream=None):
        check(self.lib.bp_batch_launch(self._h, _stream_ptr(stream, *self._tensors())),
              "bp_batch_launch", self.lib)

    def graph_launch(self, stream=None):
        check(self.lib.bp_batch_graph_launch(self._h, _stream_ptr(stream, *self._tensors())),
              "bp_batch_graph_launch", self.lib)

    def config(self):
        c = BatchConfig()
        check(self.lib.bp_batch_get_config(self._h, ct.byref(c)), "bp_batch_get_config",
              self.lib)
        return dict(time_group=c.time_group, stage_buffers=c.stage_buffers,
                    blocks_per_cu=c.blocks_per_cu, ms=c.ms)

    def set_config(self, time_group=8, stage_buffers=2, blocks_per_cu=3):
        """Launch configuration (bp_batch_config; the defaults are the
        library's): time tiles per block-order group, staging buffers, and
        1-4 = all windows in one persistent launch of that many work-groups
        per CU taking work items from a queue; 0 = all fused, one work-group
        per item; -1 = windows that fill the GPU alone in launches of their
        own, the others fused."""
        c = BatchConfig(time_group, stage_buffers, blocks_per_cu, 0.0)
        check(self.lib.bp_batch_set_config(self._h, ct.byref(c)), "bp_batch_set_config",
              self.lib)

    def tune(self, reps=2, stream=None):
        """Time every configuration on this batch, keep the fastest.  Returns
        (best, tried) as dicts."""
        tried = (BatchConfig * 64)()
        n, best = ct.c_int(), BatchConfig()
        check(self.lib.bp_batch_tune(self._h, int(reps), tried, 64, ct.byref(n), ct.byref(best),
                                     _stream_ptr(stream, *self._tensors())), "bp_batch_tune",
              self.lib)
        conv = lambda c: dict(time_group=c.time_group, stage_buffers=c.stage_buffers,
                              blocks_per_cu=c.blocks_per_cu, ms=c.ms)
        return conv(best), [conv(tried[i]) for i in range(n.value)]


class KernelTimer:
    """Accumulates the HIP-event time of the library's stack kernel."""

    def __init__(self):
        self.lib = load_library("gpu")

    def __enter__(self):
        self.lib.bp_profiling_read(None, None)
        self.lib.bp_profiling(1)
        return self

    def __exit__(self, *exc):
        ms, n = ct.c_double(), ct.c_longlong()
        self.lib.bp_profiling(0)
        self.lib.bp_profiling_read(ct.byref(ms), ct.byref(n))
        self.total_ms, self.launches = ms.value, n.value
        return False
