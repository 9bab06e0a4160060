"""ctypes front-end of the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check results.  The product package
(beampower_amd) never imports this module.

Two kinds of CPU engines are exposed:

* the oracle proper (oracle/beamform_oracle.c, a restatement of
  the reference's beampower/src/beamform.c:10-336), in the two contraction
  flavours of the reference's builds: ``"fma"`` (CMakeLists.txt:28-33, the
  pip wheels) and ``"nofma"`` (Makefile:69, -std=c99);
* the reference engine itself, compiled from beampower/src/beamform.c by
  ``make -C oracle ref`` into oracle/_ref/ (flavours ``"cmake"`` and
  ``"make"``), called through the argument tables of beampower/core.py:14-49.
"""
import ctypes as ct
import os

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

_P_F = ct.POINTER(ct.c_float)
_P_D = ct.POINTER(ct.c_double)
_P_I = ct.POINTER(ct.c_int)
_P_U8 = ct.POINTER(ct.c_ubyte)
_SZ = ct.c_size_t

_LIBS = {}


def build(quiet=True):
    """Compile the oracle (and the reference into _ref/ when its sources exist)."""
    import subprocess

    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-C", ORACLE_DIR, "all"], check=True, stdout=out)
    if os.path.exists("/root/reference/beampower/src/beamform.c"):
        subprocess.run(["make", "-C", ORACLE_DIR, "ref"], check=True, stdout=out)


def _load(path):
    if path not in _LIBS:
        if not os.path.exists(path):
            raise OSError(f"oracle library {path} is not built (run `make -C oracle`)")
        _LIBS[path] = ct.CDLL(path)
    return _LIBS[path]


def _oracle(flavour):
    if flavour not in ("fma", "nofma"):
        raise ValueError(f"flavour must be 'fma' or 'nofma', not {flavour!r}")
    lib = _load(os.path.join(ORACLE_DIR, f"liboracle_{flavour}.so"))
    if not getattr(lib, "_typed", False):
        lib.oracle_delay_bounds.argtypes = [_P_I, _P_F, _SZ, _SZ, _SZ, ct.c_int, _P_I]
        lib.oracle_prestack.argtypes = [_P_F, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int, _P_F]
        lib.oracle_beamform.argtypes = [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ,
                                        ct.c_int, ct.c_int, _P_F, _P_I]
        lib.oracle_beamform_max.argtypes = [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ,
                                            ct.c_int, ct.c_int, _P_F, _P_I, _P_I]
        lib.oracle_beamform_max_range.argtypes = [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ,
                                                  ct.c_int, ct.c_int, ct.c_longlong,
                                                  ct.c_longlong, _P_F, _P_I, _P_I]
        lib.oracle_beamform_differential.argtypes = [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ,
                                                     ct.c_int, _P_F, _P_I]
        lib.oracle_prestack_f64.argtypes = [_P_D, _P_D, _SZ, _SZ, _SZ, _SZ, ct.c_int, _P_D]
        lib.oracle_beamform_f64.argtypes = [_P_D, _P_I, _P_D, _SZ, _SZ, _SZ, _SZ,
                                            ct.c_int, ct.c_int, ct.c_longlong,
                                            ct.c_longlong, _P_D, _P_U8, _P_I]
        lib.oracle_beamform_max_f64.argtypes = [_P_D, _P_I, _P_D, _SZ, _SZ, _SZ, _SZ,
                                                ct.c_int, ct.c_int, ct.c_longlong,
                                                ct.c_longlong, _P_D, _P_D, _P_I, _P_U8, _P_I]
        for name in ("oracle_delay_bounds", "oracle_prestack", "oracle_beamform",
                     "oracle_beamform_max", "oracle_beamform_max_range",
                     "oracle_beamform_differential", "oracle_prestack_f64",
                     "oracle_beamform_f64",
                     "oracle_beamform_max_f64"):
            getattr(lib, name).restype = None
        lib._typed = True
    return lib


def _nt(n):
    """Explicit thread count: the reference library sets the process-wide
    OpenMP default (beamform.c:122-124), so never rely on it."""
    return int(n) if n and int(n) > 0 else (os.cpu_count() or 1)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a, ptype):
    return a.ctypes.data_as(ptype)


def delay_bounds(tau, beta, num_threads=0, flavour="fma"):
    """(K,2) int32 [lo, hi] per source (beamform.c:10-47)."""
    tau = _c(tau, np.int32)
    beta = _c(beta, np.float32)
    K, S, P = tau.shape
    out = np.empty((K, 2), np.int32)
    _oracle(flavour).oracle_delay_bounds(_ptr(tau, _P_I), _ptr(beta, _P_F), K, S, P,
                                         _nt(num_threads), _ptr(out, _P_I))
    return out


def prestack(u, alpha, num_threads=0, flavour="fma"):
    """(S,C,T) features, (S,C,P) weights -> (S,T,P) float32 (beamform.c:108-152)."""
    u = _c(u, np.float32)
    alpha = _c(alpha, np.float32)
    S, C, T = u.shape
    P = alpha.shape[2]
    out = np.empty((S, T, P), np.float32)
    _oracle(flavour).oracle_prestack(_ptr(u, _P_F), _ptr(alpha, _P_F), T, S, C, P,
                                     _nt(num_threads), _ptr(out, _P_F))
    return out


def beamform(F, tau, beta, out_of_bounds=0, num_threads=0, flavour="fma", init=None):
    """Full (K,T) beams from prestacked (S,T,P) features (beamform.c:154-228).

    Invalid cells keep ``init`` (default zeros, as the reference wrapper's
    zero-filled buffer, beampower.py:123)."""
    F = _c(F, np.float32)
    tau = _c(tau, np.int32)
    beta = _c(beta, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    beam = np.zeros((K, T), np.float32) if init is None else _c(init, np.float32).copy()
    scratch = np.empty(2 * K, np.int32)
    _oracle(flavour).oracle_beamform(_ptr(F, _P_F), _ptr(tau, _P_I), _ptr(beta, _P_F),
                                     T, K, S, P, int(out_of_bounds), _nt(num_threads),
                                     _ptr(beam, _P_F), _ptr(scratch, _P_I))
    return beam


def beamform_max(F, tau, beta, out_of_bounds=0, num_threads=0, flavour="fma", init=None,
                 t_begin=0, t_end=None):
    """(beam_max, beam_argmax) from prestacked (S,T,P) features
    (beamform.c:230-336) for the samples [t_begin, t_end) of the window
    (default all T).  beam_max keeps ``init`` where no source is valid."""
    F = _c(F, np.float32)
    tau = _c(tau, np.int32)
    beta = _c(beta, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    t_end = T if t_end is None else int(t_end)
    t_begin = int(t_begin)
    if not 0 <= t_begin <= t_end <= T:
        raise ValueError(f"sample range [{t_begin}, {t_end}) outside [0, {T}]")
    W = t_end - t_begin
    bmax = np.zeros(W, np.float32) if init is None else _c(init, np.float32).copy()
    barg = np.zeros(W, np.int32)
    scratch = np.empty(2 * K, np.int32)
    _oracle(flavour).oracle_beamform_max_range(_ptr(F, _P_F), _ptr(tau, _P_I), _ptr(beta, _P_F),
                                               T, K, S, P, int(out_of_bounds), _nt(num_threads),
                                               t_begin, t_end, _ptr(bmax, _P_F), _ptr(barg, _P_I),
                                               _ptr(scratch, _P_I))
    return bmax, barg


def beamform_differential(F, tau, beta, num_threads=0, flavour="fma", init=None):
    """(K,) beams at the centre sample (T - 1) // 2 from prestacked (S,T,P)
    features (beamform.c:338-394).  Sources with a used delay pointing
    outside [0, T) keep ``init`` (default zeros, beampower.py:198)."""
    F = _c(F, np.float32)
    tau = _c(tau, np.int32)
    beta = _c(beta, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    beam = np.zeros(K, np.float32) if init is None else _c(init, np.float32).copy()
    scratch = np.empty(2 * K, np.int32)
    _oracle(flavour).oracle_beamform_differential(_ptr(F, _P_F), _ptr(tau, _P_I),
                                                  _ptr(beta, _P_F), T, K, S, P,
                                                  _nt(num_threads), _ptr(beam, _P_F),
                                                  _ptr(scratch, _P_I))
    return beam


def prestack_f64(u, alpha, num_threads=0):
    """FP64 twin of prestack (beamform.c:108-152): (S,C,T), (S,C,P) ->
    (S,T,P) float64, accumulated in ascending c from 0."""
    u = _c(u, np.float64)
    alpha = _c(alpha, np.float64)
    S, C, T = u.shape
    P = alpha.shape[2]
    out = np.empty((S, T, P), np.float64)
    _oracle("fma").oracle_prestack_f64(_ptr(u, _P_D), _ptr(alpha, _P_D), T, S, C, P,
                                       _nt(num_threads), _ptr(out, _P_D))
    return out


def beamform_f64(F, tau, beta, out_of_bounds=0, t_begin=0, t_end=None, num_threads=0):
    """FP64 beams over [t_begin, t_end): returns (beam (K,W) float64, valid (K,W) bool)."""
    F = _c(F, np.float64)
    tau = _c(tau, np.int32)
    beta = _c(beta, np.float64)
    S, T, P = F.shape
    K = tau.shape[0]
    t_end = T if t_end is None else int(t_end)
    W = t_end - int(t_begin)
    beam = np.empty((K, W), np.float64)
    valid = np.empty((K, W), np.uint8)
    scratch = np.empty(2 * K, np.int32)
    _oracle("fma").oracle_beamform_f64(_ptr(F, _P_D), _ptr(tau, _P_I), _ptr(beta, _P_D),
                                       T, K, S, P, int(out_of_bounds), _nt(num_threads),
                                       int(t_begin), t_end, _ptr(beam, _P_D),
                                       _ptr(valid, _P_U8), _ptr(scratch, _P_I))
    return beam, valid.astype(bool)


def beamform_max_f64(F, tau, beta, out_of_bounds=0, t_begin=0, t_end=None, num_threads=0):
    """FP64 max/argmax over [t_begin, t_end) with the runner-up value:
    returns (max, second, argmax, valid) arrays of length t_end - t_begin."""
    F = _c(F, np.float64)
    tau = _c(tau, np.int32)
    beta = _c(beta, np.float64)
    S, T, P = F.shape
    K = tau.shape[0]
    t_end = T if t_end is None else int(t_end)
    W = t_end - int(t_begin)
    mx = np.empty(W, np.float64)
    sec = np.empty(W, np.float64)
    arg = np.empty(W, np.int32)
    valid = np.empty(W, np.uint8)
    scratch = np.empty(2 * K, np.int32)
    _oracle("fma").oracle_beamform_max_f64(_ptr(F, _P_D), _ptr(tau, _P_I), _ptr(beta, _P_D),
                                           T, K, S, P, int(out_of_bounds), _nt(num_threads),
                                           int(t_begin), t_end, _ptr(mx, _P_D), _ptr(sec, _P_D),
                                           _ptr(arg, _P_I), _ptr(valid, _P_U8),
                                           _ptr(scratch, _P_I))
    return mx, sec, arg, valid.astype(bool)


# ---------------------------------------------------------------------------
# The reference engine compiled from its own sources (oracle/_ref).
# ---------------------------------------------------------------------------

def ref_available(flavour="cmake"):
    return os.path.exists(os.path.join(REF_DIR, f"beamform_ref_{flavour}.so"))


def _ref(flavour):
    if flavour not in ("cmake", "make"):
        raise ValueError(f"flavour must be 'cmake' or 'make', not {flavour!r}")
    lib = _load(os.path.join(REF_DIR, f"beamform_ref_{flavour}.so"))
    if not getattr(lib, "_typed", False):
        # argument tables of beampower/core.py:14-49
        lib.beamform.argtypes = [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int, ct.c_int, _P_F]
        lib.beamform_max.argtypes = [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int,
                                     ct.c_int, _P_F, _P_I]
        lib.prestack_waveform_features.argtypes = [_P_F, _P_F, _SZ, _SZ, _SZ, _SZ,
                                                   ct.c_int, _P_F]
        lib.beamform_differential.argtypes = [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ,
                                              ct.c_int, _P_F]
        for name in ("beamform", "beamform_max", "prestack_waveform_features",
                     "beamform_differential"):
            getattr(lib, name).restype = None
        lib._typed = True
    return lib


def ref_prestack(u, alpha, num_threads=1, flavour="cmake"):
    u = _c(u, np.float32)
    alpha = _c(alpha, np.float32)
    S, C, T = u.shape
    P = alpha.shape[2]
    out = np.zeros((S, T, P), np.float32)
    _ref(flavour).prestack_waveform_features(_ptr(u, _P_F), _ptr(alpha, _P_F), T, S, C, P,
                                             _nt(num_threads), _ptr(out, _P_F))
    return out


def ref_beamform(F, tau, beta, out_of_bounds=0, num_threads=1, flavour="cmake"):
    F = _c(F, np.float32)
    tau = _c(tau, np.int32)
    beta = _c(beta, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    beam = np.zeros((K, T), np.float32)
    _ref(flavour).beamform(_ptr(F, _P_F), _ptr(tau, _P_I), _ptr(beta, _P_F), T, K, S, P,
                           int(out_of_bounds), _nt(num_threads), _ptr(beam, _P_F))
    return beam


def ref_beamform_max(F, tau, beta, out_of_bounds=0, num_threads=1, flavour="cmake"):
    F = _c(F, np.float32)
    tau = _c(tau, np.int32)
    beta = _c(beta, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    bmax = np.zeros(T, np.float32)
    barg = np.zeros(T, np.int32)
    _ref(flavour).beamform_max(_ptr(F, _P_F), _ptr(tau, _P_I), _ptr(beta, _P_F), T, K, S, P,
                               int(out_of_bounds), _nt(num_threads), _ptr(bmax, _P_F),
                               _ptr(barg, _P_I))
    return bmax, barg


def ref_beamform_differential(F, tau, beta, num_threads=1, flavour="cmake"):
    """The reference's beamform_differential (beamform.c:338-394).  Keep
    num_threads=1: its per-source bounds are shared between threads."""
    F = _c(F, np.float32)
    tau = _c(tau, np.int32)
    beta = _c(beta, np.float32)
    S, T, P = F.shape
    K = tau.shape[0]
    beam = np.zeros(K, np.float32)
    _ref(flavour).beamform_differential(_ptr(F, _P_F), _ptr(tau, _P_I), _ptr(beta, _P_F), T, K,
                                        S, P, _nt(num_threads), _ptr(beam, _P_F))
    return beam
