"""The tolerance contract between this engine and the reference's Makefile
build (Makefile:69: -std=c99, so `beam += w * x` is a rounded product and a
rounded sum, no fused multiply-add).

The engine computes every beam as the FMA chain of the reference's pip/CMake
build (CMakeLists.txt:28-33), so against the Makefile build its beams
differ by rounding only.  The contract:

* every beam within 1e-5 of the Makefile beam, relative to the sum of the
  absolute values of its terms (sum over used, in-range (s, p) of
  |beta[k,s] F[s,t+tau,p]|: the scale a rounding error of a sum is relative
  to, which stays meaningful where terms cancel);
* invalid cells left as they were (0 from the wrapper), non-finite beams at
  the same cells;
* reduce='max': the maximum within the same tolerance, and the argmax equal
  wherever the Makefile build's top-two gap exceeds the tolerance (closer
  candidates may swap under a last-bit change).

The Makefile beams and gaps come from the oracle's "nofma" flavour, which
the golden fixtures pin bit for bit to that build.
"""
import numpy as np

from oracle import oracle

RTOL = 1e-5


def term_scale(F, tau, beta, oob):
    """sum |beta F| over each beam's terms, by the FP64 twin (exact enough)."""
    absF = np.abs(np.nan_to_num(F.astype(np.float64), nan=0.0, posinf=0.0, neginf=0.0))
    absB = np.abs(np.nan_to_num(beta.astype(np.float64), nan=0.0, posinf=0.0, neginf=0.0))
    scale, _ = oracle.beamform_f64(absF, tau, absB, oob)
    return scale


def make_beams(F, tau, beta, oob):
    """Makefile-build beams with invalid cells at -inf (never a candidate)."""
    K, T = tau.shape[0], F.shape[1]
    return oracle.beamform(F, tau, beta, oob, flavour="nofma",
                           init=np.full((K, T), -np.inf, np.float32))


def check_beams(got, F, tau, beta, oob):
    """reduce='none' output (invalid cells 0) against the Makefile build."""
    ref = make_beams(F, tau, beta, oob)
    # validity from the FP64 twin (a valid beam may itself be -inf)
    _, valid = oracle.beamform_f64(F.astype(np.float64), tau, beta.astype(np.float64), oob)
    scale = term_scale(F, tau, beta, oob)
    assert np.all(got[~valid] == 0.0), "invalid cells must stay 0"
    g, r, sc = got[valid].astype(np.float64), ref[valid].astype(np.float64), scale[valid]
    fin = np.isfinite(r)
    assert np.array_equal(np.isfinite(g), fin), "non-finite beams differ"
    nan = np.isnan(r)
    assert np.array_equal(np.isnan(g), nan)
    inf = np.isinf(r)
    assert np.array_equal(g[inf], r[inf])
    err = np.abs(g[fin] - r[fin])
    assert np.all(err <= RTOL * sc[fin] + 1e-37), float((err / np.maximum(sc[fin], 1e-37)).max())


def check_max(got_max, got_arg, F, tau, beta, oob, init=0.0):
    """reduce='max' output against the Makefile build; returns the fraction
    of samples whose top-two gap decides the argmax."""
    beams = make_beams(F, tau, beta, oob).astype(np.float64)
    _, v64 = oracle.beamform_f64(F.astype(np.float64), tau, beta.astype(np.float64), oob)
    cand = v64 & (beams > -np.finfo(np.float32).max)  # NaN is never a candidate
    b = np.where(cand, beams, -np.inf)
    has = cand.any(0)
    scale = term_scale(F, tau, beta, oob)
    sc = np.where(cand, scale, 0.0).max(0)
    order = np.sort(b, axis=0)
    top1 = order[-1]
    top2 = order[-2] if b.shape[0] > 1 else np.full_like(top1, -np.inf)
    ref_arg = np.where(has, np.argmax(b, axis=0), 0)
    assert np.all(got_arg[~has] == 0) and np.all(got_max[~has] == init)
    g = got_max[has].astype(np.float64)
    fin = np.isfinite(top1[has])
    assert np.all(np.abs(g[fin] - top1[has][fin]) <= RTOL * sc[has][fin] + 1e-37)
    assert np.array_equal(g[~fin], top1[has][~fin])
    with np.errstate(invalid="ignore"):
        decided = has & ((top1 - top2) > 2 * RTOL * sc)
    assert np.array_equal(got_arg[decided], ref_arg[decided])
    # where the gap does not decide, the engine's pick is a candidate within
    # the tolerance of the maximum
    und = has & ~decided
    picked = b[got_arg[und], np.flatnonzero(und)]
    assert np.all(cand[got_arg[und], np.flatnonzero(und)])
    with np.errstate(invalid="ignore"):
        close = (picked == top1[und]) | (np.abs(picked - top1[und]) <= 2 * RTOL * sc[und] + 1e-37)
    assert np.all(close)
    return float(decided.sum()) / max(1, int(has.sum()))
