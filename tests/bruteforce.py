"""Independent NumPy restatement of the beamforming maths, for tiny inputs.

Written from the closed form of the reference tutorial
(notebooks/3_beampower.ipynb:400-416):
    U[s,p](t) = sum_c alpha[s,c,p] u[s,c](t)
    b_k(t)    = sum_s beta[k,s] sum_p U[s,p](t + tau[k,s,p])
with the bounds rules of beampower/src/beamform.c (stations with a zero
weight are not used; strict mode requires every used moveout in range,
flexible mode drops out-of-range terms).  Vectorised over time, looped over
sources, in float64.  `mistake` plants a single plausible bug so tests can
show that the checks catch it.
"""
import numpy as np


def prestack(u, alpha):
    return np.einsum("sct,scp->stp", u.astype(np.float64), alpha.astype(np.float64))


def beams(F, tau, beta, flexible, mistake=None):
    """Full beams (K,T) float64 and validity mask (K,T)."""
    F = np.asarray(F, np.float64)
    S, T, P = F.shape
    K = tau.shape[0]
    out = np.zeros((K, T))
    valid = np.zeros((K, T), bool)
    t = np.arange(T)
    for k in range(K):
        used = beta[k] != 0
        if mistake == "keep_zero_weights":
            used = np.ones(S, bool)
        if used.any():
            lo = tau[k, used].min()
            hi = tau[k, used].max()
            ok = (t + lo >= 0) & (t + hi < T)
        else:
            ok = np.zeros(T, bool)
        if flexible:
            ok = np.ones(T, bool)
        acc = np.zeros(T)
        for s in range(S):
            if not used[s]:
                continue
            w = float(beta[k, s])
            if mistake == "wrong_sign":
                w = -w
            for p in range(P):
                if mistake == "drop_last_phase" and p == P - 1 and P > 1:
                    continue
                d = int(tau[k, s, p])
                if mistake == "transposed_phase":
                    d = int(tau[k, s, P - 1 - p])
                if mistake == "off_by_one":
                    d += 1
                idx = t + d
                inr = (idx >= 0) & (idx < T)
                x = np.zeros(T)
                x[inr] = F[s, idx[inr], p]
                term = w * x
                if flexible:
                    term = np.where(inr, term, 0.0)
                acc = acc + term
        out[k] = np.where(ok, acc, 0.0)
        valid[k] = ok
    return out, valid


def reduce_max(b, valid):
    """Ascending-k strict '>' scan from -FLT_MAX (beamform.c:265-334)."""
    floor = -float(np.finfo(np.float32).max)
    T = b.shape[1]
    best = np.full(T, floor)
    arg = np.zeros(T, np.int64)
    for k in range(b.shape[0]):
        v = np.where(valid[k], b[k], -np.inf)
        upd = v > best
        best = np.where(upd, v, best)
        arg = np.where(upd, k, arg)
    return best, arg, best > floor
