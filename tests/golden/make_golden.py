"""Write the golden fixtures of tests/golden/ from the reference engine.

The reference's C engine (beampower/src/beamform.c) is compiled from its
own sources by `make -C oracle ref` into oracle/_ref/ in both of its build
flavours (CMake/pip: FMA-contracted; Makefile: -std=c99, no contraction),
and called here through the reference's own argument tables
(beampower/core.py:14-49).  Only inputs and outputs are stored.

    python tests/golden/make_golden.py

Cases avoid the one undefined behaviour of the reference (strict
reduce='max' with negative moveouts reads outside the feature array,
beamform.c:281-282).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import oracle  # noqa: E402
from beampower_amd import synthetic  # noqa: E402

FLAVOURS = ("cmake", "make")


def inputs_digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode() + str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def run_case(u, alpha, tau, beta, reduces=("max", "none"), oobs=(0, 1)):
    out = {}
    for fl in FLAVOURS:
        F = oracle.ref_prestack(u, alpha, num_threads=4, flavour=fl)
        out[f"F_{fl}"] = F
        for oob in oobs:
            if "max" in reduces:
                m, a = oracle.ref_beamform_max(F, tau, beta, oob, num_threads=4, flavour=fl)
                out[f"max_{oob}_{fl}"] = m
                out[f"argmax_{oob}_{fl}"] = a
            if "none" in reduces:
                out[f"none_{oob}_{fl}"] = oracle.ref_beamform(F, tau, beta, oob,
                                                              num_threads=4, flavour=fl)
    return out


def differential_case(seed, T, K=80, nonfinite=False):
    """mode='differential': moveouts on both sides of zero lag, so that some
    sources fall outside [0, T) at the centre sample on either side.  The
    reference runs single-threaded: its per-source bounds are shared between
    OpenMP threads (beamform.c:353-354,370-380)."""
    half = T // 2 + 4
    u, alpha, tau, beta = synthetic.random_case(seed, S=6, C=3, P=2, T=T, K=K,
                                                tau_range=(-half, half))
    if nonfinite:
        # plant NaN / inf where valid sources read them
        t = (T - 1) // 2
        used = beta != 0
        lo = np.where(used[:, :, None], tau, T).min(axis=(1, 2))
        hi = np.where(used[:, :, None], tau, -T).max(axis=(1, 2))
        valid = np.flatnonzero(used.any(1) & (t + lo >= 0) & (t + hi < T))
        for j, v in zip(valid[:3], (np.nan, np.inf, -np.inf)):
            s = int(np.flatnonzero(used[j])[0])
            if j == valid[2]:
                beta[j, s] = v
            else:
                u[s, 0, t + tau[j, s, 0]] = v
    out = {}
    for fl in FLAVOURS:
        F = oracle.ref_prestack(u, alpha, num_threads=1, flavour=fl)
        out[f"F_{fl}"] = F
        out[f"diff_{fl}"] = oracle.ref_beamform_differential(F, tau, beta, num_threads=1,
                                                             flavour=fl)
    return dict(u=u, alpha=alpha, tau=tau, beta=beta), out


def signed_zero_case(seed, S=5, C=2, P=2, T=64, K=40, tau_max=20):
    """Beams that end at +0.0 or -0.0.  Features and most weights are tiny
    (about 1e-30), so their products underflow to a zero whose sign is the
    product's: fmaf(-1e-30f, 1e-30f, +0.0f) is -0.0.  Unused stations (0 and
    -0.0 weights) and, in flexible mode, terms outside [0, T) are skipped by
    the reference (beamform.c:61-62,94-99) and must keep a zero's sign; -0.0
    and +0.0 tie under its '>' (beamform.c:326), so the lower index wins.  A
    few sources with ordinary weights give nonzero beams beside them."""
    rng = np.random.default_rng(seed)
    u = rng.choice(np.array([-1e-30, 1e-30, -2e-30, 0.0, -0.0], np.float32), (S, C, T),
                   p=(0.6, 0.1, 0.1, 0.1, 0.1))
    # mostly negative features: ordinary weights give negative beams, so
    # the maximum is often one of the zero beams
    alpha = rng.choice(np.array([1.0, 0.5, 0.0], np.float32), (S, C, P))
    tau = rng.integers(0, tau_max, (K, S, P)).astype(np.int32)
    beta = rng.choice(np.array([1e-30, -1e-30, 0.0, -0.0], np.float32), (K, S),
                      p=(0.6, 0.1, 0.15, 0.15))
    beta[[0, 1, 5, 17, 30]] = rng.choice(np.array([1.0, 0.0], np.float32), (5, S))
    beta[11] = 0.0   # a source with no used station
    tau[23] = tau[8]  # duplicated rows: equal beams at different indices
    beta[23] = beta[8]
    return u, alpha, tau, beta


def save(name, inputs, outputs, **meta):
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **inputs, **outputs,
                        meta=np.array(repr(meta)))
    print(f"{path}: {os.path.getsize(path) / 1024:.1f} KiB")


def main(only=None):
    if not oracle.ref_available("cmake") or not oracle.ref_available("make"):
        sys.exit("oracle/_ref is not built: run `make -C oracle ref`")

    # 9. beams that end at a signed zero
    if only in (None, "signed_zero"):
        u, alpha, tau, beta = signed_zero_case(24)
        save("signed_zero", dict(u=u, alpha=alpha, tau=tau, beta=beta),
             run_case(u, alpha, tau, beta), seed=24)
    if only == "signed_zero":
        return

    # 8. mode='differential', odd and even T, non-finite values
    for name, seed, T, nf in (("differential_odd", 17, 41, False),
                              ("differential_even", 18, 64, False),
                              ("differential_nonfinite", 19, 33, True)):
        inputs, out = differential_case(seed, T, nonfinite=nf)
        save(name, inputs, out, seed=seed)
    if only == "differential":
        return

    # 1. small random, strict + flexible, mixed weights (zeros, negatives,
    #    -0.0, a source with no used station)
    u, alpha, tau, beta = synthetic.random_case(11, S=3, C=3, P=2, T=64, K=17)
    save("random_small", dict(u=u, alpha=alpha, tau=tau, beta=beta),
         run_case(u, alpha, tau, beta), seed=11)

    # 2. negative moveouts: flexible max/none and strict none are defined
    u, alpha, tau, beta = synthetic.random_case(12, S=4, C=2, P=2, T=80, K=23,
                                                negative_delays=True)
    out = run_case(u, alpha, tau, beta, reduces=("max", "none"), oobs=(1,))
    out.update({k: v for k, v in run_case(u, alpha, tau, beta, reduces=("none",),
                                          oobs=(0,)).items() if k.startswith("none")})
    save("negative_delays", dict(u=u, alpha=alpha, tau=tau, beta=beta), out, seed=12)

    # 3. ties: duplicated moveout rows and constant features
    rng = np.random.default_rng(13)
    S, C, P, T, K = 5, 3, 2, 96, 12
    u = np.abs(rng.standard_normal((S, C, T))).astype(np.float32)
    u[:, :, 40:60] = 1.0
    alpha = np.ones((S, C, P), np.float32)
    tau = rng.integers(0, 20, (K, S, P)).astype(np.int32)
    tau[4] = tau[1]
    tau[9] = tau[1]
    tau[7] = tau[2]
    beta = np.ones((K, S), np.float32)
    save("ties", dict(u=u, alpha=alpha, tau=tau, beta=beta),
         run_case(u, alpha, tau, beta), seed=13)

    # 4. non-finite features and weights
    u, alpha, tau, beta = synthetic.random_case(14, S=4, C=3, P=2, T=72, K=15)
    u[1, 0, 10] = np.nan
    u[2, 2, 30] = np.inf
    u[0, 1, 50] = -np.inf
    beta[5, 2] = np.inf
    beta[6, 1] = np.nan
    save("nonfinite", dict(u=u, alpha=alpha, tau=tau, beta=beta),
         run_case(u, alpha, tau, beta), seed=14)

    # 5. one phase and three phases
    for P, seed in ((1, 15), (3, 16)):
        u, alpha, tau, beta = synthetic.random_case(seed, S=6, C=3, P=P, T=90, K=33)
        save(f"phases{P}", dict(u=u, alpha=alpha, tau=tau, beta=beta),
             run_case(u, alpha, tau, beta), seed=seed)

    # 6. plain-method workload, first 2,000 samples and first 700 sources
    w = synthetic.plain_c1()
    u = np.ascontiguousarray(w.waveform_features[:, :, :2000])
    tau = np.ascontiguousarray(w.time_delays[:700])
    beta = np.ascontiguousarray(w.weights_sources[:700])
    save("plain_mini", dict(u=u, alpha=w.weights_phases, tau=tau, beta=beta),
         run_case(u, w.weights_phases, tau, beta, reduces=("max",)), seed=1,
         note="plain_c1() sliced to T=2000, K=700")

    # 7. plain-method workload in full: outputs only, inputs regenerated
    #    from the seed and checked against the stored digest
    out = {}
    for fl in FLAVOURS:
        F = oracle.ref_prestack(w.waveform_features, w.weights_phases, num_threads=8,
                                flavour=fl)
        m, a = oracle.ref_beamform_max(F, w.time_delays, w.weights_sources, 0,
                                       num_threads=8, flavour=fl)
        out[f"max_0_{fl}"] = m
        out[f"argmax_0_{fl}"] = a
    digest = inputs_digest(w.waveform_features, w.time_delays, w.weights_phases,
                           w.weights_sources)
    save("plain_full", {"events": np.array(w.events, np.int64)}, out, seed=1,
         inputs_sha256=digest)


if __name__ == "__main__":
    # `make_golden.py differential` writes only the mode='differential' cases,
    # `make_golden.py signed_zero` only the signed-zero case
    main(sys.argv[1] if len(sys.argv) > 1 else None)
