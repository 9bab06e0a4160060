"""Pins the CPU oracle (oracle/beamform_oracle.c).

* against the golden fixtures written by the reference engine compiled from
  its own sources (tests/golden/make_golden.py), bit for bit, per build
  flavour;
* against the reference engine itself when oracle/_ref is built;
* against an independent NumPy FP64 brute force on tiny inputs, with
  planted mistakes that the checks must catch;
* against closed forms (impulse and constant features, exact prestack).
"""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, bits_equal, load_golden
from oracle import oracle
import bruteforce as bf
from beampower_amd import synthetic

FLAV = {"cmake": "fma", "make": "nofma"}


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("ref_flavour", ["cmake", "make"])
def test_oracle_matches_golden(case, ref_flavour):
    g = load_golden(case)
    fl = FLAV[ref_flavour]
    F = oracle.prestack(g["u"], g["alpha"], flavour=fl)
    assert bits_equal(F, g[f"F_{ref_flavour}"])
    for oob in (0, 1):
        key = f"max_{oob}_{ref_flavour}"
        if key in g:
            m, a = oracle.beamform_max(F, g["tau"], g["beta"], oob, flavour=fl)
            assert bits_equal(m, g[key]), key
            assert np.array_equal(a, g[f"argmax_{oob}_{ref_flavour}"]), key
        key = f"none_{oob}_{ref_flavour}"
        if key in g:
            b = oracle.beamform(F, g["tau"], g["beta"], oob, flavour=fl)
            assert bits_equal(b, g[key]), key


def test_oracle_matches_golden_plain_full():
    g = load_golden("plain_full")
    w = synthetic.plain_c1()
    from make_golden import inputs_digest  # noqa: E402  (tests/golden on sys.path below)
    assert inputs_digest(w.waveform_features, w.time_delays, w.weights_phases,
                         w.weights_sources) == eval(str(g["meta"]))["inputs_sha256"]
    for ref_flavour, fl in FLAV.items():
        F = oracle.prestack(w.waveform_features, w.weights_phases, flavour=fl)
        m, a = oracle.beamform_max(F, w.time_delays, w.weights_sources, 0, flavour=fl)
        assert bits_equal(m, g[f"max_0_{ref_flavour}"])
        assert np.array_equal(a, g[f"argmax_0_{ref_flavour}"])


@pytest.mark.skipif(not oracle.ref_available("cmake"), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [101, 102, 103])
def test_oracle_matches_reference_build_live(seed):
    u, alpha, tau, beta = synthetic.random_case(seed, S=7, C=3, P=2, T=300, K=90,
                                                tau_range=(0, 120))
    for ref_flavour, fl in FLAV.items():
        Fr = oracle.ref_prestack(u, alpha, num_threads=2, flavour=ref_flavour)
        F = oracle.prestack(u, alpha, flavour=fl)
        assert bits_equal(F, Fr)
        for oob in (0, 1):
            mr, ar = oracle.ref_beamform_max(Fr, tau, beta, oob, num_threads=3,
                                             flavour=ref_flavour)
            m, a = oracle.beamform_max(F, tau, beta, oob, flavour=fl)
            assert bits_equal(m, mr) and np.array_equal(a, ar)
            br = oracle.ref_beamform(Fr, tau, beta, oob, num_threads=3, flavour=ref_flavour)
            assert bits_equal(oracle.beamform(F, tau, beta, oob, flavour=fl), br)


def test_flavours_differ():
    """Contraction changes rounding: the two flavours are distinct oracles."""
    u, alpha, tau, beta = synthetic.random_case(5, S=8, C=3, P=2, T=400, K=50)
    Fa = oracle.prestack(u, alpha, flavour="fma")
    Fb = oracle.prestack(u, alpha, flavour="nofma")
    ba = oracle.beamform(Fa, tau, beta, 1, flavour="fma")
    bb = oracle.beamform(Fa, tau, beta, 1, flavour="nofma")
    assert not bits_equal(ba, bb)
    assert np.allclose(ba, bb, rtol=1e-4, atol=1e-4)
    assert np.allclose(Fa, Fb, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("oob", [0, 1])
def test_oracle_vs_bruteforce_fp64(seed, oob):
    u, alpha, tau, beta = synthetic.random_case(seed, S=4, C=3, P=2, T=50, K=9,
                                                tau_range=(0, 15),
                                                negative_delays=bool(oob))
    F64 = bf.prestack(u, alpha)
    b_ref, v_ref = bf.beams(F64, tau, beta, flexible=bool(oob))
    # FP64 twin of the oracle: same sums in double, ~exact agreement
    b64, v64 = oracle.beamform_f64(F64, tau, beta, oob)
    assert np.array_equal(v64, v_ref)
    np.testing.assert_allclose(b64[v64], b_ref[v_ref], rtol=1e-12, atol=1e-12)
    # FP32 oracle: within float32 rounding of the FP64 sums
    F32 = oracle.prestack(u, alpha)
    np.testing.assert_allclose(F32, F64, rtol=1e-6, atol=1e-6)
    b32 = oracle.beamform(F32, tau, beta, oob)
    scale = np.abs(beta).sum(1, keepdims=True) * np.abs(F64).max() * tau.shape[2] + 1e-30
    assert np.all(np.abs(b32 - b_ref) <= 1e-6 * scale)
    # reduction rule: max value and its lowest index
    best, arg, any_ = bf.reduce_max(b_ref, v_ref)
    m32, a32 = oracle.beamform_max(F32, tau, beta, oob)
    gap_ok = np.ones_like(any_)
    srt = np.sort(np.where(v_ref, b_ref, -np.inf), axis=0)
    if srt.shape[0] > 1:
        with np.errstate(invalid="ignore"):
            gap_ok = (srt[-1] - srt[-2]) > 1e-4 * (np.abs(srt[-1]) + 1)
    assert np.array_equal(a32[any_ & gap_ok], arg[any_ & gap_ok])
    assert np.all(m32[~any_] == 0.0) and np.all(a32[~any_] == 0)


@pytest.mark.parametrize("mistake", ["keep_zero_weights", "wrong_sign", "drop_last_phase",
                                     "transposed_phase", "off_by_one"])
def test_planted_mistakes_are_caught(mistake):
    """Each plausible bug, planted in the NumPy restatement, disagrees with
    the reference's golden beams -- so these checks can fail."""
    g = load_golden("random_small")
    F64 = bf.prestack(g["u"], g["alpha"])
    caught = []
    for oob in (0, 1):
        gold = g[f"none_{oob}_cmake"].astype(np.float64)
        good, _ = bf.beams(F64, g["tau"], g["beta"], flexible=bool(oob))
        bad, _ = bf.beams(F64, g["tau"], g["beta"], flexible=bool(oob), mistake=mistake)
        assert np.allclose(good, gold, rtol=1e-5, atol=1e-5)
        caught.append(not np.allclose(bad, gold, rtol=1e-5, atol=1e-5))
    assert any(caught)


def test_closed_form_impulse():
    """A unit impulse at (s0, t0, p0) puts beta[k,s0] at t0 - tau[k,s0,p0]."""
    rng = np.random.default_rng(7)
    S, T, P, K = 5, 120, 2, 30
    tau = rng.integers(0, 30, (K, S, P)).astype(np.int32)
    beta = rng.uniform(0.5, 2.0, (K, S)).astype(np.float32)
    F = np.zeros((S, T, P), np.float32)
    s0, t0, p0 = 3, 70, 1
    F[s0, t0, p0] = 1.0
    b = oracle.beamform(F, tau, beta, 1)
    for k in range(K):
        expect = np.zeros(T, np.float32)
        hits = [(s0, p) for p in range(P) if p == p0]
        for s, p in hits:
            expect[t0 - tau[k, s, p]] += beta[k, s]
        assert bits_equal(b[k], expect)


def test_closed_form_constant():
    """Constant features c give b = c * P * sum(used beta) (exact in binary)."""
    S, T, P, K = 6, 80, 2, 20
    rng = np.random.default_rng(8)
    tau = rng.integers(0, 10, (K, S, P)).astype(np.int32)
    beta = rng.choice([0.0, 0.5, 1.0, 2.0], (K, S)).astype(np.float32)
    F = np.full((S, T, P), 0.25, np.float32)
    expect = (0.25 * P * beta.sum(1)).astype(np.float32)
    for oob in (0, 1):
        b = oracle.beamform(F, tau, beta, oob)
        # every moveout is in [0, 10): samples below T-10 see all terms
        assert np.array_equal(b[:, :T - 10], np.repeat(expect[:, None], T - 10, 1))


def test_closed_form_prestack():
    rng = np.random.default_rng(9)
    S, C, T, P = 4, 3, 64, 2
    u = rng.integers(-8, 8, (S, C, T)).astype(np.float32)
    alpha = rng.choice([0.0, 0.5, 1.0], (S, C, P)).astype(np.float32)
    for fl in ("fma", "nofma"):
        F = oracle.prestack(u, alpha, flavour=fl)
        np.testing.assert_array_equal(F, np.einsum("sct,scp->stp", u, alpha))


# ---- FP64 twins (the ground truth of the 16-bit precision reports) ---------

def test_prestack_f64_vs_bruteforce():
    """oracle.prestack_f64 (beamform.c:108-152 in double) equals the NumPy
    contraction sum_c alpha[s,c,p] u[s,c,t] to double rounding.  Distinct
    dimensions (S=5, C=3, P=2, T=37) and non-symmetric values make a wrong
    index or a transposed operand fail."""
    rng = np.random.default_rng(91)
    S, C, P, T = 5, 3, 2, 37
    u = rng.standard_normal((S, C, T))
    alpha = rng.uniform(-2, 2, (S, C, P))
    F = oracle.prestack_f64(u, alpha)
    assert F.shape == (S, T, P) and F.dtype == np.float64
    np.testing.assert_allclose(F, bf.prestack(u, alpha), rtol=1e-13, atol=1e-13)
    # each plausible mistake moves the result away from it
    mistakes = {
        "phases_swapped": np.einsum("sct,scp->stp", u, alpha[:, :, ::-1]),
        "channels_reversed": np.einsum("sct,scp->stp", u[:, ::-1], alpha),
        "last_channel_dropped": np.einsum("sct,scp->stp", u[:, :-1], alpha[:, :-1]),
        "stations_shifted": np.einsum("sct,scp->stp", np.roll(u, 1, 0), alpha),
        "time_shifted": np.einsum("sct,scp->stp", np.roll(u, 1, 2), alpha),
    }
    for name, bad in mistakes.items():
        assert not np.allclose(F, bad, rtol=1e-6, atol=1e-6), name


def test_prestack_f64_closed_form_and_fp32_twin():
    """Exact inputs (small integers, alpha in {0, +-0.5, 1, 2}): the FP64
    prestack is the exact integer sum, and equals the FP32 oracle bit for
    bit; on the reference's golden prestacks it is within FP32 rounding."""
    rng = np.random.default_rng(92)
    S, C, P, T = 4, 3, 3, 50
    u = rng.integers(-64, 64, (S, C, T)).astype(np.float64)
    alpha = rng.choice([0.0, 0.5, -0.5, 1.0, 2.0], (S, C, P))
    exact = np.einsum("sct,scp->stp", (u * 2).astype(np.int64),
                      (alpha * 2).astype(np.int64)) / 4.0
    F = oracle.prestack_f64(u, alpha)
    assert np.array_equal(F, exact)
    F32 = oracle.prestack(u.astype(np.float32), alpha.astype(np.float32))
    assert np.array_equal(F32.astype(np.float64), F)
    for case in GOLDEN_CASES:
        g = load_golden(case)
        F64 = oracle.prestack_f64(g["u"], g["alpha"])
        gold = g["F_cmake"].astype(np.float64)
        fin = np.isfinite(gold)
        np.testing.assert_allclose(F64[fin], gold[fin], rtol=2e-6, atol=1e-6)


def test_delay_bounds_vs_numpy():
    u, alpha, tau, beta = synthetic.random_case(21, S=6, C=1, P=3, T=10, K=40,
                                                negative_delays=True)
    lohi = oracle.delay_bounds(tau, beta)
    for k in range(tau.shape[0]):
        used = beta[k] != 0
        if used.any():
            assert lohi[k, 0] == tau[k, used].min() and lohi[k, 1] == tau[k, used].max()
        else:
            assert lohi[k, 0] == np.iinfo(np.int32).max
            assert lohi[k, 1] == np.iinfo(np.int32).min


def test_plain_known_answer():
    """Injected events are recovered: at each event origin the argmax is the
    event's source or one with an identical moveout row."""
    w = synthetic.plain_c1()
    F = oracle.prestack(w.waveform_features, w.weights_phases)
    m, a = oracle.beamform_max(F, w.time_delays, w.weights_sources, 0)
    tau = w.time_delays
    for k_e, t_e in w.events:
        k = a[t_e]
        assert np.array_equal(tau[k], tau[k_e])
        assert k == np.flatnonzero((tau == tau[k_e]).all(axis=(1, 2)))[0]


@pytest.mark.parametrize("oob", [0, 1])
def test_fp64_max_vs_bruteforce(oob):
    """FP64 max/argmax/runner-up over a time range agrees with the NumPy
    brute force (reduction rule, bounds, and the range offset)."""
    u, alpha, tau, beta = synthetic.random_case(77, S=5, C=3, P=2, T=90, K=23,
                                                tau_range=(0, 20), negative_delays=bool(oob))
    F = bf.prestack(u, alpha)
    b, v = bf.beams(F, tau, beta, flexible=bool(oob))
    best, arg, any_ = bf.reduce_max(b, v)
    mx, sec, a, val = oracle.beamform_max_f64(F, tau, beta.astype(np.float64), oob,
                                              t_begin=10, t_end=70)
    assert np.array_equal(val, any_[10:70])
    np.testing.assert_allclose(mx[val], best[10:70][val], rtol=1e-12)
    assert np.array_equal(a[val], arg[10:70][val])
    masked = np.where(v, b, -np.inf)[:, 10:70]
    srt = np.sort(masked, axis=0)
    ok = val & np.isfinite(srt[-2])
    np.testing.assert_allclose(sec[ok], srt[-2][ok], rtol=1e-12)


# ---- mode='differential' (beamform.c:338-394) ------------------------------

DIFF_CASES = ("differential_odd", "differential_even", "differential_nonfinite")


@pytest.mark.parametrize("case", DIFF_CASES)
@pytest.mark.parametrize("ref_flavour", ["cmake", "make"])
def test_differential_oracle_matches_golden(case, ref_flavour):
    g = load_golden(case)
    fl = FLAV[ref_flavour]
    F = oracle.prestack(g["u"], g["alpha"], flavour=fl)
    assert bits_equal(F, g[f"F_{ref_flavour}"])
    b = oracle.beamform_differential(F, g["tau"], g["beta"], flavour=fl)
    assert bits_equal(b, g[f"diff_{ref_flavour}"])


@pytest.mark.parametrize("case", DIFF_CASES)
def test_differential_is_the_centre_column_of_the_direct_beams(case):
    """The differential beam of a source is its direct strict beam at the
    centre sample (T - 1) // 2, and the same sources are left untouched."""
    g = load_golden(case)
    F = oracle.prestack(g["u"], g["alpha"])
    T = F.shape[1]
    K = g["tau"].shape[0]
    direct = oracle.beamform(F, g["tau"], g["beta"], 0, init=np.full((K, T), -7.0, np.float32))
    diff = oracle.beamform_differential(F, g["tau"], g["beta"], init=np.full(K, -7.0, np.float32))
    assert bits_equal(diff, np.ascontiguousarray(direct[:, (T - 1) // 2]))


def test_differential_centre_planted_mistake_is_caught():
    """An FP64 brute force at the right centre sample matches the reference's
    golden beams; the same at T // 2 (a plausible off-by-one for even T)
    does not."""
    g = load_golden("differential_even")
    F64 = bf.prestack(g["u"], g["alpha"])
    T = F64.shape[1]
    gold = g["diff_cmake"].astype(np.float64)
    beams, valid = bf.beams(F64, g["tau"], g["beta"], flexible=False)
    good = np.where(valid[:, (T - 1) // 2], beams[:, (T - 1) // 2], 0.0)
    bad = np.where(valid[:, T // 2], beams[:, T // 2], 0.0)
    assert np.allclose(good, gold, rtol=1e-5, atol=1e-5)
    assert not np.allclose(bad, gold, rtol=1e-5, atol=1e-5)


@pytest.mark.skipif(not oracle.ref_available("cmake"), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed,T", [(201, 1), (202, 2), (203, 257), (204, 1000)])
def test_differential_oracle_matches_reference_build_live(seed, T):
    half = T // 2 + 3
    u, alpha, tau, beta = synthetic.random_case(seed, S=9, C=3, P=2, T=T, K=300,
                                                tau_range=(-half, half))
    for ref_flavour, fl in FLAV.items():
        F = oracle.prestack(u, alpha, flavour=fl)
        br = oracle.ref_beamform_differential(F, tau, beta, num_threads=1, flavour=ref_flavour)
        assert bits_equal(oracle.beamform_differential(F, tau, beta, flavour=fl), br)


# ---- the Makefile build's tolerance contract (tests/contract.py) -----------

@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_fma_chain_meets_makefile_contract(case):
    """The FMA chain (the pip/CMake build, which the GPU engine reproduces
    bit for bit) against the Makefile build's golden outputs: beams within
    1e-5 of the term scale, argmax equal where the top-two gap decides."""
    import contract
    g = load_golden(case)
    Fm = g["F_make"]
    F = oracle.prestack(g["u"], g["alpha"], flavour="fma")
    for oob in (0, 1):
        if f"none_{oob}_make" in g:
            contract.check_beams(g[f"none_{oob}_make"], Fm, g["tau"], g["beta"], oob)
            contract.check_beams(oracle.beamform(F, g["tau"], g["beta"], oob), Fm,
                                 g["tau"], g["beta"], oob)
        if f"max_{oob}_make" in g:
            m, a = oracle.beamform_max(F, g["tau"], g["beta"], oob)
            contract.check_max(m, a, Fm, g["tau"], g["beta"], oob)
            # the checker reproduces the golden Makefile maxima themselves
            contract.check_max(g[f"max_{oob}_make"], g[f"argmax_{oob}_make"], Fm, g["tau"],
                               g["beta"], oob)


def test_makefile_contract_catches_a_planted_error():
    """A last-bit change passes; an error of 1e-4 of the term scale in one
    beam, or a wrong argmax at a decided sample, fails."""
    import contract
    g = load_golden("random_small")
    Fm, tau, beta = g["F_make"], g["tau"], g["beta"]
    b = oracle.beamform(Fm, tau, beta, 1, flavour="nofma")
    contract.check_beams(b, Fm, tau, beta, 1)
    contract.check_beams(np.where(b != 0, np.nextafter(b, np.float32(np.inf)), b), Fm, tau,
                         beta, 1)
    scale = contract.term_scale(Fm, tau, beta, 1)
    k, t = np.unravel_index(np.argmax(scale), scale.shape)
    bad = b.copy()
    bad[k, t] += np.float32(1e-4 * scale[k, t])
    with pytest.raises(AssertionError):
        contract.check_beams(bad, Fm, tau, beta, 1)
    m, a = g["max_1_make"], g["argmax_1_make"]
    frac = contract.check_max(m, a, Fm, tau, beta, 1)
    assert frac > 0.5
    a_bad = a.copy()
    a_bad[np.argmax(m)] = (a[np.argmax(m)] + 1) % tau.shape[0]
    with pytest.raises(AssertionError):
        contract.check_max(m, a_bad, Fm, tau, beta, 1)


@pytest.mark.parametrize("flavour", ["fma", "nofma"])
def test_max_over_a_sample_range_is_the_full_calls_slice(flavour):
    """oracle.beamform_max over [t_begin, t_end) equals the full call's
    values there (validity uses the whole window), for both bounds modes,
    ranges at both ends, in the middle and empty."""
    for case in ("random_small", "negative_delays", "ties", "nonfinite", "signed_zero"):
        g = load_golden(case)
        F = oracle.prestack(g["u"], g["alpha"], flavour=flavour)
        T = F.shape[1]
        for oob in (0, 1):
            if case == "negative_delays" and oob == 0:
                continue
            m, a = oracle.beamform_max(F, g["tau"], g["beta"], oob, flavour=flavour)
            for t0, t1 in ((0, T), (0, 7), (T - 9, T), (T // 3, T // 2), (5, 5)):
                mr, ar = oracle.beamform_max(F, g["tau"], g["beta"], oob, flavour=flavour,
                                             t_begin=t0, t_end=t1)
                assert bits_equal(mr, m[t0:t1]) and np.array_equal(ar, a[t0:t1]), (case, t0)
    with pytest.raises(ValueError):
        oracle.beamform_max(F, g["tau"], g["beta"], 0, t_begin=3, t_end=T + 1)
