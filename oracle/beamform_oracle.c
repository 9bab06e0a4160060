/*
 * beamform_oracle.c -- CPU oracle for the beamforming hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and never by the product library (beampower_amd/lib/beamform_gpu.so),
 * which has no CPU fallback.
 *
 * It restates the algorithm of the reference C engine
 * (ebeauce/beampower v1.0.4, beampower/src/beamform.c) as plain C99 +
 * OpenMP, function by function:
 *
 *   oracle_delay_bounds   <- _find_minmax_time_delays   beamform.c:10-47
 *   stack_strict          <- _beam                      beamform.c:49-72
 *   stack_flexible        <- _beam_check_out_of_bounds  beamform.c:74-105
 *   oracle_prestack       <- prestack_waveform_features beamform.c:108-152
 *   oracle_beamform       <- beamform (reduce='none')   beamform.c:154-228
 *   oracle_beamform_max   <- beamform_max (reduce='max') beamform.c:230-336
 *   oracle_beamform_differential <- beamform_differential beamform.c:338-394
 *
 * Numerics.  The per-(source, time) beam is accumulated in float from
 * 0.0f, stations ascending, phases ascending, exactly as beamform.c:59-70.
 * The reference's two builds differ in floating-point contraction:
 * the CMake/pip build (gnu mode) contracts `beam += w * x` into a fused
 * multiply-add, the Makefile build (-std=c99) does not.  This file is
 * compiled twice: with -DORACLE_FMA the multiply-add is an explicit fmaf
 * (CMake/pip flavour), without it a separate multiply and add under
 * -ffp-contract=off (Makefile flavour).  Each flavour is pinned bit-exactly
 * against the reference compiled from its own sources with its own flags
 * (oracle/Makefile target `ref`).
 *
 * Bounds semantics (normative for the library and its tests):
 *   used(k,s)   <=> beta[k,s] != 0           (beamform.c:31,61,91)
 *   lo_k, hi_k  =  min/max of tau over used stations and all phases,
 *                  INT_MAX/INT_MIN when no station is used (beamform.c:22-23)
 *   strict      :  (k,t) valid iff t+lo_k >= 0 and t+hi_k < T
 *                  (beamform.c:196-201).  The reference's beamform_max
 *                  tests t+lo_k < 0 on a size_t (beamform.c:281-282), which
 *                  is always false, so with negative delays it reads outside
 *                  the feature array; this oracle applies the intended test.
 *                  With delays >= 0 (relative moveouts) both agree.
 *   flexible    :  every (k,t) is valid; a term with t+tau outside [0,T)
 *                  is dropped (beamform.c:97-98 through unsigned wrap-around).
 *   reduce=max  :  ascending k, running max with strict '>' from -FLT_MAX,
 *                  argmax 0; beam_max[t] written only if some candidate beat
 *                  -FLT_MAX, argmax always written (beamform.c:269-334).
 *   reduce=none :  invalid (k,t) cells are left untouched (beamform.c:196-201).
 *
 * The FP64 entry points (suffix _f64) evaluate the same sums in double for
 * the reduced-precision feature forms; they share the bounds semantics.
 */
#include <stddef.h>
#include <stdint.h>
#include <float.h>
#include <limits.h>
#include <math.h>
#include <omp.h>

#ifdef ORACLE_FMA
#define MAC32(acc, w, x) ((acc) = fmaf((w), (x), (acc)))
#else
#define MAC32(acc, w, x) ((acc) = (acc) + (w) * (x))
#endif
#define MAC64(acc, w, x) ((acc) = (acc) + (w) * (x))

static int nthreads_or_default(int num_threads)
{
    return num_threads > 0 ? num_threads : omp_get_max_threads();
}

/* beamform.c:10-47: per-source min/max moveout over used stations. */
void oracle_delay_bounds(const int *tau, const float *beta, size_t n_sources,
                         size_t n_stations, size_t n_phases, int num_threads,
                         int *lohi)
{
#pragma omp parallel for num_threads(nthreads_or_default(num_threads)) schedule(static)
    for (size_t k = 0; k < n_sources; k++) {
        int lo = INT_MAX, hi = INT_MIN;
        const int *tk = tau + k * n_stations * n_phases;
        const float *bk = beta + k * n_stations;
        for (size_t s = 0; s < n_stations; s++) {
            if (bk[s] == 0.0f)
                continue;
            for (size_t p = 0; p < n_phases; p++) {
                int d = tk[s * n_phases + p];
                if (d > hi) hi = d;
                if (d < lo) lo = d;
            }
        }
        lohi[2 * k + 0] = lo;
        lohi[2 * k + 1] = hi;
    }
}

static int strict_valid(long long t, int lo, int hi, size_t n_samples)
{
    /* A source with no used station keeps the sentinels lo=INT_MAX,
     * hi=INT_MIN and is never valid: the reference's test t + hi >= T is
     * evaluated in size_t, where t + INT_MIN wraps to a huge value
     * (beamform.c:196-197, 279-280). */
    if (lo > hi)
        return 0;
    return (t + (long long)lo >= 0) && (t + (long long)hi < (long long)n_samples);
}

/* beamform.c:49-72 (strict: every gathered index is in range by construction) */
static float stack_strict(const float *F, const int *tk, const float *bk,
                          long long t, size_t n_samples, size_t n_stations,
                          size_t n_phases)
{
    float acc = 0.0f;
    for (size_t s = 0; s < n_stations; s++) {
        float w = bk[s];
        if (w == 0.0f)
            continue;
        const float *Fs = F + s * n_samples * n_phases;
        for (size_t p = 0; p < n_phases; p++) {
            long long idx = t + tk[s * n_phases + p];
            MAC32(acc, w, Fs[idx * (long long)n_phases + (long long)p]);
        }
    }
    return acc;
}

/* beamform.c:74-105 (flexible: drop terms with t+tau outside [0,T)) */
static float stack_flexible(const float *F, const int *tk, const float *bk,
                            long long t, size_t n_samples, size_t n_stations,
                            size_t n_phases)
{
    float acc = 0.0f;
    for (size_t s = 0; s < n_stations; s++) {
        float w = bk[s];
        if (w == 0.0f)
            continue;
        const float *Fs = F + s * n_samples * n_phases;
        for (size_t p = 0; p < n_phases; p++) {
            long long idx = t + tk[s * n_phases + p];
            if (idx < 0 || idx >= (long long)n_samples)
                continue;
            MAC32(acc, w, Fs[idx * (long long)n_phases + (long long)p]);
        }
    }
    return acc;
}

/* beamform.c:108-152: F[s,t,p] = sum_c alpha[s,c,p] * u[s,c,t], c ascending,
 * output layout (S,T,P). */
void oracle_prestack(const float *u, const float *alpha, size_t n_samples,
                     size_t n_stations, size_t n_channels, size_t n_phases,
                     int num_threads, float *F)
{
#pragma omp parallel for num_threads(nthreads_or_default(num_threads)) schedule(static)
    for (size_t s = 0; s < n_stations; s++) {
        for (size_t t = 0; t < n_samples; t++) {
            for (size_t p = 0; p < n_phases; p++) {
                float acc = 0.0f;
                for (size_t c = 0; c < n_channels; c++) {
                    float w = alpha[(s * n_channels + c) * n_phases + p];
                    float x = u[(s * n_channels + c) * n_samples + t];
                    MAC32(acc, w, x);
                }
                F[(s * n_samples + t) * n_phases + p] = acc;
            }
        }
    }
}

/* beamform.c:154-228: full (K,T) beams; OpenMP over sources. */
void oracle_beamform(const float *F, const int *tau, const float *beta,
                     size_t n_samples, size_t n_sources, size_t n_stations,
                     size_t n_phases, int out_of_bounds, int num_threads,
                     float *beam, int *lohi_scratch)
{
    int nt = nthreads_or_default(num_threads);
    oracle_delay_bounds(tau, beta, n_sources, n_stations, n_phases, nt, lohi_scratch);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t k = 0; k < n_sources; k++) {
        const int *tk = tau + k * n_stations * n_phases;
        const float *bk = beta + k * n_stations;
        int lo = lohi_scratch[2 * k], hi = lohi_scratch[2 * k + 1];
        for (long long t = 0; t < (long long)n_samples; t++) {
            if (out_of_bounds == 0) {
                if (!strict_valid(t, lo, hi, n_samples))
                    continue;
                beam[k * n_samples + t] =
                    stack_strict(F, tk, bk, t, n_samples, n_stations, n_phases);
            } else {
                beam[k * n_samples + t] =
                    stack_flexible(F, tk, bk, t, n_samples, n_stations, n_phases);
            }
        }
    }
}

/* beamform.c:230-336: per-t max over sources, lowest index on ties, for the
 * samples [t_begin, t_end) of the window (outputs indexed from t_begin).
 * Each sample's maximum depends only on the whole window's features and
 * bounds, so a sub-range gives exactly the full call's values there. */
void oracle_beamform_max_range(const float *F, const int *tau, const float *beta,
                               size_t n_samples, size_t n_sources, size_t n_stations,
                               size_t n_phases, int out_of_bounds, int num_threads,
                               long long t_begin, long long t_end, float *beam_max,
                               int *beam_argmax, int *lohi_scratch)
{
    int nt = nthreads_or_default(num_threads);
    oracle_delay_bounds(tau, beta, n_sources, n_stations, n_phases, nt, lohi_scratch);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long long t = t_begin; t < t_end; t++) {
        float best = -FLT_MAX;
        int best_k = 0;
        for (size_t k = 0; k < n_sources; k++) {
            const int *tk = tau + k * n_stations * n_phases;
            const float *bk = beta + k * n_stations;
            float b;
            if (out_of_bounds == 0) {
                if (!strict_valid(t, lohi_scratch[2 * k], lohi_scratch[2 * k + 1], n_samples))
                    continue;
                b = stack_strict(F, tk, bk, t, n_samples, n_stations, n_phases);
            } else {
                b = stack_flexible(F, tk, bk, t, n_samples, n_stations, n_phases);
            }
            if (b > best) {
                best = b;
                best_k = (int)k;
            }
        }
        if (best > -FLT_MAX)
            beam_max[t - t_begin] = best;
        beam_argmax[t - t_begin] = best_k;
    }
}

void oracle_beamform_max(const float *F, const int *tau, const float *beta,
                         size_t n_samples, size_t n_sources, size_t n_stations,
                         size_t n_phases, int out_of_bounds, int num_threads,
                         float *beam_max, int *beam_argmax, int *lohi_scratch)
{
    oracle_beamform_max_range(F, tau, beta, n_samples, n_sources, n_stations, n_phases,
                              out_of_bounds, num_threads, 0, (long long)n_samples, beam_max,
                              beam_argmax, lohi_scratch);
}

/* beamform.c:338-394 (mode='differential'): one beam per source at the
 * centre sample t = (T - 1) / 2, for sources whose used delays all point
 * inside [0, T); other entries of `beam` are left untouched.  The
 * reference shares its per-source bounds between OpenMP threads
 * (time_delay_min/max are declared outside the parallel loop and not
 * private), a data race with more than one thread; here they are private,
 * which is the reference's single-thread behaviour. */
void oracle_beamform_differential(const float *F, const int *tau, const float *beta,
                                  size_t n_samples, size_t n_sources, size_t n_stations,
                                  size_t n_phases, int num_threads, float *beam,
                                  int *lohi_scratch)
{
    int nt = nthreads_or_default(num_threads);
    if (n_samples == 0)
        return;
    long long t = ((long long)n_samples - 1) / 2;
    oracle_delay_bounds(tau, beta, n_sources, n_stations, n_phases, nt, lohi_scratch);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t k = 0; k < n_sources; k++) {
        if (!strict_valid(t, lohi_scratch[2 * k], lohi_scratch[2 * k + 1], n_samples))
            continue;
        beam[k] = stack_strict(F, tau + k * n_stations * n_phases, beta + k * n_stations,
                               t, n_samples, n_stations, n_phases);
    }
}

/* ---------------------------------------------------------------------- */
/* FP64 twins: same loops, double arithmetic, for the reduced-precision    */
/* feature forms (features stored narrow, summed in float on the device).  */
/* ---------------------------------------------------------------------- */

void oracle_prestack_f64(const double *u, const double *alpha, size_t n_samples,
                         size_t n_stations, size_t n_channels, size_t n_phases,
                         int num_threads, double *F)
{
#pragma omp parallel for num_threads(nthreads_or_default(num_threads)) schedule(static)
    for (size_t s = 0; s < n_stations; s++)
        for (size_t t = 0; t < n_samples; t++)
            for (size_t p = 0; p < n_phases; p++) {
                double acc = 0.0;
                for (size_t c = 0; c < n_channels; c++)
                    MAC64(acc, alpha[(s * n_channels + c) * n_phases + p],
                          u[(s * n_channels + c) * n_samples + t]);
                F[(s * n_samples + t) * n_phases + p] = acc;
            }
}

static double stack_f64(const double *F, const int *tk, const double *bk,
                        long long t, size_t n_samples, size_t n_stations,
                        size_t n_phases, int flexible)
{
    double acc = 0.0;
    for (size_t s = 0; s < n_stations; s++) {
        double w = bk[s];
        if (w == 0.0)
            continue;
        const double *Fs = F + s * n_samples * n_phases;
        for (size_t p = 0; p < n_phases; p++) {
            long long idx = t + tk[s * n_phases + p];
            if (flexible && (idx < 0 || idx >= (long long)n_samples))
                continue;
            MAC64(acc, w, Fs[idx * (long long)n_phases + (long long)p]);
        }
    }
    return acc;
}

static void delay_bounds_f64(const int *tau, const double *beta, size_t n_sources,
                             size_t n_stations, size_t n_phases, int nt, int *lohi)
{
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t k = 0; k < n_sources; k++) {
        int lo = INT_MAX, hi = INT_MIN;
        for (size_t s = 0; s < n_stations; s++) {
            if (beta[k * n_stations + s] == 0.0)
                continue;
            for (size_t p = 0; p < n_phases; p++) {
                int d = tau[(k * n_stations + s) * n_phases + p];
                if (d > hi) hi = d;
                if (d < lo) lo = d;
            }
        }
        lohi[2 * k] = lo;
        lohi[2 * k + 1] = hi;
    }
}

/* Full beams in double on a sub-range of time [t_begin, t_end); the
 * output row stride is (t_end - t_begin).  Used on short slices of long
 * windows: the slice is exact because beam(k,t) only reads
 * F[:, t+lo_k .. t+hi_k, :]. */
void oracle_beamform_f64(const double *F, const int *tau, const double *beta,
                         size_t n_samples, size_t n_sources, size_t n_stations,
                         size_t n_phases, int out_of_bounds, int num_threads,
                         long long t_begin, long long t_end, double *beam,
                         unsigned char *valid, int *lohi_scratch)
{
    int nt = nthreads_or_default(num_threads);
    long long width = t_end - t_begin;
    delay_bounds_f64(tau, beta, n_sources, n_stations, n_phases, nt, lohi_scratch);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t k = 0; k < n_sources; k++) {
        const int *tk = tau + k * n_stations * n_phases;
        const double *bk = beta + k * n_stations;
        for (long long t = t_begin; t < t_end; t++) {
            size_t o = k * (size_t)width + (size_t)(t - t_begin);
            int ok = out_of_bounds ? 1
                                   : strict_valid(t, lohi_scratch[2 * k],
                                                  lohi_scratch[2 * k + 1], n_samples);
            valid[o] = (unsigned char)ok;
            beam[o] = ok ? stack_f64(F, tk, bk, t, n_samples, n_stations, n_phases,
                                     out_of_bounds)
                         : 0.0;
        }
    }
}

/* Max/argmax in double over [t_begin, t_end) (same candidate and tie rules
 * as oracle_beamform_max), plus the runner-up value of each sample so that
 * callers can tell ties from clear winners.  out_valid[t] = 0 where no
 * source is a candidate. */
void oracle_beamform_max_f64(const double *F, const int *tau, const double *beta,
                             size_t n_samples, size_t n_sources, size_t n_stations,
                             size_t n_phases, int out_of_bounds, int num_threads,
                             long long t_begin, long long t_end, double *out_max,
                             double *out_second, int *out_arg, unsigned char *out_valid,
                             int *lohi_scratch)
{
    int nt = nthreads_or_default(num_threads);
    delay_bounds_f64(tau, beta, n_sources, n_stations, n_phases, nt, lohi_scratch);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4)
    for (long long t = t_begin; t < t_end; t++) {
        double best = -DBL_MAX, second = -DBL_MAX;
        int best_k = 0, any = 0;
        for (size_t k = 0; k < n_sources; k++) {
            if (!out_of_bounds &&
                !strict_valid(t, lohi_scratch[2 * k], lohi_scratch[2 * k + 1], n_samples))
                continue;
            double b = stack_f64(F, tau + k * n_stations * n_phases, beta + k * n_stations, t,
                                 n_samples, n_stations, n_phases, out_of_bounds);
            if (!(b > -FLT_MAX))
                continue;
            any = 1;
            if (b > best) {
                second = best;
                best = b;
                best_k = (int)k;
            } else if (b > second) {
                second = b;
            }
        }
        long long o = t - t_begin;
        out_max[o] = best;
        out_second[o] = second;
        out_arg[o] = best_k;
        out_valid[o] = (unsigned char)any;
    }
}
