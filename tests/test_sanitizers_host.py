"""The host code that builds without HIP, under the compilers' sanitizers.

* beampower_amd/csrc/bp_order.h (the similarity source order, bisected on
  host threads): a driver built with -fsanitize=thread and one with
  -fsanitize=address,undefined orders shuffled and degenerate delay tables
  on 1 to 16 threads and checks that every thread count gives the
  one-thread permutation.
* beampower_amd/csrc/bp_host_pool.h (the worker threads behind the
  library's host copies and comparisons): a driver built with each of the
  two sanitizer sets runs jobs from eight caller threads at once, nested
  jobs, jobs whose tasks throw, and pools of no and one worker, and checks
  that every task ran exactly once.
* oracle/beamform_oracle.c (OpenMP): a driver built with
  -fsanitize=address,undefined runs every oracle entry point (prestack,
  beams, maxima over the whole window and a sample range, differential mode,
  the FP64 twins) in both bounds modes, including negative delays, unused
  stations and sources, and empty shapes.  The OpenMP runtime (libgomp) is
  not built with ThreadSanitizer, which would then report its own
  synchronisation as races, so the oracle has no TSan run; its parallel
  loops write disjoint outputs.

The drivers run as child processes (a sanitized library cannot be loaded
into an uninstrumented Python); any sanitizer report fails the test.
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "beampower_amd", "csrc")
ORACLE_C = os.path.join(REPO, "oracle", "beamform_oracle.c")

ORDER_DRIVER = r'''
#include "bp_order.h"
#include <cstdio>
#include <cstdlib>
#include <vector>

static unsigned rng_state = 12345u;
static int rnd(int n) { rng_state = rng_state * 1103515245u + 12345u; return (rng_state >> 8) % n; }

int main() {
  // a grid-like table with shuffled rows and duplicated rows, and edge shapes
  const long K = 6000, D = 24;
  std::vector<int> tau(K * D);
  for (long k = 0; k < K; ++k)
    for (long d = 0; d < D; ++d) tau[k * D + d] = (int)((k * 7 + d * 13) % 300) + rnd(5);
  for (long k = 0; k < K; ++k) {  // shuffle rows
    long j = rnd((int)K);
    for (long d = 0; d < D; ++d) std::swap(tau[k * D + d], tau[j * D + d]);
  }
  for (int i = 0; i < 200; ++i) {
    long k = rnd((int)K);
    for (long d = 0; d < D; ++d) tau[k * D + d] = tau[d];
  }
  const std::vector<int> one = bp::similarity_order(tau.data(), K, D, 48, 1);
  std::vector<char> seen(K, 0);
  for (long i = 0; i < K; ++i) {
    if (one[i] < 0 || one[i] >= K || seen[one[i]]) { std::puts("not a permutation"); return 2; }
    seen[one[i]] = 1;
  }
  for (unsigned th : {2u, 3u, 8u, 16u}) {
    const std::vector<int> o = bp::similarity_order(tau.data(), K, D, 48, th);
    if (o != one) { std::printf("%u threads differ\n", th); return 3; }
  }
  const std::vector<int> lib = bp::similarity_order(tau.data(), K, D, 64);
  if ((long)lib.size() != K) return 4;
  const std::vector<int> small = bp::similarity_order(tau.data(), 1, D, 48, 4);
  const std::vector<int> none = bp::similarity_order(tau.data(), 100, 0, 48, 4);
  if (small.size() != 1 || none.size() != 100) return 5;
  std::puts("order ok");
  return 0;
}
'''

POOL_DRIVER = r'''
#include "bp_host_pool.h"
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

static int check(bp::HostPool &pool, size_t n, int rounds) {
  for (int r = 0; r < rounds; ++r) {
    std::vector<std::atomic<int>> hits(n);
    for (auto &h : hits) h.store(0);
    pool.run(n, [&](size_t i) { hits[i].fetch_add(1); });
    for (size_t i = 0; i < n; ++i)
      if (hits[i].load() != 1) return 1;
  }
  return 0;
}

int main() {
  bp::HostPool &pool = bp::HostPool::shared();
  int bad = check(pool, 0, 2) + check(pool, 1, 3) + check(pool, 37, 20);
  // eight callers at once, each with its own jobs
  std::vector<std::thread> callers;
  std::atomic<int> caller_bad{0};
  for (int c = 0; c < 8; ++c)
    callers.emplace_back([&, c] { caller_bad += check(pool, 5 + 3 * c, 30); });
  for (auto &t : callers) t.join();
  bad += caller_bad.load();
  // nested: tasks that run jobs of their own on the same pool
  std::atomic<long> total{0};
  pool.run(6, [&](size_t i) {
    pool.run(10, [&](size_t j) { total += (long)(i * 10 + j); });
  });
  if (total.load() != 1770) bad += 1;
  // a throwing task: the others still run, the first exception reaches the caller
  std::atomic<int> ran{0};
  bool caught = false;
  try {
    pool.run(12, [&](size_t i) {
      ran += 1;
      if (i % 5 == 3) throw std::runtime_error("task failed");
    });
  } catch (const std::runtime_error &) {
    caught = true;
  }
  if (!caught || ran.load() != 12) bad += 1;
  // pools of no and one worker
  bp::HostPool none(0), one(1);
  bad += check(none, 9, 3) + check(one, 9, 30);
  if (bad) { std::printf("pool failures: %d\n", bad); return 2; }
  std::puts("pool ok");
  return 0;
}
'''

ORACLE_DRIVER = r'''
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
void oracle_delay_bounds(const int *, const float *, size_t, size_t, size_t, int, int *);
void oracle_prestack(const float *, const float *, size_t, size_t, size_t, size_t, int, float *);
void oracle_beamform(const float *, const int *, const float *, size_t, size_t, size_t, size_t,
                     int, int, float *, int *);
void oracle_beamform_max(const float *, const int *, const float *, size_t, size_t, size_t,
                         size_t, int, int, float *, int *, int *);
void oracle_beamform_max_range(const float *, const int *, const float *, size_t, size_t,
                               size_t, size_t, int, int, long long, long long, float *, int *,
                               int *);
void oracle_beamform_differential(const float *, const int *, const float *, size_t, size_t,
                                  size_t, size_t, int, float *, int *);
void oracle_prestack_f64(const double *, const double *, size_t, size_t, size_t, size_t, int,
                         double *);
void oracle_beamform_f64(const double *, const int *, const double *, size_t, size_t, size_t,
                         size_t, int, int, long long, long long, double *, unsigned char *,
                         int *);
void oracle_beamform_max_f64(const double *, const int *, const double *, size_t, size_t,
                             size_t, size_t, int, int, long long, long long, double *, double *,
                             int *, unsigned char *, int *);

static unsigned st = 7u;
static float frand(void) { st = st * 1664525u + 1013904223u; return (float)(st >> 8) / 16777216.0f; }

static int run(size_t S, size_t C, size_t P, size_t T, size_t K, int lo, int hi) {
  float *u = malloc(sizeof(float) * (S * C * T + 1)), *a = malloc(sizeof(float) * (S * C * P + 1));
  float *F = malloc(sizeof(float) * (S * T * P + 1)), *b = malloc(sizeof(float) * (K * S + 1));
  int *tau = malloc(sizeof(int) * (K * S * P + 1)), *scr = malloc(sizeof(int) * (2 * K + 2));
  for (size_t i = 0; i < S * C * T; ++i) u[i] = frand();
  for (size_t i = 0; i < S * C * P; ++i) a[i] = frand() - 0.3f;
  for (size_t i = 0; i < K * S * P; ++i) tau[i] = lo + (int)(frand() * (float)(hi - lo));
  for (size_t i = 0; i < K * S; ++i) b[i] = frand() < 0.2f ? 0.0f : frand();
  if (K > 3) for (size_t s = 0; s < S; ++s) b[3 * S + s] = 0.0f;  /* a source with no station */
  oracle_prestack(u, a, T, S, C, P, 3, F);
  float *beam = calloc(K * T + 1, sizeof(float)), *m = calloc(T + 1, sizeof(float));
  int *am = calloc(T + 1, sizeof(int));
  for (int oob = 0; oob < 2; ++oob) {
    if (oob == 0 && lo < 0) continue;  /* the reference's strict max needs delays >= 0 */
    oracle_beamform(F, tau, b, T, K, S, P, oob, 3, beam, scr);
    oracle_beamform_max(F, tau, b, T, K, S, P, oob, 3, m, am, scr);
    if (T > 4) oracle_beamform_max_range(F, tau, b, T, K, S, P, oob, 3, 1, (long long)T - 2,
                                         m, am, scr);
  }
  float *dbeam = calloc(K + 1, sizeof(float));
  oracle_beamform_differential(F, tau, b, T, K, S, P, 3, dbeam, scr);
  free(dbeam);
  double *u64 = malloc(sizeof(double) * (S * C * T + 1)), *a64 = malloc(sizeof(double) * (S * C * P + 1));
  double *F64 = malloc(sizeof(double) * (S * T * P + 1)), *b64 = malloc(sizeof(double) * (K * S + 1));
  for (size_t i = 0; i < S * C * T; ++i) u64[i] = u[i];
  for (size_t i = 0; i < S * C * P; ++i) a64[i] = a[i];
  for (size_t i = 0; i < K * S; ++i) b64[i] = b[i];
  oracle_prestack_f64(u64, a64, T, S, C, P, 3, F64);
  double *beam64 = malloc(sizeof(double) * (K * T + 1)), *mx = malloc(sizeof(double) * (T + 1));
  double *sec = malloc(sizeof(double) * (T + 1));
  unsigned char *valid = malloc((K > 1 ? K : 1) * T + 1);  /* (K,T) beams, (T) maxima */
  for (int oob = 0; oob < 2; ++oob) {
    if (oob == 0 && lo < 0) continue;
    oracle_beamform_f64(F64, tau, b64, T, K, S, P, oob, 3, 0, (long long)T, beam64, valid, scr);
    oracle_beamform_max_f64(F64, tau, b64, T, K, S, P, oob, 3, 0, (long long)T, mx, sec, am,
                            valid, scr);
  }
  free(u); free(a); free(F); free(b); free(tau); free(scr); free(beam); free(m); free(am);
  free(u64); free(a64); free(F64); free(b64); free(beam64); free(mx); free(sec); free(valid);
  return 0;
}

int main(void) {
  run(5, 3, 2, 300, 70, 0, 60);     /* ordinary */
  run(4, 2, 3, 90, 33, -40, 40);    /* negative delays (flexible and differential) */
  run(3, 1, 1, 17, 5, 0, 30);       /* delays beyond the window */
  run(2, 3, 2, 0, 4, 0, 5);         /* no samples */
  run(2, 3, 2, 40, 0, 0, 5);        /* no sources */
  puts("oracle ok");
  return 0;
}
'''


def _build_and_run(tmp_path, name, source, compiler, flags, env_extra):
    src = tmp_path / name
    src.write_text(source)
    exe = tmp_path / (name + ".bin")
    subprocess.run([compiler, *flags, str(src), "-o", str(exe)], check=True,
                   capture_output=True, text=True)
    env = dict(os.environ, **env_extra)
    res = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    for marker in ("Sanitizer", "runtime error", "data race"):
        assert marker not in out, out[-4000:]
    return out


@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_similarity_order_under_sanitizers(tmp_path, san):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    flags = ["-O1", "-g", "-std=c++17", f"-fsanitize={san}", "-fno-omit-frame-pointer",
             "-pthread", "-I", CSRC]
    if "undefined" in san:
        flags.append("-fno-sanitize-recover=undefined")
    out = _build_and_run(tmp_path, "order_driver.cpp", ORDER_DRIVER, "g++", flags,
                         {"TSAN_OPTIONS": "halt_on_error=1", "ASAN_OPTIONS": "halt_on_error=1"})
    assert "order ok" in out


@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_host_pool_under_sanitizers(tmp_path, san):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    flags = ["-O1", "-g", "-std=c++17", f"-fsanitize={san}", "-fno-omit-frame-pointer",
             "-pthread", "-I", CSRC]
    if "undefined" in san:
        flags.append("-fno-sanitize-recover=undefined")
    out = _build_and_run(tmp_path, "pool_driver.cpp", POOL_DRIVER, "g++", flags,
                         {"TSAN_OPTIONS": "halt_on_error=1", "ASAN_OPTIONS": "halt_on_error=1"})
    assert "pool ok" in out


def test_oracle_under_address_and_undefined_sanitizers(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    drv = tmp_path / "oracle_driver.c"
    drv.write_text(ORACLE_DRIVER)
    exe = tmp_path / "oracle_driver.bin"
    for fma in ([], ["-DORACLE_FMA"]):
        subprocess.run(["gcc", "-O1", "-g", "-std=c99", "-fopenmp", "-ffp-contract=off", *fma,
                        "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                        "-fno-omit-frame-pointer", ORACLE_C, str(drv), "-o", str(exe), "-lm"],
                       check=True, capture_output=True, text=True)
        res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600,
                             env=dict(os.environ, ASAN_OPTIONS="halt_on_error=1"))
        out = res.stdout + res.stderr
        assert res.returncode == 0, out[-4000:]
        assert "oracle ok" in out and "runtime error" not in out and "Sanitizer" not in out
