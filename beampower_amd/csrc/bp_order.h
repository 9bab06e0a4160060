// bp_order.h -- source order of a network plan (host code).
//
// The LDS kernel stages, per (tile, station, phase), one window of features
// that covers every delay of the tile's sources, so a tile's cost grows with
// the spread of its sources' delays (the window is the time tile plus the
// spread).  The reference takes sources in the caller's order
// (beamform.c:255-300 scans k ascending), and nothing obliges a caller to
// number neighbouring grid points consecutively: a subset of a grid, a list
// of sources sorted by some other key, or a shuffled grid puts sources from
// all over the volume into each tile of consecutive indices.
//
// similarity_order() returns a permutation (position -> original index)
// that groups sources of similar delay vectors into the same tile: a
// recursive bisection of the K delay vectors (dimension S*P), each range
// split at a whole number of tiles near its middle along the delay entry
// whose range over the range's sources is widest.  Every range starts at a
// multiple of the tile size and ends at one or at K, so every tile but the
// last is whole, and each leaf is one tile.  Ties in the split key are
// broken by the original index, so the order is a deterministic function
// of the delays.  Within a tile, positions are in ascending original index:
// the kernels scan a tile's sources in position order with a strict '>',
// which then keeps the lowest original index among equal values, as the
// reference's ascending scan does.
//
// The order only changes which sources share a tile.  Each source's beam
// is the same chain of FMAs in any order, and across tiles the kernels
// break max ties on the original index, so results do not depend on it.
#pragma once

#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstdlib>
#include <exception>
#include <numeric>
#include <system_error>
#include <thread>
#include <utility>
#include <vector>

namespace bp {

namespace order_detail {

// Bisection state of one thread: scratch for the split of one range.
struct Splitter {
  const int *tau;
  size_t D, tile;
  std::vector<std::pair<int, int>> key;
  std::vector<int> lo, hi;

  Splitter(const int *t, size_t d, size_t tl, size_t max_range)
      : tau(t), D(d), tile(tl), lo(d), hi(d) {
    key.reserve(max_range);
  }

  // Splits idx[b, e) in place at a whole number of tiles and returns the
  // left part's length, or sorts a leaf (one tile at most) and returns 0.
  size_t split(int *idx, size_t b, size_t e) {
    const size_t n = e - b;
    if (n <= tile) {
      std::sort(idx + b, idx + e);
      return 0;
    }
    // the split entry is chosen on at most this many evenly spaced sources
    // of a range (all of them once ranges are small)
    constexpr size_t kSample = 512;
    std::fill(lo.begin(), lo.end(), INT_MAX);
    std::fill(hi.begin(), hi.end(), INT_MIN);
    const size_t step = n > kSample ? n / kSample : 1;
    for (size_t i = b; i < e; i += step) {
      const int *row = tau + (size_t)idx[i] * D;
      for (size_t d = 0; d < D; ++d) {
        lo[d] = std::min(lo[d], row[d]);
        hi[d] = std::max(hi[d], row[d]);
      }
    }
    size_t dbest = 0;
    long long wbest = -1;
    for (size_t d = 0; d < D; ++d) {
      const long long w = (long long)hi[d] - lo[d];
      if (w > wbest) {
        wbest = w;
        dbest = d;
      }
    }
    // left part: a whole number of tiles near half, at least one, fewer than n
    size_t nl = (n / 2 + tile / 2) / tile * tile;
    nl = std::max(tile, std::min(nl, n - 1));
    key.resize(n);
    for (size_t i = 0; i < n; ++i) {
      const int k = idx[b + i];
      key[i] = {tau[(size_t)k * D + dbest], k};
    }
    std::nth_element(key.begin(), key.begin() + (std::ptrdiff_t)nl, key.end());
    for (size_t i = 0; i < n; ++i) idx[b + i] = key[i].second;
    return nl;
  }

  // Bisects idx[b, e) down to its leaves.
  void run(int *idx, size_t b, size_t e) {
    std::vector<std::pair<size_t, size_t>> todo{{b, e}};
    while (!todo.empty()) {
      const auto r = todo.back();
      todo.pop_back();
      const size_t nl = split(idx, r.first, r.second);
      if (nl == 0) continue;
      todo.push_back({r.first + nl, r.second});
      todo.push_back({r.first, r.first + nl});
    }
  }
};

// Host threads for an order over `entries` delay entries: one below about
// a million entries (a thread start costs more than it saves there), else
// the hardware's threads, at most OMP_NUM_THREADS when that is set, and at
// most 16.
inline unsigned order_threads(size_t entries) {
  if (entries < ((size_t)1 << 20)) return 1;
  unsigned n = std::thread::hardware_concurrency();
  if (const char *env = std::getenv("OMP_NUM_THREADS")) {
    const long v = std::strtol(env, nullptr, 10);
    if (v > 0) n = std::min<unsigned>(n, (unsigned)v);
  }
  return std::max(1u, std::min(n, 16u));
}

// Bisects idx[b, e) with the splitters parts[t0, t0 + threads): the range
// is split on this thread, and its halves go to two threads (this one and
// a new one) each with half of the splitters, until one splitter is left.
inline void bisect(std::vector<Splitter> &parts, int *idx, size_t b, size_t e, unsigned t0,
                   unsigned threads) {
  Splitter &sp = parts[t0];
  if (threads <= 1) {
    sp.run(idx, b, e);
    return;
  }
  const size_t nl = sp.split(idx, b, e);
  if (nl == 0) return;
  const unsigned lt = threads / 2;
  // an exception of the new thread is carried to this one and rethrown
  // after the join (none may leave a thread's function)
  std::exception_ptr right_error;
  std::thread right;
  try {
    right = std::thread([&parts, &right_error, idx, b, e, nl, t0, lt, threads] {
      try {
        bisect(parts, idx, b + nl, e, t0 + lt, threads - lt);
      } catch (...) {
        right_error = std::current_exception();
      }
    });
  } catch (const std::system_error &) {
    sp.run(idx, b + nl, e);  // no thread: the right half is bisected here
  }
  try {
    bisect(parts, idx, b, b + nl, t0, lt);
  } catch (...) {
    if (right.joinable()) right.join();
    throw;
  }
  if (right.joinable()) right.join();
  if (right_error) std::rethrow_exception(right_error);
}

}  // namespace order_detail

// Each range is split by a function of its own sources and their order
// only, and ranges are disjoint, so subtrees are bisected on `threads` host
// threads with the same result as on one thread.
inline std::vector<int> similarity_order(const int *tau, size_t K, size_t D, size_t tile,
                                         unsigned threads) {
  std::vector<int> idx(K);
  std::iota(idx.begin(), idx.end(), 0);
  if (K <= tile || D == 0) return idx;
  const unsigned nt = std::max(1u, std::min(threads, 64u));
  // the first range of every splitter but the first is at most half of K
  // and a tile
  std::vector<order_detail::Splitter> parts;
  parts.reserve(nt);
  for (unsigned t = 0; t < nt; ++t) parts.emplace_back(tau, D, tile, t ? K / 2 + tile : K);
  order_detail::bisect(parts, idx.data(), 0, K, 0, nt);
  return idx;
}

inline std::vector<int> similarity_order(const int *tau, size_t K, size_t D, size_t tile) {
  return similarity_order(tau, K, D, tile, order_detail::order_threads(K * D));
}

}  // namespace bp
