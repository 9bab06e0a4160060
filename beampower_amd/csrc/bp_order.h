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
#include <numeric>
#include <utility>
#include <vector>

namespace bp {

inline std::vector<int> similarity_order(const int *tau, size_t K, size_t D, size_t tile) {
  std::vector<int> idx(K);
  std::iota(idx.begin(), idx.end(), 0);
  if (K <= tile || D == 0) return idx;
  std::vector<std::pair<size_t, size_t>> todo{{0, K}};
  std::vector<std::pair<int, int>> key;
  std::vector<int> lo(D), hi(D);
  // the split entry is chosen on at most this many evenly spaced sources of
  // a range (all of them once ranges are small)
  constexpr size_t kSample = 512;
  while (!todo.empty()) {
    const size_t b = todo.back().first, e = todo.back().second;
    todo.pop_back();
    const size_t n = e - b;
    if (n <= tile) {
      std::sort(idx.begin() + (std::ptrdiff_t)b, idx.begin() + (std::ptrdiff_t)e);
      continue;
    }
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
    todo.push_back({b + nl, e});
    todo.push_back({b, b + nl});
  }
  return idx;
}

}  // namespace bp
