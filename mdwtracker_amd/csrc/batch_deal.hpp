// Host-side dealing of launch form 2's work items to the persistent grid
// (no counterpart in the reference, which renders one camera per launch).
//
// The persistent grid's workgroup g takes list positions g, g + G, g + 2G, ...
// in order. After each call the list is re-dealt from the per-item costs the
// last frame measured: longest processing time first with a static
// assignment. Items go by decreasing cost, in rounds of G, each round's
// direction alternating (g, then G - 1 - g), so each workgroup's sum of costs
// is about the same.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace mdwt {

// Item i's cost from its wall-clock stamps ticks[2i] (start) and ticks[2i+1]
// (end); a pair out of order counts as 0.
inline unsigned long long item_cost(const unsigned long long* ticks, size_t i) {
  const unsigned long long a = ticks[2 * i], b = ticks[2 * i + 1];
  return b >= a ? b - a : 0ull;
}

// Positions 0 .. n-1 by decreasing cost, equal costs in position order (the
// order a stable comparison sort by decreasing cost gives). A stable
// least-significant-digit radix sort on ~cost over the span of bits in which
// the keys differ, 11 bits a pass: two passes for the costs of a frame
// under 20 ms (2^21 ticks at 100 MHz), O(n) each.
inline std::vector<uint32_t> order_by_decreasing_cost(const unsigned long long* ticks, size_t n) {
  constexpr int kDigit = 11;
  constexpr uint64_t kMask = (uint64_t(1) << kDigit) - 1;
  std::vector<uint32_t> pos(n), pos_tmp(n);
  std::vector<uint64_t> key(n), key_tmp(n);
  uint64_t differ = 0;
  for (size_t i = 0; i < n; ++i) {
    pos[i] = uint32_t(i);
    key[i] = ~uint64_t(item_cost(ticks, i));
    differ |= key[i] ^ key[0];
  }
  int lo = 0, hi = -1;  // lowest and highest bit in which two keys differ
  for (int b = 0; b < 64; ++b)
    if ((differ >> b) & 1u) {
      if (hi < 0) lo = b;
      hi = b;
    }
  std::vector<size_t> start(kMask + 2);
  for (int shift = lo; shift <= hi; shift += kDigit) {
    std::fill(start.begin(), start.end(), size_t(0));
    for (size_t i = 0; i < n; ++i) ++start[((key[i] >> shift) & kMask) + 1];
    for (uint64_t d = 0; d <= kMask; ++d) start[d + 1] += start[d];
    for (size_t i = 0; i < n; ++i) {
      const size_t at = start[(key[i] >> shift) & kMask]++;
      pos_tmp[at] = pos[i];
      key_tmp[at] = key[i];
    }
    pos.swap(pos_tmp);
    key.swap(key_tmp);
  }
  return pos;
}

// The dealt list's source positions: dealt[idx] = list[src[idx]] for a list
// of n items with the given stamps, dealt to G workgroups (G >= 1).
inline std::vector<uint32_t> deal_sources(const unsigned long long* ticks, size_t n, size_t G) {
  const std::vector<uint32_t> pos = order_by_decreasing_cost(ticks, n);
  std::vector<uint32_t> src(n);
  for (size_t idx = 0; idx < n; ++idx) {
    const size_t r = idx / G, g = idx % G;
    const size_t cnt = (n - r * G) < G ? (n - r * G) : G;
    src[idx] = pos[r * G + ((r & 1) ? cnt - 1 - g : g)];
  }
  return src;
}

}  // namespace mdwt
