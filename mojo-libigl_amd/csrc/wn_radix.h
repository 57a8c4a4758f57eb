// Stable LSD radix sort of 64-bit keys with a 64-bit payload (host), used by
// the edge matching of the stripifier and the boundary-edge extraction.
// 16-bit digits; a digit position on which all keys agree is skipped.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace wn {

// payload may be null (keys only).
inline void radix_sort(std::vector<uint64_t>& keys, std::vector<uint64_t>* payload) {
  const size_t n = keys.size();
  if (n < 2) return;
  uint64_t varies = 0;
  for (size_t i = 1; i < n; ++i) varies |= keys[i] ^ keys[0];
  std::vector<uint64_t> k2(n), p2(payload ? n : 0);
  std::vector<size_t> count(65537);
  for (int shift = 0; shift < 64; shift += 16) {
    if (((varies >> shift) & 0xFFFF) == 0) continue;
    std::fill(count.begin(), count.end(), 0);
    for (size_t i = 0; i < n; ++i) ++count[((keys[i] >> shift) & 0xFFFF) + 1];
    for (size_t d = 1; d < count.size(); ++d) count[d] += count[d - 1];
    for (size_t i = 0; i < n; ++i) {
      const size_t at = count[(keys[i] >> shift) & 0xFFFF]++;
      k2[at] = keys[i];
      if (payload) p2[at] = (*payload)[i];
    }
    keys.swap(k2);
    if (payload) payload->swap(p2);
  }
}

}  // namespace wn
