// Stable LSD radix sort of 64-bit keys with a 64-bit payload (host), used by
// the edge matching of the stripifier and the boundary-edge extraction.
// 16-bit digits; a digit position on which all keys agree is skipped. Large
// arrays are split into contiguous chunks, one per thread: each pass counts
// the digits per chunk, places chunk t's keys of digit d after every key of a
// smaller digit and after chunk t' < t's keys of digit d, and scatters the
// chunks in parallel -- the same order as the serial pass, so the result does
// not depend on the thread count.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <utility>
#include <vector>

namespace wn {

// payload may be null (keys only).
inline void radix_sort(std::vector<uint64_t>& keys, std::vector<uint64_t>* payload) {
  const size_t n = keys.size();
  if (n < 2) return;
  constexpr size_t kBuckets = 65536;
  const int threads = n < (size_t)1 << 20
                          ? 1
                          : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  uint64_t varies = 0;
  for (size_t i = 1; i < n; ++i) varies |= keys[i] ^ keys[0];
  std::vector<uint64_t> k2(n), p2(payload ? n : 0);
  std::vector<size_t> count((size_t)threads * kBuckets);
  std::vector<size_t> cut((size_t)threads + 1);
  for (int t = 0; t <= threads; ++t) cut[(size_t)t] = n * (size_t)t / (size_t)threads;
  auto run = [threads](auto&& fn) {
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(fn, t);
    fn(0);
    for (auto& th : pool) th.join();
  };
  for (int shift = 0; shift < 64; shift += 16) {
    if (((varies >> shift) & 0xFFFF) == 0) continue;
    std::fill(count.begin(), count.end(), 0);
    run([&](int t) {
      size_t* c = count.data() + (size_t)t * kBuckets;
      for (size_t i = cut[(size_t)t]; i < cut[(size_t)t + 1]; ++i) ++c[(keys[i] >> shift) & 0xFFFF];
    });
    // exclusive offsets, digit-major then chunk: count[t][d] becomes the
    // first output slot of chunk t's keys with digit d
    size_t at = 0;
    for (size_t d = 0; d < kBuckets; ++d)
      for (int t = 0; t < threads; ++t) {
        size_t& c = count[(size_t)t * kBuckets + d];
        const size_t here = c;
        c = at;
        at += here;
      }
    run([&](int t) {
      size_t* c = count.data() + (size_t)t * kBuckets;
      for (size_t i = cut[(size_t)t]; i < cut[(size_t)t + 1]; ++i) {
        const size_t slot = c[(keys[i] >> shift) & 0xFFFF]++;
        k2[slot] = keys[i];
        if (payload) p2[slot] = (*payload)[i];
      }
    });
    keys.swap(k2);
    if (payload) payload->swap(p2);
  }
}

}  // namespace wn
