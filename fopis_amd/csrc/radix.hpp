// Stable device radix sort and scan helpers (radix.hip).
#pragma once
#include "fopis_internal.hpp"

namespace fopis {

// Reusable scratch for stable LSD radix sorts of uint32 keys with uint32
// payloads. sort() leaves keys ascending and vals permuted alongside; equal
// keys keep their input order. init_vals fills vals with 0..n-1 first.
struct RadixSorter {
  DBuf<uint32_t> k2, v2, hist;
  DBuf<int64_t> offs, tot;
  void sort(hipStream_t st, uint32_t* keys, uint32_t* vals, int64_t n, int bits, bool init_vals);
};

// Exclusive prefix sum of n uint32 values into int64 offsets.
void exclusive_scan_u32(hipStream_t st, const uint32_t* in, int64_t n, int64_t* out, DBuf<int64_t>& tmp);

}  // namespace fopis
