// Stable LSD radix sort of (key, index) pairs on the device, 8 bits per pass.
// Used to bin particles by tile, to sort staged particles, and for the
// explicit sort by reference cell key. Ranks within a block are computed in
// item order, so equal keys keep their input order (std::stable_sort).
#include "radix.hpp"

namespace fopis {

namespace {

constexpr int kRBlock = 256;
constexpr int kRItems = 16;                    // items per thread per block chunk
constexpr int kRChunk = kRBlock * kRItems;     // items per block

__global__ __launch_bounds__(kRBlock) void k_radix_count(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                         uint32_t* __restrict__ hist, int nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRChunk;
  for (int q = 0; q < kRItems; ++q) {
    const int64_t i = base + (int64_t)q * kRBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of a uint32 array into int64 offsets: per-block scan with
// block totals, then a scan of the totals, then add-back.
__global__ __launch_bounds__(kRBlock) void k_scan_block(const uint32_t* __restrict__ in, int64_t n,
                                                        int64_t* __restrict__ out, int64_t* __restrict__ totals) {
  __shared__ int64_t s[kRBlock];
  const int64_t per = kRItems;
  const int64_t base = (int64_t)blockIdx.x * kRChunk + (int64_t)threadIdx.x * per;
  int64_t sum = 0;
  for (int q = 0; q < per; ++q) {
    const int64_t i = base + q;
    if (i < n) sum += in[i];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < kRBlock; off <<= 1) {
    int64_t v = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t run = s[threadIdx.x] - sum;
  for (int q = 0; q < per; ++q) {
    const int64_t i = base + q;
    if (i < n) {
      out[i] = run;
      run += in[i];
    }
  }
  if (threadIdx.x == kRBlock - 1) totals[blockIdx.x] = s[kRBlock - 1];
}

__global__ __launch_bounds__(kRBlock) void k_scan_totals(int64_t* __restrict__ t, int nb) {
  // single block: sequential chunks per thread, then block scan
  __shared__ int64_t s[kRBlock];
  const int per = (nb + kRBlock - 1) / kRBlock;
  const int b0 = threadIdx.x * per;
  int64_t sum = 0;
  for (int q = 0; q < per; ++q)
    if (b0 + q < nb) sum += t[b0 + q];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < kRBlock; off <<= 1) {
    int64_t v = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t run = s[threadIdx.x] - sum;
  for (int q = 0; q < per; ++q)
    if (b0 + q < nb) {
      const int64_t v = t[b0 + q];
      t[b0 + q] = run;
      run += v;
    }
}

__global__ __launch_bounds__(kRBlock) void k_scan_add(int64_t* __restrict__ out, int64_t n,
                                                      const int64_t* __restrict__ totals) {
  const int64_t add = totals[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kRChunk + (int64_t)threadIdx.x * kRItems;
  for (int q = 0; q < kRItems; ++q) {
    const int64_t i = base + q;
    if (i < n) out[i] += add;
  }
}

__global__ __launch_bounds__(kRBlock) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin, int64_t n, int shift,
                                                           const int64_t* __restrict__ off, int nblocks,
                                                           uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  __shared__ uint32_t cnt[4][256];
  __shared__ uint32_t pre[4][256];
  __shared__ int64_t run[256];
  run[threadIdx.x] = off[(int64_t)threadIdx.x * nblocks + blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int64_t base = (int64_t)blockIdx.x * kRChunk;
  for (int q = 0; q < kRItems; ++q) {
    for (int ww = 0; ww < 4; ++ww) cnt[ww][threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = base + (int64_t)q * kRBlock + threadIdx.x;
    const bool act = i < n;
    const uint32_t k = act ? kin[i] : 0u;
    const uint32_t v = act ? vin[i] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    unsigned long long m = __ballot(act);
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const int rank = __popcll(m & lt);
    if (act && rank == 0) cnt[w][d] = (uint32_t)__popcll(m);
    __syncthreads();
    {
      const int dd = threadIdx.x;
      uint32_t p = 0;
      for (int ww = 0; ww < 4; ++ww) {
        pre[ww][dd] = p;
        p += cnt[ww][dd];
      }
      // run[] is advanced after everyone has read it below
      cnt[0][dd] = p;  // reuse as the block-chunk total of digit dd
    }
    __syncthreads();
    if (act) {
      const int64_t pos = run[d] + pre[w][d] + rank;
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    run[threadIdx.x] += cnt[0][threadIdx.x];
    __syncthreads();
  }
}

__global__ void k_iota(uint32_t* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

}  // namespace

void RadixSorter::sort(hipStream_t st, uint32_t* keys, uint32_t* vals, int64_t n, int bits, bool init_vals) {
  if (n <= 0) return;
  if (n > (int64_t)0xffffffffll) throw Error(FOPIS_E_ARG, "radix sort: more than 2^32 items");
  const int nblocks = div_up(n, kRChunk);
  k2.alloc(std::max<size_t>(k2.n, n));
  v2.alloc(std::max<size_t>(v2.n, n));
  const int64_t nh = (int64_t)256 * nblocks;
  hist.alloc(std::max<size_t>(hist.n, nh));
  offs.alloc(std::max<size_t>(offs.n, nh));
  const int nsb = div_up(nh, kRChunk);
  tot.alloc(std::max<size_t>(tot.n, nsb));
  if (nsb > kRBlock * 4096) throw Error(FOPIS_E_ARG, "radix sort: too many items");
  if (init_vals) k_iota<<<div_up(n, 256), 256, 0, st>>>(vals, n);
  uint32_t *ka = keys, *va = vals, *kb = k2.p, *vb = v2.p;
  int passes = (bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    k_radix_count<<<nblocks, kRBlock, 0, st>>>(ka, n, shift, hist.p, nblocks);
    k_scan_block<<<nsb, kRBlock, 0, st>>>(hist.p, nh, offs.p, tot.p);
    k_scan_totals<<<1, kRBlock, 0, st>>>(tot.p, nsb);
    k_scan_add<<<nsb, kRBlock, 0, st>>>(offs.p, nh, tot.p);
    k_radix_scatter<<<nblocks, kRBlock, 0, st>>>(ka, va, n, shift, offs.p, nblocks, kb, vb);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    FOPIS_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    FOPIS_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
  }
  FOPIS_HIP(hipGetLastError());
}

void exclusive_scan_u32(hipStream_t st, const uint32_t* in, int64_t n, int64_t* out, DBuf<int64_t>& tmp) {
  if (n <= 0) return;
  const int nsb = div_up(n, kRChunk);
  tmp.alloc(std::max<size_t>(tmp.n, nsb));
  k_scan_block<<<nsb, kRBlock, 0, st>>>(in, n, out, tmp.p);
  k_scan_totals<<<1, kRBlock, 0, st>>>(tmp.p, nsb);
  k_scan_add<<<nsb, kRBlock, 0, st>>>(out, n, tmp.p);
  FOPIS_HIP(hipGetLastError());
}

}  // namespace fopis
