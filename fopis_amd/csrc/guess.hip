// Learned initial guess for the Poisson solve: a 3-D convolutional
// encoder-decoder that maps (charge density, conductor mask) to the potential,
// trained on (rho, phi) pairs that the multigrid solver itself produces. The
// multigrid solve always iterates from the guess to the same residual, so the
// network only changes where the iteration starts. The reference has no
// counterpart (its solver is SOR from the previous potential,
// RootGridField.cpp:41-103).
//
// Network (levels 0..3 at n, n/2, n/4, n/8 voxels per axis; 3x3x3 kernels,
// zero padding, BF16 weights and activations, FP32 accumulation):
//   e0a 8->16, e0b 16->16 = S0 | pool | e1a 16->32, e1b 32->32 = S1 | pool |
//   e2a 32->48, e2b 48->48 = S2 | pool | e3a 48->64, e3b 64->64 = B3 |
//   d2a up(B3) 64->48 (+S2), d2b 48->48 | d1a up 48->32 (+S1), d1b 32->32 |
//   d0a up 32->16 (+S0), d0b 16->16 | out 16->1 (linear).
// Every layer but `out` is relu(conv(x) + b [+ skip]); pool is a 2x2x2 mean,
// up is nearest-neighbour (read through the index map, never materialised).
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16: activations
// are channels-last [voxel][channel], a K step of 32 is four 8-channel chunks
// of (tap, channel) pairs, one chunk per 16-lane group, loaded as 16 bytes.
// The data-gradient is the same kernel with the taps mirrored and the channel
// roles swapped; the weight-gradient reduces over voxels with the voxels as
// the MFMA K dimension.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "fopis_internal.hpp"

namespace fopis {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __host__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ __host__ inline uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline bf16x8 load8(const uint16_t* p) {
  const u32x4 raw = *reinterpret_cast<const u32x4*>(p);
  return __builtin_bit_cast(bf16x8, raw);
}

__device__ inline bf16x8 zero8() {
  const u32x4 z = {0u, 0u, 0u, 0u};
  return __builtin_bit_cast(bf16x8, z);
}

// See poisson.hip: XCD x runs a contiguous run of the logical workgroup ids.
__device__ inline int xcd_block(int p, int n) {
  const int per = n >> 3, rem = n & 7, x = p & 7, slot = p >> 3;
  return x * per + min(x, rem) + slot;
}

// ---------------------------------------------------------------- conv forward

enum { EPI_BF16 = 0, EPI_F32 = 1, EPI_PHI = 2 };

struct ConvArgs {
  const uint16_t* x;     // input [voxel][cin] (at half resolution when UP)
  const uint16_t* w;     // packed weights [nqp][cout][8]
  const float* bias;     // [cout]
  const uint16_t* add;   // [voxel][cout] added before the activation, or null
  uint16_t* y;           // EPI_BF16: [voxel][cout]
  float* y32;            // EPI_F32: channel 0, [voxel]
  double* phi;           // EPI_PHI: glue-layout potential
  const double* scal;    // EPI_PHI: scal[0] = input scale sigma
  double phi_k;          // EPI_PHI: phi = out * sigma * phi_k
  int phi_add;           // EPI_PHI: add to phi instead of replacing it
  int64_t gsx, gsy, goff;
  int vn0, vn1, vn2;     // EPI_PHI: voxels that map to interior nodes
  int n0, n1, n2;        // output grid
  int cin, cout, nq, nqp;
  int relu;
};

// Output bricks of kBI x kBJ x 16 voxels per workgroup: the brick's input
// with a one-voxel halo (kBI+2, kBJ+2, 18) is staged once in LDS, channel rows
// padded by 8 elements so that the 16 lanes of a fragment read hit distinct
// banks; each of the 4 waves then owns kBI*kBJ/4 rows of 16 voxels (its M
// tiles) and reads its A fragments from LDS, B fragments (weights) from
// global memory through the caches.
constexpr int kBI = 4, kBJ = 4, kBK = 16;
constexpr int kHI = kBI + 2, kHJ = kBJ + 2, kHK = kBK + 2;
constexpr int kHalo = kHI * kHJ * kHK;
constexpr int kRowsPerWave = kBI * kBJ / 4;

inline size_t conv_lds_bytes(int cin) { return (size_t)kHalo * (cin + 8) * sizeof(uint16_t); }

template <int NT, int EPI, bool UP>
__global__ __launch_bounds__(256) void k_conv(ConvArgs A) {
  extern __shared__ __attribute__((aligned(16))) uint16_t xs[];
  constexpr int MT = kRowsPerWave;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, grp = lane >> 4;
  const int nb1 = (A.n1 + kBJ - 1) / kBJ, nb2 = (A.n2 + kBK - 1) / kBK;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int b2 = blk % nb2, b1 = (blk / nb2) % nb1, b0 = blk / (nb2 * nb1);
  const int i0 = b0 * kBI, j0 = b1 * kBJ, k0 = b2 * kBK;
  const int cpt = A.cin >> 3;
  const int ls = A.cin + 8;  // LDS elements per staged voxel
  const int in1 = UP ? (A.n1 >> 1) : A.n1, in2 = UP ? (A.n2 >> 1) : A.n2;

  // stage the input brick with its halo (zeros outside the grid)
  for (int e = threadIdx.x; e < kHalo * cpt; e += 256) {
    const int ch = e % cpt, hv = e / cpt;
    const int hk = hv % kHK, hj = (hv / kHK) % kHJ, hi = hv / (kHK * kHJ);
    int gi = i0 - 1 + hi, gj = j0 - 1 + hj, gk = k0 - 1 + hk;
    u32x4 v = {0u, 0u, 0u, 0u};
    if (gi >= 0 && gi < A.n0 && gj >= 0 && gj < A.n1 && gk >= 0 && gk < A.n2) {
      if (UP) {
        gi >>= 1;
        gj >>= 1;
        gk >>= 1;
      }
      v = *reinterpret_cast<const u32x4*>(A.x + (((size_t)gi * in1 + gj) * in2 + gk) * A.cin + ch * 8);
    }
    *reinterpret_cast<u32x4*>(xs + hv * ls + ch * 8) = v;
  }
  __syncthreads();

  f32x4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = (f32x4){0.f, 0.f, 0.f, 0.f};

  const int nsteps = A.nqp >> 2;
  for (int s = 0; s < nsteps; ++s) {
    const int q = 4 * s + grp;
    bf16x8 bfr[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) bfr[nt] = load8(A.w + ((size_t)q * A.cout + nt * 16 + row) * 8);
    // chunks past the last tap (K padding) read chunk 0 of tap 0 and are zeroed
    const bool qok = q < A.nq;
    const int tap = qok ? q / cpt : 0;
    const int c0 = qok ? (q - tap * cpt) * 8 : 0;
    const int di = tap / 9, dj = (tap / 3) % 3, dk = tap % 3;  // halo-relative offsets 0..2
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int r = wave * MT + mt;
      const int hv = ((r / kBJ + di) * kHJ + (r % kBJ + dj)) * kHK + row + dk;
      bf16x8 af = qok ? __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(xs + hv * ls + c0)) : zero8();
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr[nt], acc[mt][nt], 0, 0, 0);
    }
  }

  // D layout: lane holds output channel nt*16 + row for voxels k0 + 4*grp + r of the row
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int rr = wave * MT + mt;
    const int i = i0 + rr / kBJ, j = j0 + rr % kBJ;
    if (i >= A.n0 || j >= A.n1) continue;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int co = nt * 16 + row;
      if (EPI != EPI_BF16 && co != 0) continue;
      const float bco = A.bias[co];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = k0 + 4 * grp + r;
        if (k >= A.n2) continue;
        const int64_t v = ((int64_t)i * A.n1 + j) * A.n2 + k;
        float val = acc[mt][nt][r] + bco;
        if (EPI == EPI_BF16) {
          if (A.add) val += bf2f(A.add[v * A.cout + co]);
          if (A.relu) val = fmaxf(val, 0.f);
          A.y[v * A.cout + co] = f2bf(val);
        } else if (EPI == EPI_F32) {
          A.y32[v] = val;
        } else if (i < A.vn0 && j < A.vn1 && k < A.vn2) {
          double* p = A.phi + A.goff + i * A.gsx + j * A.gsy + k;
          *p = (A.phi_add ? *p : 0.0) + (double)val * A.scal[0] * A.phi_k;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- weight gradient

struct WgradArgs {
  const uint16_t* x;  // layer input [voxel][cin] (half resolution when UP)
  const uint16_t* g;  // output gradient after the activation [voxel][cout]
  float* part;        // [pair][chunk][27][16][16]
  int n0, n1, n2;     // output grid
  int cin, cout, nco;  // nco = cout / 16 tiles
  int64_t nsteps;      // 32-voxel K steps
};

constexpr int kWTile = 27 * 256;

template <bool UP>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs A) {
  __shared__ float red[kWTile];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, grp = lane >> 4;
  const int pair = blockIdx.y;
  const int cit = pair / A.nco, cot = pair % A.nco;
  const int ci = cit * 16 + row, co = cot * 16 + row;
  const bool ciok = ci < A.cin;
  const int64_t nvox = (int64_t)A.n0 * A.n1 * A.n2;
  const int in1 = UP ? (A.n1 >> 1) : A.n1, in2 = UP ? (A.n2 >> 1) : A.n2;
  f32x4 acc[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};

  for (int64_t s = (int64_t)blockIdx.x * 4 + wave; s < A.nsteps; s += (int64_t)gridDim.x * 4) {
    const int64_t vb = s * 32 + grp * 8;
    int vi[8], vj[8], vk[8];
    bool vin[8];
    uint16_t gb[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int64_t v = vb + j;
      vin[j] = v < nvox;
      vk[j] = (int)(v % A.n2);
      vj[j] = (int)((v / A.n2) % A.n1);
      vi[j] = (int)(v / ((int64_t)A.n1 * A.n2));
      gb[j] = vin[j] ? A.g[v * A.cout + co] : (uint16_t)0;
    }
    u32x4 braw;
#pragma unroll
    for (int h = 0; h < 4; ++h) braw[h] = (uint32_t)gb[2 * h] | ((uint32_t)gb[2 * h + 1] << 16);
    const bf16x8 bfr = __builtin_bit_cast(bf16x8, braw);
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
      uint16_t xa[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        int ii = vi[j] + di, jj = vj[j] + dj, kk = vk[j] + dk;
        const bool ok = ciok && vin[j] && ii >= 0 && ii < A.n0 && jj >= 0 && jj < A.n1 && kk >= 0 && kk < A.n2;
        uint16_t xv = 0;
        if (ok) {
          if (UP) {
            ii >>= 1;
            jj >>= 1;
            kk >>= 1;
          }
          xv = A.x[(((size_t)ii * in1 + jj) * in2 + kk) * A.cin + ci];
        }
        xa[j] = xv;
      }
      u32x4 araw;
#pragma unroll
      for (int h = 0; h < 4; ++h) araw[h] = (uint32_t)xa[2 * h] | ((uint32_t)xa[2 * h + 1] << 16);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, araw), bfr, acc[t], 0, 0, 0);
    }
  }
  // D[m = ci local][n = co local]: lane holds n = row, m = 4*grp + r. Sum the four waves.
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int t = 0; t < 27; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int idx = t * 256 + (4 * grp + r) * 16 + row;
          red[idx] = (w == 0 ? 0.f : red[idx]) + acc[t][r];
        }
    }
    __syncthreads();
  }
  float* out = A.part + ((size_t)pair * gridDim.x + blockIdx.x) * kWTile;
  for (int i = threadIdx.x; i < kWTile; i += 256) out[i] = red[i];
}

// dW[tap][ci][co] = sum over chunks (fixed order).
__global__ void k_wgrad_reduce(const float* part, int nchunk, int cin, int cout, int nco, float* dw) {
  const int64_t n = 27LL * cin * cout;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int co = (int)(e % cout), ci = (int)((e / cout) % cin), t = (int)(e / ((int64_t)cout * cin));
  const int pair = (ci / 16) * nco + co / 16;
  const float* p = part + (size_t)pair * nchunk * kWTile + t * 256 + (ci % 16) * 16 + (co % 16);
  float s = 0.f;
  for (int k = 0; k < nchunk; ++k) s += p[(size_t)k * kWTile];
  dw[e] = s;
}

// Column sums of g [nvox][cout] -> per-block partials.
__global__ __launch_bounds__(256) void k_colsum(const uint16_t* g, int64_t nvox, int cout, float* part) {
  __shared__ float s[256];
  const int lanes = 256 / cout;  // cout in {16, 32, 48, 64}: 48 leaves 16 idle threads
  const int co = threadIdx.x % cout, vl = threadIdx.x / cout;
  float acc = 0.f;
  if (vl < lanes)
    for (int64_t v = (int64_t)blockIdx.x * lanes + vl; v < nvox; v += (int64_t)gridDim.x * lanes)
      acc += bf2f(g[v * cout + co]);
  s[threadIdx.x] = vl < lanes ? acc : 0.f;
  __syncthreads();
  if (threadIdx.x < cout) {
    float t = 0.f;
    for (int l = 0; l < lanes; ++l) t += s[l * cout + threadIdx.x];
    part[(size_t)blockIdx.x * cout + threadIdx.x] = t;
  }
}

__global__ void k_colsum_reduce(const float* part, int nb, int cout, float* db) {
  const int co = threadIdx.x;
  if (co >= cout) return;
  float t = 0.f;
  for (int b = 0; b < nb; ++b) t += part[(size_t)b * cout + co];
  db[co] = t;
}

// ---------------------------------------------------------------- elementwise

__global__ void k_relu_bwd(uint16_t* g, const uint16_t* y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(bf2f(y[i]) > 0.f)) g[i] = 0;
}

// 2x2x2 mean: coarse [nc][C] from fine [2nc][C]; one thread per 8 channels.
__global__ void k_pool(const uint16_t* xf, uint16_t* xc, int c0, int c1, int c2, int C) {
  const int cc = C >> 3;
  const int64_t n = (int64_t)c0 * c1 * c2 * cc;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int h = (int)(e % cc);
  const int64_t v = e / cc;
  const int k = (int)(v % c2), j = (int)((v / c2) % c1), i = (int)(v / ((int64_t)c1 * c2));
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const size_t f = ((size_t)(2 * i + a) * (2 * c1) + (2 * j + b)) * (2 * c2) + (2 * k + c);
        const uint16_t* p = xf + f * C + h * 8;
        for (int q = 0; q < 8; ++q) s[q] += bf2f(p[q]);
      }
  uint16_t* o = xc + v * C + h * 8;
  for (int q = 0; q < 8; ++q) o[q] = f2bf(s[q] * 0.125f);
}

// Gradient through the mean pool, plus the skip-connection gradient:
// gf[fine] = 0.125 * gc[fine / 2] + skip[fine].
__global__ void k_pool_bwd(const uint16_t* gc, const uint16_t* skip, uint16_t* gf, int f0, int f1, int f2, int C) {
  const int64_t n = (int64_t)f0 * f1 * f2 * C;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int ch = (int)(e % C);
  const int64_t v = e / C;
  const int k = (int)(v % f2), j = (int)((v / f2) % f1), i = (int)(v / ((int64_t)f1 * f2));
  const size_t vc = ((size_t)(i >> 1) * (f1 >> 1) + (j >> 1)) * (f2 >> 1) + (k >> 1);
  gf[e] = f2bf(0.125f * bf2f(gc[vc * C + ch]) + bf2f(skip[e]));
}

// Gradient through nearest upsampling: coarse = sum of its 8 fine children.
__global__ void k_upsum(const uint16_t* gf, uint16_t* gc, int c0, int c1, int c2, int C) {
  const int64_t n = (int64_t)c0 * c1 * c2 * C;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int ch = (int)(e % C);
  const int64_t v = e / C;
  const int k = (int)(v % c2), j = (int)((v / c2) % c1), i = (int)(v / ((int64_t)c1 * c2));
  float s = 0.f;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const size_t f = ((size_t)(2 * i + a) * (2 * c1) + (2 * j + b)) * (2 * c2) + (2 * k + c);
        s += bf2f(gf[f * C + ch]);
      }
  gc[e] = f2bf(s);
}

// Weight packing. Forward: wp[q][co][j] = W[tap][c0 + j][co] with q = (tap, c0/8)
// (zero for q >= nq). Data-gradient (taps mirrored, channel roles swapped):
// wd[q'][ci][j] = W[26 - tap'][ci][o0 + j] with q' = (tap', o0/8).
__global__ void k_pack(const float* W, uint16_t* wp, int cin, int cout, int nqp, int mirror) {
  const int ci_in = mirror ? cout : cin;   // channels of the GEMM's K side
  const int co_out = mirror ? cin : cout;  // output channels
  const int cpt = ci_in >> 3;
  const int64_t n = (int64_t)nqp * co_out * 8;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int j = (int)(e % 8), o = (int)((e / 8) % co_out), q = (int)(e / (8 * (int64_t)co_out));
  float v = 0.f;
  if (q < 27 * cpt) {
    const int tap = q / cpt, c = (q % cpt) * 8 + j;
    if (!mirror)
      v = W[((size_t)tap * cin + c) * cout + o];
    else
      v = W[((size_t)(26 - tap) * cin + o) * cout + c];
  }
  wp[e] = f2bf(v);
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------- inputs, targets, loss

struct Geo {
  int n0, n1, n2;       // network grid
  int vn0, vn1, vn2;    // voxels mapped to interior nodes
  int64_t gsx, gsy;     // glue strides
  int64_t goff;         // glue index of voxel (0,0,0): node (2,2,2)
  int mcy, mcz;         // cell-mask strides: (ny+1), (nz+1)
};

// Sum of rho^2 over interior nodes (per-block partials).
__global__ __launch_bounds__(256) void k_rho2(const double* rho, Geo G, double* part) {
  __shared__ double s[256];
  const int64_t n = (int64_t)G.vn0 * G.vn1 * G.vn2;
  double acc = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (int64_t)gridDim.x * 256) {
    const int k = (int)(q % G.vn2), j = (int)((q / G.vn2) % G.vn1), i = (int)(q / ((int64_t)G.vn1 * G.vn2));
    const double r = rho[G.goff + i * G.gsx + j * G.gsy + k];
    acc += r * r;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// scal[0] = rms of rho over interior nodes (1 when rho vanishes), scal[1] = 1/scal[0].
// One block of 256 threads, fixed-order tree.
__global__ __launch_bounds__(256) void k_rho_scale(const double* part, int nb, int64_t n, double* scal) {
  __shared__ double sh[256];
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) t += part[b];
  sh[threadIdx.x] = t;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double s = sqrt(sh[0] / (double)n);
    if (!(s > 0.0)) s = 1.0;
    scal[0] = s;
    scal[1] = 1.0 / s;
  }
}

// Node mask: 1 if any of the 8 cells around the node belongs to a conductor.
__device__ inline uint8_t node_mask(const uint8_t* cm, const Geo& G, int i, int j, int k) {
  if (!cm) return 0;
  // voxel (i,j,k) is glue node (i+2, j+2, k+2); its cells are glue cells i+1..i+2
  uint8_t m = 0;
  for (int a = 1; a <= 2; ++a)
    for (int b = 1; b <= 2; ++b)
      for (int c = 1; c <= 2; ++c) m |= cm[((size_t)(i + a) * G.mcy + (j + b)) * G.mcz + (k + c)];
  return m ? 1 : 0;
}

// Training pair from the context fields: rho_hat = rho/sigma (BF16), node
// mask, phi_hat = phi * eps0 / sigma (FP32 in, BF16 stored).
__global__ void k_store_pair(const double* rho, const double* phi, const uint8_t* cm, Geo G, const double* scal,
                             double eps0, uint16_t* drho, uint8_t* dmask, uint16_t* dphi) {
  const int64_t n = (int64_t)G.n0 * G.n1 * G.n2;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int k = (int)(v % G.n2), j = (int)((v / G.n2) % G.n1), i = (int)(v / ((int64_t)G.n1 * G.n2));
  float r = 0.f, p = 0.f;
  uint8_t m = 0;
  if (i < G.vn0 && j < G.vn1 && k < G.vn2) {
    const int64_t g = G.goff + i * G.gsx + j * G.gsy + k;
    r = (float)(rho[g] * scal[1]);
    p = (float)(phi[g] * eps0 * scal[1]);
    m = node_mask(cm, G, i, j, k);
  }
  drho[v] = f2bf(r);
  dmask[v] = m;
  dphi[v] = f2bf(p);
}

// Network input [voxel][8]: channel 0 rho_hat, channel 1 mask, 2..7 zero.
__global__ void k_features_pair(const uint16_t* drho, const uint8_t* dmask, uint16_t* x0, int64_t n) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  u32x4 o = {(uint32_t)drho[v] | ((uint32_t)(dmask[v] ? 0x3f80u : 0u) << 16), 0u, 0u, 0u};
  *reinterpret_cast<u32x4*>(x0 + v * 8) = o;
}

__global__ void k_features_field(const double* rho, const uint8_t* cm, Geo G, const double* scal, uint16_t* x0) {
  const int64_t n = (int64_t)G.n0 * G.n1 * G.n2;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int k = (int)(v % G.n2), j = (int)((v / G.n2) % G.n1), i = (int)(v / ((int64_t)G.n1 * G.n2));
  uint16_t r = 0;
  uint8_t m = 0;
  if (i < G.vn0 && j < G.vn1 && k < G.vn2) {
    r = f2bf((float)(rho[G.goff + i * G.gsx + j * G.gsy + k] * scal[1]));
    m = node_mask(cm, G, i, j, k);
  }
  u32x4 o = {(uint32_t)r | ((uint32_t)(m ? 0x3f80u : 0u) << 16), 0u, 0u, 0u};
  *reinterpret_cast<u32x4*>(x0 + v * 8) = o;
}

// eps0 (b - A phi) at the network's interior nodes, b = rho / eps0 and A the
// 7-point operator (6 phi minus the neighbours, dx = 1): the charge whose
// potential the warm start phi is still missing.
__global__ void k_warm_resid(const double* phi, const double* rho, double eps0, Geo G, double* out) {
  const int64_t n = (int64_t)G.vn0 * G.vn1 * G.vn2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q % G.vn2), j = (int)((q / G.vn2) % G.vn1), i = (int)(q / ((int64_t)G.vn1 * G.vn2));
  const int64_t g = G.goff + i * G.gsx + j * G.gsy + k;
  const double ap = 6.0 * phi[g] - (phi[g - G.gsx] + phi[g + G.gsx] + phi[g - G.gsy] + phi[g + G.gsy] + phi[g - 1] +
                                    phi[g + 1]);
  out[g] = rho[g] - eps0 * ap;
}

// d = phi - phi_start at the interior nodes (in place into phi_start).
__global__ void k_phi_delta(const double* phi, double* d, Geo G) {
  const int64_t n = (int64_t)G.vn0 * G.vn1 * G.vn2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q % G.vn2), j = (int)((q / G.vn2) % G.vn1), i = (int)(q / ((int64_t)G.vn1 * G.vn2));
  const int64_t g = G.goff + i * G.gsx + j * G.gsy + k;
  d[g] = phi[g] - d[g];
}

// Loss 0.5/N * sum (out - target/k)^2 over voxels of interior nodes, N their count;
// gradient into channel 0 of g [voxel][16].
__global__ __launch_bounds__(256) void k_loss(const float* out, const uint16_t* target, float inv_k, Geo G,
                                              float inv_n, uint16_t* g, double* part) {
  __shared__ double s[256];
  const int64_t n = (int64_t)G.n0 * G.n1 * G.n2;
  double acc = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (int64_t)gridDim.x * 256) {
    const int k = (int)(v % G.n2), j = (int)((v / G.n2) % G.n1), i = (int)(v / ((int64_t)G.n1 * G.n2));
    float d = 0.f;
    if (i < G.vn0 && j < G.vn1 && k < G.vn2) d = out[v] - bf2f(target[v]) * inv_k;
    acc += 0.5 * (double)d * d;
    u32x4 z = {(uint32_t)f2bf(d * inv_n), 0u, 0u, 0u};
    *reinterpret_cast<u32x4*>(g + v * 16) = z;
    *reinterpret_cast<u32x4*>(g + v * 16 + 8) = (u32x4){0u, 0u, 0u, 0u};
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Residual loss 0.5/N * sum q^2 with q = A (k out) - rho_hat over voxels of
// interior nodes (A the 7-point operator 6 u - sum of neighbours, the output
// taken as 0 outside those voxels): the initial relative residual the
// multigrid starts from. First pass: q and the loss partials.
__global__ __launch_bounds__(256) void k_res_loss(const float* out, const uint16_t* rho_hat, float k, Geo G,
                                                  float* q, double* part) {
  __shared__ double s[256];
  const int64_t n = (int64_t)G.n0 * G.n1 * G.n2;
  double acc = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (int64_t)gridDim.x * 256) {
    const int kk = (int)(v % G.n2), j = (int)((v / G.n2) % G.n1), i = (int)(v / ((int64_t)G.n1 * G.n2));
    float qv = 0.f;
    if (i < G.vn0 && j < G.vn1 && kk < G.vn2) {
      auto o = [&](int a, int b, int c) -> float {
        if (a < 0 || b < 0 || c < 0 || a >= G.vn0 || b >= G.vn1 || c >= G.vn2) return 0.f;
        return out[((int64_t)a * G.n1 + b) * G.n2 + c];
      };
      const float au = 6.f * out[v] - (o(i - 1, j, kk) + o(i + 1, j, kk) + o(i, j - 1, kk) + o(i, j + 1, kk) +
                                       o(i, j, kk - 1) + o(i, j, kk + 1));
      qv = k * au - bf2f(rho_hat[v]);
    }
    q[v] = qv;
    acc += 0.5 * (double)qv * qv;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o2 = 128; o2 > 0; o2 >>= 1) {
    if (threadIdx.x < o2) s[threadIdx.x] += s[threadIdx.x + o2];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Second pass: d loss / d out = (k/N) A q (A is symmetric), into channel 0 of g [voxel][16].
__global__ void k_res_grad(const float* q, float k_over_n, Geo G, uint16_t* g) {
  const int64_t n = (int64_t)G.n0 * G.n1 * G.n2;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int kk = (int)(v % G.n2), j = (int)((v / G.n2) % G.n1), i = (int)(v / ((int64_t)G.n1 * G.n2));
  float gv = 0.f;
  if (i < G.vn0 && j < G.vn1 && kk < G.vn2) {
    auto Q = [&](int a, int b, int c) -> float {
      if (a < 0 || b < 0 || c < 0 || a >= G.vn0 || b >= G.vn1 || c >= G.vn2) return 0.f;
      return q[((int64_t)a * G.n1 + b) * G.n2 + c];
    };
    gv = k_over_n * (6.f * q[v] - (Q(i - 1, j, kk) + Q(i + 1, j, kk) + Q(i, j - 1, kk) + Q(i, j + 1, kk) +
                                   Q(i, j, kk - 1) + Q(i, j, kk + 1)));
  }
  u32x4 z = {(uint32_t)f2bf(gv), 0u, 0u, 0u};
  *reinterpret_cast<u32x4*>(g + v * 16) = z;
  *reinterpret_cast<u32x4*>(g + v * 16 + 8) = (u32x4){0u, 0u, 0u, 0u};
}

__global__ void k_sum_into(const double* part, int nb, double scale, double* out) {
  if (threadIdx.x != 0) return;
  double t = 0.0;
  for (int b = 0; b < nb; ++b) t += part[b];
  *out = t * scale;
}

// Training-data objects: axis-aligned boxes of cells. Cells inside get the
// box id in the cell mask; rho at nodes with all 8 cells inside is cleared,
// rho at surface nodes (some but not all cells inside) gets the box's surface
// charge times a smooth modulation 1 + 0.5 sin(...) with per-box phases.
struct Box {
  int lo[3], hi[3];  // glue cell ranges [lo, hi)
  double q;          // surface charge per node
  double ph[3];      // modulation phases
};

__global__ void k_box_mask(uint8_t* cm, int c0, int c1, int c2, const Box* boxes, int nbox) {
  const int64_t n = (int64_t)c0 * c1 * c2;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int k = (int)(e % c2), j = (int)((e / c2) % c1), i = (int)(e / ((int64_t)c1 * c2));
  uint8_t m = 0;
  for (int b = 0; b < nbox; ++b) {
    const Box& B = boxes[b];
    if (i >= B.lo[0] && i < B.hi[0] && j >= B.lo[1] && j < B.hi[1] && k >= B.lo[2] && k < B.hi[2]) m = (uint8_t)(b + 1);
  }
  cm[e] = m;
}

__global__ void k_box_charge(double* rho, int cx, int cy, int cz, const uint8_t* cm, const Box* boxes, int nbox) {
  const int64_t n = (int64_t)cx * cy * cz;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int k = (int)(e % cz), j = (int)((e / cz) % cy), i = (int)(e / ((int64_t)cy * cz));
  if (i < 2 || j < 2 || k < 2 || i > cx - 3 || j > cy - 3 || k > cz - 3) return;  // interior nodes only
  int inside = 0, owner = 0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        // the cell map is (cx-1) x (cy-1) x (cz-1): cell (p, q, r) spans nodes p..p+1
        const uint8_t m = cm[((size_t)(i - 1 + a) * (cy - 1) + (j - 1 + b)) * (cz - 1) + (k - 1 + c)];
        if (m) {
          ++inside;
          owner = m;
        }
      }
  if (inside == 0) return;
  if (inside == 8) {
    rho[e] = 0.0;
    return;
  }
  const Box& B = boxes[owner - 1];
  const double mod = 1.0 + 0.5 * sin(0.37 * i + B.ph[0]) * sin(0.41 * j + B.ph[1]) * sin(0.43 * k + B.ph[2]);
  rho[e] += B.q * mod;
}

}  // namespace

// ---------------------------------------------------------------- host side

struct GuessLayer {
  int cin, cout, lvl, up, relu;
  int64_t woff, boff;  // offsets in the flat parameter vector
  int nq, nqp, nqd, nqpd;
  DBuf<uint16_t> wp, wd;
};

struct GuessNet {
  Geo G{};
  int lv[4][3];
  std::vector<GuessLayer> L;
  int64_t nparams = 0;
  DBuf<float> master, grad, m1, m2;
  double k_out = 1.0;
  int64_t adam_t = 0;
  // activations (BF16 channels-last) and gradients
  DBuf<uint16_t> X0, A0, S0, P0, A1, S1, P1, A2, S2, P2, A3, B3, D2, E2, D1, E1, D0, E0;
  DBuf<float> OUT;
  DBuf<uint16_t> gOUT, gE0, gD0, gE1, gD1, gE2, gD2, gB3, gA3, gP2, gS2, gA2, gP1, gS1, gA1, gP0, gS0, gA0;
  DBuf<uint16_t> T0, T1, T2;
  bool train_alloc = false;
  DBuf<float> wpart, cpart;  // weight- and bias-gradient partial sums
  DBuf<double> dpart, scal, loss_acc;
  DBuf<uint8_t> cellmask;  // training-data objects
  DBuf<double> rwarm;      // eps0 (b - A phi) of the warm start at the interior nodes (modes 2 and 3)
  DBuf<double> phi_start;  // mode 3: the warm start, to form phi_solution - phi_start after the solve
  int record_slot = 0;     // mode 3: the dataset slot the next pair goes to
  bool use_cellmask = false;
  DBuf<Box> boxes;
  // dataset of training pairs
  int capacity = 0;
  DBuf<uint16_t> ds_rho, ds_phi;
  DBuf<uint8_t> ds_mask;
  std::vector<double> ds_phi_rms;
  int mode = 0;
  int loss_kind = 0;  // 0: potential MSE, 1: initial multigrid residual
  DBuf<float> q;      // residual of the guess (loss_kind 1)
  uint64_t seed = 1;
};

namespace {

int64_t nvox_of(const int* d) { return (int64_t)d[0] * d[1] * d[2]; }

void add_layer(GuessNet& g, int cin, int cout, int lvl, int up, int relu) {
  GuessLayer l;
  l.cin = cin;
  l.cout = cout;
  l.lvl = lvl;
  l.up = up;
  l.relu = relu;
  l.woff = g.nparams;
  g.nparams += 27LL * cin * cout;
  l.boff = g.nparams;
  g.nparams += cout;
  l.nq = 27 * cin / 8;
  l.nqp = (l.nq + 3) / 4 * 4;
  l.nqd = 27 * cout / 8;
  l.nqpd = (l.nqd + 3) / 4 * 4;
  g.L.push_back(std::move(l));
}

enum { E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, D2A, D2B, D1A, D1B, D0A, D0B, OUTL, NLAYER };

void pack_all(Ctx& c, GuessNet& g) {
  for (auto& l : g.L) {
    const int64_t nf = (int64_t)l.nqp * l.cout * 8;
    k_pack<<<div_up(nf, 256), 256, 0, c.stream>>>(g.master.p + l.woff, l.wp.p, l.cin, l.cout, l.nqp, 0);
    if (l.cin % 16 == 0) {
      const int64_t nd = (int64_t)l.nqpd * l.cin * 8;
      k_pack<<<div_up(nd, 256), 256, 0, c.stream>>>(g.master.p + l.woff, l.wd.p, l.cin, l.cout, l.nqpd, 1);
    }
  }
}

template <int NT, int EPI, bool UP>
void launch_conv_t(hipStream_t st, const ConvArgs& a) {
  const int64_t nb = (int64_t)div_up(a.n0, kBI) * div_up(a.n1, kBJ) * div_up(a.n2, kBK);
  const size_t lds = conv_lds_bytes(a.cin);
  static size_t granted = 65536;  // dynamic LDS this instance may use (raised once as needed)
  const size_t most = conv_lds_bytes(64);
  if (lds > granted) {
    FOPIS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv<NT, EPI, UP>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)most));
    granted = most;
  }
  k_conv<NT, EPI, UP><<<(unsigned)nb, 256, lds, st>>>(a);
}

template <int EPI, bool UP>
void launch_conv_nt(hipStream_t st, const ConvArgs& a) {
  switch (a.cout / 16) {
    case 1: launch_conv_t<1, EPI, UP>(st, a); break;
    case 2: launch_conv_t<2, EPI, UP>(st, a); break;
    case 3: launch_conv_t<3, EPI, UP>(st, a); break;
    case 4: launch_conv_t<4, EPI, UP>(st, a); break;
    default: throw Error(FOPIS_E_ARG, "conv: output channels must be 16, 32, 48 or 64");
  }
}

void launch_conv(hipStream_t st, const ConvArgs& a, int epi, bool up) {
  if (a.cin % 8 != 0 || a.cin > 64 || a.cin < 8) throw Error(FOPIS_E_ARG, "conv: input channels must be 8..64, multiple of 8");
  if (up && (a.n0 % 2 || a.n1 % 2 || a.n2 % 2)) throw Error(FOPIS_E_ARG, "conv: upsampled grid must be even");
  if (epi == EPI_BF16) {
    if (up) launch_conv_nt<EPI_BF16, true>(st, a); else launch_conv_nt<EPI_BF16, false>(st, a);
  } else if (epi == EPI_F32) {
    if (up) launch_conv_nt<EPI_F32, true>(st, a); else launch_conv_nt<EPI_F32, false>(st, a);
  } else {
    if (up) launch_conv_nt<EPI_PHI, true>(st, a); else launch_conv_nt<EPI_PHI, false>(st, a);
  }
}

ConvArgs conv_args(const int* d, int cin, int cout, int nq, int nqp, const uint16_t* x, const uint16_t* w,
                   const float* b, int relu) {
  ConvArgs a{};
  a.x = x;
  a.w = w;
  a.bias = b;
  a.n0 = d[0];
  a.n1 = d[1];
  a.n2 = d[2];
  a.cin = cin;
  a.cout = cout;
  a.nq = nq;
  a.nqp = nqp;
  a.relu = relu;
  return a;
}

// Weight and bias gradient of one layer into the flat gradient vector.
void wgrad(Ctx& c, GuessNet& g, const int* d, int cin, int cout, bool up, const uint16_t* x, const uint16_t* gy,
           float* dw, float* db) {
  const int64_t nvox = nvox_of(d);
  const int nci = (cin + 15) / 16, nco = cout / 16, npairs = nci * nco;
  WgradArgs a{};
  a.x = x;
  a.g = gy;
  a.n0 = d[0];
  a.n1 = d[1];
  a.n2 = d[2];
  a.cin = cin;
  a.cout = cout;
  a.nco = nco;
  a.nsteps = (nvox + 31) / 32;
  int nchunk = (int)std::min<int64_t>(std::max(1, 1024 / npairs), (a.nsteps + 3) / 4);
  nchunk = std::max(nchunk, 1);
  const size_t need = (size_t)npairs * nchunk * kWTile;
  if (g.wpart.n < need) {
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    g.wpart.alloc(need);
  }
  a.part = g.wpart.p;
  if (up)
    k_wgrad<true><<<dim3(nchunk, npairs), 256, 0, c.stream>>>(a);
  else
    k_wgrad<false><<<dim3(nchunk, npairs), 256, 0, c.stream>>>(a);
  const int64_t nw = 27LL * cin * cout;
  k_wgrad_reduce<<<div_up(nw, 256), 256, 0, c.stream>>>(g.wpart.p, nchunk, cin, cout, nco, dw);
  const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, nvox / 256));
  if (!g.cpart.p) g.cpart.alloc((size_t)1024 * 64);
  k_colsum<<<nb, 256, 0, c.stream>>>(gy, nvox, cout, g.cpart.p);
  k_colsum_reduce<<<1, 64, 0, c.stream>>>(g.cpart.p, nb, cout, db);
}

void alloc_train(GuessNet& g) {
  if (g.train_alloc) return;
  const int64_t n0 = nvox_of(g.lv[0]), n1 = nvox_of(g.lv[1]), n2 = nvox_of(g.lv[2]), n3 = nvox_of(g.lv[3]);
  g.gOUT.alloc(n0 * 16);
  g.gE0.alloc(n0 * 16);
  g.gD0.alloc(n0 * 16);
  g.gS0.alloc(n0 * 16);
  g.gA0.alloc(n0 * 16);
  g.T0.alloc(n0 * 32);
  g.gE1.alloc(n1 * 32);
  g.gD1.alloc(n1 * 32);
  g.gS1.alloc(n1 * 32);
  g.gA1.alloc(n1 * 32);
  g.gP0.alloc(n1 * 16);
  g.T1.alloc(n1 * 48);
  g.gE2.alloc(n2 * 48);
  g.gD2.alloc(n2 * 48);
  g.gS2.alloc(n2 * 48);
  g.gA2.alloc(n2 * 48);
  g.gP1.alloc(n2 * 32);
  g.T2.alloc(n2 * 64);
  g.gB3.alloc(n3 * 64);
  g.gA3.alloc(n3 * 64);
  g.gP2.alloc(n3 * 48);
  g.m1.alloc(g.nparams);
  g.m2.alloc(g.nparams);
  FOPIS_HIP(hipMemset(g.m1.p, 0, g.nparams * sizeof(float)));
  FOPIS_HIP(hipMemset(g.m2.p, 0, g.nparams * sizeof(float)));
  g.train_alloc = true;
}

// Forward pass from g.X0. epi = EPI_F32 writes g.OUT, EPI_PHI writes phi.
void forward(Ctx& c, GuessNet& g, int epi, bool phi_add = false) {
  hipStream_t st = c.stream;
  auto conv = [&](int li, const uint16_t* x, const uint16_t* add, uint16_t* y) {
    GuessLayer& l = g.L[li];
    ConvArgs a = conv_args(g.lv[l.lvl], l.cin, l.cout, l.nq, l.nqp, x, l.wp.p, g.master.p + l.boff, l.relu);
    a.add = add;
    a.y = y;
    launch_conv(st, a, EPI_BF16, l.up != 0);
  };
  auto pool = [&](const uint16_t* xf, uint16_t* xc, int lc, int C) {
    const int* d = g.lv[lc];
    k_pool<<<div_up(nvox_of(d) * (C / 8), 256), 256, 0, st>>>(xf, xc, d[0], d[1], d[2], C);
  };
  conv(E0A, g.X0.p, nullptr, g.A0.p);
  conv(E0B, g.A0.p, nullptr, g.S0.p);
  pool(g.S0.p, g.P0.p, 1, 16);
  conv(E1A, g.P0.p, nullptr, g.A1.p);
  conv(E1B, g.A1.p, nullptr, g.S1.p);
  pool(g.S1.p, g.P1.p, 2, 32);
  conv(E2A, g.P1.p, nullptr, g.A2.p);
  conv(E2B, g.A2.p, nullptr, g.S2.p);
  pool(g.S2.p, g.P2.p, 3, 48);
  conv(E3A, g.P2.p, nullptr, g.A3.p);
  conv(E3B, g.A3.p, nullptr, g.B3.p);
  conv(D2A, g.B3.p, g.S2.p, g.D2.p);
  conv(D2B, g.D2.p, nullptr, g.E2.p);
  conv(D1A, g.E2.p, g.S1.p, g.D1.p);
  conv(D1B, g.D1.p, nullptr, g.E1.p);
  conv(D0A, g.E1.p, g.S0.p, g.D0.p);
  conv(D0B, g.D0.p, nullptr, g.E0.p);
  GuessLayer& lo = g.L[OUTL];
  ConvArgs a = conv_args(g.lv[0], lo.cin, lo.cout, lo.nq, lo.nqp, g.E0.p, lo.wp.p, g.master.p + lo.boff, 0);
  if (epi == EPI_F32) {
    a.y32 = g.OUT.p;
  } else {
    a.phi = c.phi.p;
    a.scal = g.scal.p;
    a.phi_k = g.k_out / c.d.eps0;
    a.phi_add = phi_add ? 1 : 0;
    a.gsx = g.G.gsx;
    a.gsy = g.G.gsy;
    a.goff = g.G.goff;
    a.vn0 = g.G.vn0;
    a.vn1 = g.G.vn1;
    a.vn2 = g.G.vn2;
  }
  launch_conv(st, a, epi, false);
}

// Backward pass from g.gOUT; fills g.grad.
void backward(Ctx& c, GuessNet& g) {
  hipStream_t st = c.stream;
  auto relu_bwd = [&](uint16_t* gy, const uint16_t* y, int lvl, int C) {
    const int64_t n = nvox_of(g.lv[lvl]) * C;
    k_relu_bwd<<<div_up(n, 256), 256, 0, st>>>(gy, y, n);
  };
  auto wg = [&](int li, const uint16_t* x, const uint16_t* gy) {
    GuessLayer& l = g.L[li];
    wgrad(c, g, g.lv[l.lvl], l.cin, l.cout, l.up != 0, x, gy, g.grad.p + l.woff, g.grad.p + l.boff);
  };
  // data gradient of layer li at its output grid into dx ([out voxels][cin])
  auto dg = [&](int li, const uint16_t* gy, uint16_t* dx) {
    GuessLayer& l = g.L[li];
    const float* zero_bias = g.master.p + g.nparams;  // cin zeros kept after the parameters
    ConvArgs a = conv_args(g.lv[l.lvl], l.cout, l.cin, l.nqd, l.nqpd, gy, l.wd.p, zero_bias, 0);
    a.y = dx;
    launch_conv(st, a, EPI_BF16, false);
  };
  auto upsum = [&](const uint16_t* gf, uint16_t* gcoarse, int lc, int C) {
    const int* d = g.lv[lc];
    k_upsum<<<div_up(nvox_of(d) * C, 256), 256, 0, st>>>(gf, gcoarse, d[0], d[1], d[2], C);
  };
  auto poolb = [&](const uint16_t* gcoarse, const uint16_t* skip, uint16_t* gf, int lf, int C) {
    const int* d = g.lv[lf];
    k_pool_bwd<<<div_up(nvox_of(d) * C, 256), 256, 0, st>>>(gcoarse, skip, gf, d[0], d[1], d[2], C);
  };
  wg(OUTL, g.E0.p, g.gOUT.p);
  dg(OUTL, g.gOUT.p, g.gE0.p);
  relu_bwd(g.gE0.p, g.E0.p, 0, 16);
  wg(D0B, g.D0.p, g.gE0.p);
  dg(D0B, g.gE0.p, g.gD0.p);
  relu_bwd(g.gD0.p, g.D0.p, 0, 16);
  wg(D0A, g.E1.p, g.gD0.p);
  dg(D0A, g.gD0.p, g.T0.p);
  upsum(g.T0.p, g.gE1.p, 1, 32);
  relu_bwd(g.gE1.p, g.E1.p, 1, 32);
  wg(D1B, g.D1.p, g.gE1.p);
  dg(D1B, g.gE1.p, g.gD1.p);
  relu_bwd(g.gD1.p, g.D1.p, 1, 32);
  wg(D1A, g.E2.p, g.gD1.p);
  dg(D1A, g.gD1.p, g.T1.p);
  upsum(g.T1.p, g.gE2.p, 2, 48);
  relu_bwd(g.gE2.p, g.E2.p, 2, 48);
  wg(D2B, g.D2.p, g.gE2.p);
  dg(D2B, g.gE2.p, g.gD2.p);
  relu_bwd(g.gD2.p, g.D2.p, 2, 48);
  wg(D2A, g.B3.p, g.gD2.p);
  dg(D2A, g.gD2.p, g.T2.p);
  upsum(g.T2.p, g.gB3.p, 3, 64);
  relu_bwd(g.gB3.p, g.B3.p, 3, 64);
  wg(E3B, g.A3.p, g.gB3.p);
  dg(E3B, g.gB3.p, g.gA3.p);
  relu_bwd(g.gA3.p, g.A3.p, 3, 64);
  wg(E3A, g.P2.p, g.gA3.p);
  dg(E3A, g.gA3.p, g.gP2.p);
  poolb(g.gP2.p, g.gD2.p, g.gS2.p, 2, 48);
  relu_bwd(g.gS2.p, g.S2.p, 2, 48);
  wg(E2B, g.A2.p, g.gS2.p);
  dg(E2B, g.gS2.p, g.gA2.p);
  relu_bwd(g.gA2.p, g.A2.p, 2, 48);
  wg(E2A, g.P1.p, g.gA2.p);
  dg(E2A, g.gA2.p, g.gP1.p);
  poolb(g.gP1.p, g.gD1.p, g.gS1.p, 1, 32);
  relu_bwd(g.gS1.p, g.S1.p, 1, 32);
  wg(E1B, g.A1.p, g.gS1.p);
  dg(E1B, g.gS1.p, g.gA1.p);
  relu_bwd(g.gA1.p, g.A1.p, 1, 32);
  wg(E1A, g.P0.p, g.gA1.p);
  dg(E1A, g.gA1.p, g.gP0.p);
  poolb(g.gP0.p, g.gD0.p, g.gS0.p, 0, 16);
  relu_bwd(g.gS0.p, g.S0.p, 0, 16);
  wg(E0B, g.A0.p, g.gS0.p);
  dg(E0B, g.gS0.p, g.gA0.p);
  relu_bwd(g.gA0.p, g.A0.p, 0, 16);
  wg(E0A, g.X0.p, g.gA0.p);
}

GuessNet& net(Ctx& c) {
  if (!c.guess) throw Error(FOPIS_E_STATE, "the learned guess is not initialised (fopis_guess_init)");
  return *c.guess;
}

void features_from_slot(Ctx& c, GuessNet& g, int slot) {
  const int64_t n = nvox_of(g.lv[0]);
  k_features_pair<<<div_up(n, 256), 256, 0, c.stream>>>(g.ds_rho.p + (size_t)slot * n, g.ds_mask.p + (size_t)slot * n,
                                                         g.X0.p, n);
}

// Loss of pair `slot` on g.OUT into *dst (device) and its gradient into g.gOUT.
void launch_loss(Ctx& c, GuessNet& g, int slot, double* dst) {
  const int64_t n = nvox_of(g.lv[0]);
  const int nb = 256;
  const double nvalid = (double)g.G.vn0 * g.G.vn1 * g.G.vn2;
  if (g.loss_kind == 1) {
    if (!g.q.p) g.q.alloc(n);
    k_res_loss<<<nb, 256, 0, c.stream>>>(g.OUT.p, g.ds_rho.p + (size_t)slot * n, (float)g.k_out, g.G, g.q.p,
                                         g.dpart.p);
    k_res_grad<<<div_up(n, 256), 256, 0, c.stream>>>(g.q.p, (float)(g.k_out / nvalid), g.G, g.gOUT.p);
  } else {
    k_loss<<<nb, 256, 0, c.stream>>>(g.OUT.p, g.ds_phi.p + (size_t)slot * n, (float)(1.0 / g.k_out), g.G,
                                     (float)(1.0 / nvalid), g.gOUT.p, g.dpart.p);
  }
  k_sum_into<<<1, 1, 0, c.stream>>>(g.dpart.p, nb, 1.0 / nvalid, dst);
}

double loss_and_grad(Ctx& c, GuessNet& g, int slot, bool sync) {
  launch_loss(c, g, slot, g.loss_acc.p);
  if (!sync) return 0.0;
  double l = 0.0;
  FOPIS_HIP(hipMemcpyAsync(&l, g.loss_acc.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  return l;
}

void compute_scale(Ctx& c, GuessNet& g, const double* rho) {
  const int nb = 1024;
  k_rho2<<<nb, 256, 0, c.stream>>>(rho, g.G, g.dpart.p);
  k_rho_scale<<<1, 256, 0, c.stream>>>(g.dpart.p, nb, (int64_t)g.G.vn0 * g.G.vn1 * g.G.vn2, g.scal.p);
}

}  // namespace

// ---------------------------------------------------------------- entry points

void guess_free(Ctx& c) {
  delete c.guess;
  c.guess = nullptr;
}

void guess_init(Ctx& c, uint64_t seed) {
  guess_free(c);
  GuessNet* gp = new GuessNet();
  c.guess = gp;
  GuessNet& g = *gp;
  g.seed = seed;
  const int vn[3] = {c.nx - 2, c.ny - 2, c.nz - 2};
  for (int q = 0; q < 3; ++q) {
    const int n = (vn[q] + 7) / 8 * 8;
    for (int l = 0; l < 4; ++l) g.lv[l][q] = n >> l;
  }
  g.G.n0 = g.lv[0][0];
  g.G.n1 = g.lv[0][1];
  g.G.n2 = g.lv[0][2];
  g.G.vn0 = vn[0];
  g.G.vn1 = vn[1];
  g.G.vn2 = vn[2];
  g.G.gsx = (int64_t)c.cy * c.cz;
  g.G.gsy = c.cz;
  g.G.goff = (int64_t)c.id3(2, 2, 2);
  g.G.mcy = c.cy - 1;
  g.G.mcz = c.cz - 1;
  add_layer(g, 8, 16, 0, 0, 1);
  add_layer(g, 16, 16, 0, 0, 1);
  add_layer(g, 16, 32, 1, 0, 1);
  add_layer(g, 32, 32, 1, 0, 1);
  add_layer(g, 32, 48, 2, 0, 1);
  add_layer(g, 48, 48, 2, 0, 1);
  add_layer(g, 48, 64, 3, 0, 1);
  add_layer(g, 64, 64, 3, 0, 1);
  add_layer(g, 64, 48, 2, 1, 1);
  add_layer(g, 48, 48, 2, 0, 1);
  add_layer(g, 48, 32, 1, 1, 1);
  add_layer(g, 32, 32, 1, 0, 1);
  add_layer(g, 32, 16, 0, 1, 1);
  add_layer(g, 16, 16, 0, 0, 1);
  add_layer(g, 16, 16, 0, 0, 0);  // out: only channel 0 is used
  for (auto& l : g.L) {
    l.wp.alloc((size_t)l.nqp * l.cout * 8);
    if (l.cin % 16 == 0) l.wd.alloc((size_t)l.nqpd * l.cin * 8);
  }
  // He-normal initialisation (std::mt19937_64, fixed seed); biases zero; the
  // output layer starts small and only channel 0 is live.
  std::vector<float> h(g.nparams + 64, 0.f);
  std::mt19937_64 rng(seed);
  for (int li = 0; li < NLAYER; ++li) {
    GuessLayer& l = g.L[li];
    const int cin_live = li == E0A ? 2 : l.cin;
    const double sd = std::sqrt(2.0 / (27.0 * cin_live)) * (li == OUTL ? 0.1 : 1.0);
    std::normal_distribution<double> nd(0.0, sd);
    for (int t = 0; t < 27; ++t)
      for (int ci = 0; ci < l.cin; ++ci)
        for (int co = 0; co < l.cout; ++co) {
          float v = (float)nd(rng);
          if (li == E0A && ci >= 2) v = 0.f;
          if (li == OUTL && co != 0) v = 0.f;
          h[l.woff + ((size_t)t * l.cin + ci) * l.cout + co] = v;
        }
  }
  // the parameter buffer carries 64 zeros after the parameters (zero bias of the data-gradient convs)
  g.master.alloc(g.nparams + 64);
  g.grad.alloc(g.nparams + 64);
  FOPIS_HIP(hipMemcpy(g.master.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  FOPIS_HIP(hipMemset(g.grad.p, 0, (g.nparams + 64) * sizeof(float)));
  const int64_t n0 = nvox_of(g.lv[0]), n1 = nvox_of(g.lv[1]), n2 = nvox_of(g.lv[2]), n3 = nvox_of(g.lv[3]);
  g.X0.alloc(n0 * 8);
  g.A0.alloc(n0 * 16);
  g.S0.alloc(n0 * 16);
  g.D0.alloc(n0 * 16);
  g.E0.alloc(n0 * 16);
  g.P0.alloc(n1 * 16);
  g.A1.alloc(n1 * 32);
  g.S1.alloc(n1 * 32);
  g.D1.alloc(n1 * 32);
  g.E1.alloc(n1 * 32);
  g.P1.alloc(n2 * 32);
  g.A2.alloc(n2 * 48);
  g.S2.alloc(n2 * 48);
  g.D2.alloc(n2 * 48);
  g.E2.alloc(n2 * 48);
  g.P2.alloc(n3 * 48);
  g.A3.alloc(n3 * 64);
  g.B3.alloc(n3 * 64);
  g.OUT.alloc(n0);
  g.dpart.alloc(1024);
  g.scal.alloc(4);
  g.loss_acc.alloc(4096);
  pack_all(c, g);
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

int64_t guess_num_params(Ctx& c) { return net(c).nparams; }

void guess_get_params(Ctx& c, float* out, double* k_out) {
  GuessNet& g = net(c);
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  FOPIS_HIP(hipMemcpy(out, g.master.p, g.nparams * sizeof(float), hipMemcpyDeviceToHost));
  if (k_out) *k_out = g.k_out;
}

void guess_set_params(Ctx& c, const float* in, double k_out) {
  GuessNet& g = net(c);
  if (!(k_out > 0.0)) throw Error(FOPIS_E_ARG, "output scale must be positive");
  FOPIS_HIP(hipMemcpyAsync(g.master.p, in, g.nparams * sizeof(float), hipMemcpyHostToDevice, c.stream));
  g.k_out = k_out;
  pack_all(c, g);
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

void guess_set_loss(Ctx& c, int kind) {
  if (kind != 0 && kind != 1) throw Error(FOPIS_E_ARG, "loss kind must be 0 or 1");
  net(c).loss_kind = kind;
}

void guess_set_mode(Ctx& c, int mode) {
  if (mode < 0 || mode > 3) throw Error(FOPIS_E_ARG, "guess mode must be 0, 1, 2 or 3");
  if (mode != 0) net(c);
  if (c.guess) {
    c.guess->mode = mode;
    if (mode == 3) c.guess->record_slot = 0;
  }
}

int guess_mode(const Ctx& c) { return c.guess ? c.guess->mode : 0; }

// Modes 2 and 3 before the first solve of a step: the warm start's residual
// (in charge units) into rwarm. Mode 2 then adds the network's potential for
// that charge to phi; mode 3 keeps phi for the pair recorded after the solve.
void guess_before_solve(Ctx& c) {
  GuessNet& g = net(c);
  Phase ph(c, "guess");
  g.rwarm.alloc(c.nnode);
  const int64_t n = (int64_t)g.G.vn0 * g.G.vn1 * g.G.vn2;
  k_warm_resid<<<div_up(n, 256), 256, 0, c.stream>>>(c.phi.p, c.rho.p, c.d.eps0, g.G, g.rwarm.p);
  if (g.mode == 3) {
    g.phi_start.alloc(c.nnode);
    FOPIS_HIP(hipMemcpyAsync(g.phi_start.p, c.phi.p, c.nnode * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return;
  }
  compute_scale(c, g, g.rwarm.p);
  const uint8_t* cm = g.use_cellmask ? g.cellmask.p : (c.have_mask ? c.mask.p : nullptr);
  k_features_field<<<div_up(nvox_of(g.lv[0]), 256), 256, 0, c.stream>>>(g.rwarm.p, cm, g.G, g.scal.p, g.X0.p);
  forward(c, g, EPI_PHI, true);
  FOPIS_HIP(hipGetLastError());
}

// Mode 3 after the solve: the pair (eps0 r0 / sigma, mask, (phi - phi_start) eps0 / sigma)
// into the next dataset slot, while there is room.
void guess_after_solve(Ctx& c) {
  GuessNet& g = net(c);
  if (g.mode != 3 || g.record_slot >= g.capacity || !g.phi_start.p) return;
  const int64_t n = (int64_t)g.G.vn0 * g.G.vn1 * g.G.vn2;
  k_phi_delta<<<div_up(n, 256), 256, 0, c.stream>>>(c.phi.p, g.phi_start.p, g.G);
  compute_scale(c, g, g.rwarm.p);
  const int slot = g.record_slot++;
  const int64_t nv = nvox_of(g.lv[0]);
  const uint8_t* cm = g.use_cellmask ? g.cellmask.p : (c.have_mask ? c.mask.p : nullptr);
  k_store_pair<<<div_up(nv, 256), 256, 0, c.stream>>>(g.rwarm.p, g.phi_start.p, cm, g.G, g.scal.p, c.d.eps0,
                                                     g.ds_rho.p + (size_t)slot * nv, g.ds_mask.p + (size_t)slot * nv,
                                                     g.ds_phi.p + (size_t)slot * nv);
  std::vector<uint16_t> h(nv);
  FOPIS_HIP(hipMemcpyAsync(h.data(), g.ds_phi.p + (size_t)slot * nv, nv * sizeof(uint16_t), hipMemcpyDeviceToHost,
                           c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (int64_t v = 0; v < nv; ++v) {
    const double x = bf2f(h[v]);
    s += x * x;
  }
  g.ds_phi_rms[slot] = std::sqrt(s / ((double)g.G.vn0 * g.G.vn1 * g.G.vn2));
}

int guess_recorded(const Ctx& c) { return c.guess ? c.guess->record_slot : 0; }

// phi at interior nodes <- network(rho[0], conductor mask).
void guess_apply(Ctx& c) {
  GuessNet& g = net(c);
  Phase ph(c, "guess");
  compute_scale(c, g, c.rho.p);
  const uint8_t* cm = g.use_cellmask ? g.cellmask.p : (c.have_mask ? c.mask.p : nullptr);
  const int64_t n = nvox_of(g.lv[0]);
  k_features_field<<<div_up(n, 256), 256, 0, c.stream>>>(c.rho.p, cm, g.G, g.scal.p, g.X0.p);
  forward(c, g, EPI_PHI);
  FOPIS_HIP(hipGetLastError());
}

void guess_dataset_alloc(Ctx& c, int capacity) {
  GuessNet& g = net(c);
  if (capacity < 1) throw Error(FOPIS_E_ARG, "dataset capacity must be >= 1");
  const int64_t n = nvox_of(g.lv[0]);
  g.ds_rho.alloc((size_t)capacity * n);
  g.ds_phi.alloc((size_t)capacity * n);
  g.ds_mask.alloc((size_t)capacity * n);
  g.capacity = capacity;
  g.ds_phi_rms.assign(capacity, 0.0);
}

static void check_slot(const GuessNet& g, int slot) {
  if (slot < 0 || slot >= g.capacity) throw Error(FOPIS_E_ARG, "dataset slot out of range");
}

// Store (rho[0], node mask, phi) of the context as pair `slot`; returns the rms of phi_hat.
double guess_dataset_store(Ctx& c, int slot) {
  GuessNet& g = net(c);
  check_slot(g, slot);
  compute_scale(c, g, c.rho.p);
  const int64_t n = nvox_of(g.lv[0]);
  const uint8_t* cm = g.use_cellmask ? g.cellmask.p : (c.have_mask ? c.mask.p : nullptr);
  k_store_pair<<<div_up(n, 256), 256, 0, c.stream>>>(c.rho.p, c.phi.p, cm, g.G, g.scal.p, c.d.eps0,
                                                    g.ds_rho.p + (size_t)slot * n, g.ds_mask.p + (size_t)slot * n,
                                                    g.ds_phi.p + (size_t)slot * n);
  std::vector<uint16_t> h(n);
  FOPIS_HIP(hipMemcpyAsync(h.data(), g.ds_phi.p + (size_t)slot * n, n * sizeof(uint16_t), hipMemcpyDeviceToHost,
                           c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (int64_t v = 0; v < n; ++v) {
    const double x = bf2f(h[v]);
    s += x * x;
  }
  const double nvalid = (double)g.G.vn0 * g.G.vn1 * g.G.vn2;
  g.ds_phi_rms[slot] = std::sqrt(s / nvalid);
  return g.ds_phi_rms[slot];
}

void guess_dataset_get(Ctx& c, int slot, uint16_t* rho_hat, uint8_t* mask, uint16_t* phi_hat) {
  GuessNet& g = net(c);
  check_slot(g, slot);
  const int64_t n = nvox_of(g.lv[0]);
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  if (rho_hat) FOPIS_HIP(hipMemcpy(rho_hat, g.ds_rho.p + (size_t)slot * n, n * 2, hipMemcpyDeviceToHost));
  if (mask) FOPIS_HIP(hipMemcpy(mask, g.ds_mask.p + (size_t)slot * n, n, hipMemcpyDeviceToHost));
  if (phi_hat) FOPIS_HIP(hipMemcpy(phi_hat, g.ds_phi.p + (size_t)slot * n, n * 2, hipMemcpyDeviceToHost));
}

void guess_dataset_set(Ctx& c, int slot, const uint16_t* rho_hat, const uint8_t* mask, const uint16_t* phi_hat) {
  GuessNet& g = net(c);
  check_slot(g, slot);
  const int64_t n = nvox_of(g.lv[0]);
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  FOPIS_HIP(hipMemcpy(g.ds_rho.p + (size_t)slot * n, rho_hat, n * 2, hipMemcpyHostToDevice));
  FOPIS_HIP(hipMemcpy(g.ds_mask.p + (size_t)slot * n, mask, n, hipMemcpyHostToDevice));
  FOPIS_HIP(hipMemcpy(g.ds_phi.p + (size_t)slot * n, phi_hat, n * 2, hipMemcpyHostToDevice));
  double s = 0.0;
  for (int64_t v = 0; v < n; ++v) {
    const double x = bf2f(phi_hat[v]);
    s += x * x;
  }
  g.ds_phi_rms[slot] = std::sqrt(s / ((double)g.G.vn0 * g.G.vn1 * g.G.vn2));
}

// Loss and flat parameter gradient on pair `slot` (no update).
double guess_grad(Ctx& c, int slot, float* grad_out) {
  GuessNet& g = net(c);
  check_slot(g, slot);
  alloc_train(g);
  features_from_slot(c, g, slot);
  forward(c, g, EPI_F32);
  loss_and_grad(c, g, slot, false);
  backward(c, g);
  double l = 0.0;
  FOPIS_HIP(hipMemcpyAsync(&l, g.loss_acc.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  if (grad_out) FOPIS_HIP(hipMemcpy(grad_out, g.grad.p, g.nparams * sizeof(float), hipMemcpyDeviceToHost));
  FOPIS_HIP(hipGetLastError());
  return l;
}

// Adam over pairs [0, nsamples) for `epochs` passes in shuffled order; the
// learning rate follows a cosine from lr to lr_end. loss_out[e] = mean loss of epoch e.
void guess_train(Ctx& c, int nsamples, int epochs, double lr, double lr_end, uint64_t seed, double* loss_out) {
  GuessNet& g = net(c);
  if (nsamples < 1 || nsamples > g.capacity) throw Error(FOPIS_E_ARG, "nsamples out of range of the dataset");
  if (epochs < 1) throw Error(FOPIS_E_ARG, "epochs must be >= 1");
  alloc_train(g);
  {
    // output scale: rms of phi_hat over the training pairs
    double s = 0.0;
    for (int i = 0; i < nsamples; ++i) s += g.ds_phi_rms[i] * g.ds_phi_rms[i];
    s = std::sqrt(s / nsamples);
    if (g.adam_t == 0 && s > 0.0) g.k_out = s;
  }
  std::mt19937_64 rng(seed);
  std::vector<int> order(nsamples);
  for (int i = 0; i < nsamples; ++i) order[i] = i;
  const int64_t total = (int64_t)epochs * nsamples;
  int64_t step = 0;
  const int nacc = 4096;
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(order.begin(), order.end(), rng);
    double esum = 0.0;
    int pending = 0;
    auto flush = [&]() {
      if (!pending) return;
      std::vector<double> h(pending);
      FOPIS_HIP(hipMemcpyAsync(h.data(), g.loss_acc.p + 1, pending * sizeof(double), hipMemcpyDeviceToHost,
                               c.stream));
      FOPIS_HIP(hipStreamSynchronize(c.stream));
      for (double x : h) esum += x;
      pending = 0;
    };
    for (int i = 0; i < nsamples; ++i) {
      const int slot = order[i];
      features_from_slot(c, g, slot);
      forward(c, g, EPI_F32);
      launch_loss(c, g, slot, g.loss_acc.p + 1 + pending);
      ++pending;
      backward(c, g);
      ++g.adam_t;
      const double frac = (double)step / std::max<int64_t>(1, total - 1);
      const double lr_t = lr_end + (lr - lr_end) * 0.5 * (1.0 + std::cos(M_PI * frac));
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)g.adam_t), c2 = 1.f - std::pow(b2, (float)g.adam_t);
      k_adam<<<div_up(g.nparams, 256), 256, 0, c.stream>>>(g.master.p, g.grad.p, g.m1.p, g.m2.p, g.nparams,
                                                           (float)lr_t, b1, b2, 1e-8f, c1, c2);
      pack_all(c, g);
      ++step;
      if (pending == nacc - 1) flush();
    }
    flush();
    if (loss_out) loss_out[e] = esum / nsamples;
  }
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  FOPIS_HIP(hipGetLastError());
}

// Mean loss over pairs [begin, end) with the current parameters.
double guess_eval(Ctx& c, int begin, int end) {
  GuessNet& g = net(c);
  if (begin < 0 || end > g.capacity || begin >= end) throw Error(FOPIS_E_ARG, "bad pair range");
  alloc_train(g);
  double s = 0.0;
  for (int i = begin; i < end; ++i) {
    features_from_slot(c, g, i);
    forward(c, g, EPI_F32);
    s += loss_and_grad(c, g, i, true);
  }
  return s / (end - begin);
}

// Training-data conductor boxes: sets the guess's own cell mask and edits rho[0].
void guess_synth_boxes(Ctx& c, int nbox, const int32_t* lohi, const double* q, uint64_t seed) {
  GuessNet& g = net(c);
  if (nbox < 0 || nbox > 8) throw Error(FOPIS_E_ARG, "nbox must be 0..8");
  const int c0 = c.cx - 1, c1 = c.cy - 1, c2 = c.cz - 1;
  std::vector<Box> hb(std::max(nbox, 1));
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> ud(0.0, 6.283185307179586);
  for (int b = 0; b < nbox; ++b) {
    for (int a = 0; a < 3; ++a) {
      hb[b].lo[a] = lohi[6 * b + a];
      hb[b].hi[a] = lohi[6 * b + 3 + a];
      const int lim = a == 0 ? c0 : (a == 1 ? c1 : c2);
      if (hb[b].lo[a] < 1 || hb[b].hi[a] > lim - 1 || hb[b].lo[a] >= hb[b].hi[a])
        throw Error(FOPIS_E_ARG, "box outside the grid's interior cells");
    }
    hb[b].q = q[b];
    for (int a = 0; a < 3; ++a) hb[b].ph[a] = ud(rng);
  }
  g.cellmask.alloc((size_t)c0 * c1 * c2);
  g.boxes.alloc(hb.size());
  FOPIS_HIP(hipMemcpyAsync(g.boxes.p, hb.data(), hb.size() * sizeof(Box), hipMemcpyHostToDevice, c.stream));
  const int64_t nm = (int64_t)c0 * c1 * c2;
  k_box_mask<<<div_up(nm, 256), 256, 0, c.stream>>>(g.cellmask.p, c0, c1, c2, g.boxes.p, nbox);
  k_box_charge<<<div_up((int64_t)c.nnode, 256), 256, 0, c.stream>>>(c.rho.p, c.cx, c.cy, c.cz, g.cellmask.p,
                                                                     g.boxes.p, nbox);
  g.use_cellmask = nbox > 0;
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  FOPIS_HIP(hipGetLastError());
}

void guess_clear_boxes(Ctx& c) { net(c).use_cellmask = false; }

// ---- single-layer entry points (host buffers) for tests and benchmarks

namespace {
struct TmpStream {
  hipStream_t s = nullptr;
  TmpStream() { FOPIS_HIP(hipStreamCreate(&s)); }
  ~TmpStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

void check_conv_shape(int n0, int n1, int n2, int cin, int cout, int up) {
  if (n0 < 1 || n1 < 1 || n2 < 1) throw Error(FOPIS_E_ARG, "grid must be positive");
  if (cin < 8 || cin > 64 || cin % 8) throw Error(FOPIS_E_ARG, "cin must be 8..64, multiple of 8");
  if (cout < 16 || cout > 64 || cout % 16) throw Error(FOPIS_E_ARG, "cout must be 16, 32, 48 or 64");
  if (up && (n0 % 2 || n1 % 2 || n2 % 2)) throw Error(FOPIS_E_ARG, "upsampled grid must be even");
}
}  // namespace

void nn_conv3d(int n0, int n1, int n2, int cin, int cout, int up, int relu, const uint16_t* x, const float* w,
               const float* b, const uint16_t* add, uint16_t* y, int reps, float* ms) {
  check_conv_shape(n0, n1, n2, cin, cout, up);
  TmpStream ts;
  const int64_t nv = (int64_t)n0 * n1 * n2, nin = up ? nv / 8 : nv;
  const int nq = 27 * cin / 8, nqp = (nq + 3) / 4 * 4;
  DBuf<uint16_t> dx, dy, dadd, dwp;
  DBuf<float> dw, db;
  dx.alloc(nin * cin);
  dy.alloc(nv * cout);
  dw.alloc(27LL * cin * cout);
  db.alloc(cout);
  dwp.alloc((size_t)nqp * cout * 8);
  FOPIS_HIP(hipMemcpy(dx.p, x, nin * cin * 2, hipMemcpyHostToDevice));
  FOPIS_HIP(hipMemcpy(dw.p, w, 27LL * cin * cout * 4, hipMemcpyHostToDevice));
  FOPIS_HIP(hipMemcpy(db.p, b, cout * 4, hipMemcpyHostToDevice));
  if (add) {
    dadd.alloc(nv * cout);
    FOPIS_HIP(hipMemcpy(dadd.p, add, nv * cout * 2, hipMemcpyHostToDevice));
  }
  k_pack<<<div_up((int64_t)nqp * cout * 8, 256), 256, 0, ts.s>>>(dw.p, dwp.p, cin, cout, nqp, 0);
  const int dims[3] = {n0, n1, n2};
  ConvArgs a = conv_args(dims, cin, cout, nq, nqp, dx.p, dwp.p, db.p, relu);
  a.add = add ? dadd.p : nullptr;
  a.y = dy.p;
  launch_conv(ts.s, a, EPI_BF16, up != 0);
  if (reps > 0 && ms) {
    hipEvent_t e0, e1;
    FOPIS_HIP(hipEventCreate(&e0));
    FOPIS_HIP(hipEventCreate(&e1));
    FOPIS_HIP(hipEventRecord(e0, ts.s));
    for (int r = 0; r < reps; ++r) launch_conv(ts.s, a, EPI_BF16, up != 0);
    FOPIS_HIP(hipEventRecord(e1, ts.s));
    FOPIS_HIP(hipEventSynchronize(e1));
    FOPIS_HIP(hipEventElapsedTime(ms, e0, e1));
    *ms /= reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  FOPIS_HIP(hipStreamSynchronize(ts.s));
  FOPIS_HIP(hipGetLastError());
  FOPIS_HIP(hipMemcpy(y, dy.p, nv * cout * 2, hipMemcpyDeviceToHost));
}

void nn_conv3d_backward(int n0, int n1, int n2, int cin, int cout, int up, const uint16_t* x, const float* w,
                        const uint16_t* gy, float* dw_out, float* db_out, uint16_t* dx_out) {
  check_conv_shape(n0, n1, n2, cin, cout, up);
  if (dx_out && cin % 16) throw Error(FOPIS_E_ARG, "data gradient needs cin multiple of 16");
  TmpStream ts;
  const int64_t nv = (int64_t)n0 * n1 * n2, nin = up ? nv / 8 : nv;
  DBuf<uint16_t> dx, dg, dwd, ddx;
  DBuf<float> dw, dgw, dgb, zb;
  dx.alloc(nin * cin);
  dg.alloc(nv * cout);
  dw.alloc(27LL * cin * cout);
  dgw.alloc(27LL * cin * cout);
  dgb.alloc(cout);
  zb.alloc(64);
  FOPIS_HIP(hipMemcpy(dx.p, x, nin * cin * 2, hipMemcpyHostToDevice));
  FOPIS_HIP(hipMemcpy(dg.p, gy, nv * cout * 2, hipMemcpyHostToDevice));
  FOPIS_HIP(hipMemcpy(dw.p, w, 27LL * cin * cout * 4, hipMemcpyHostToDevice));
  FOPIS_HIP(hipMemset(zb.p, 0, 64 * 4));
  // weight and bias gradients through a throwaway network object
  GuessNet tmp;
  Ctx cc;
  cc.stream = ts.s;
  const int dims[3] = {n0, n1, n2};
  wgrad(cc, tmp, dims, cin, cout, up != 0, dx.p, dg.p, dgw.p, dgb.p);
  cc.stream = nullptr;
  if (dx_out) {
    const int nqd = 27 * cout / 8, nqpd = (nqd + 3) / 4 * 4;
    dwd.alloc((size_t)nqpd * cin * 8);
    k_pack<<<div_up((int64_t)nqpd * cin * 8, 256), 256, 0, ts.s>>>(dw.p, dwd.p, cin, cout, nqpd, 1);
    ddx.alloc(nv * cin);
    ConvArgs a = conv_args(dims, cout, cin, nqd, nqpd, dg.p, dwd.p, zb.p, 0);
    a.y = ddx.p;
    launch_conv(ts.s, a, EPI_BF16, false);
    FOPIS_HIP(hipStreamSynchronize(ts.s));
    FOPIS_HIP(hipMemcpy(dx_out, ddx.p, nv * cin * 2, hipMemcpyDeviceToHost));
  }
  FOPIS_HIP(hipStreamSynchronize(ts.s));
  FOPIS_HIP(hipGetLastError());
  FOPIS_HIP(hipMemcpy(dw_out, dgw.p, 27LL * cin * cout * 4, hipMemcpyDeviceToHost));
  FOPIS_HIP(hipMemcpy(db_out, dgb.p, cout * 4, hipMemcpyDeviceToHost));
}

}  // namespace fopis
