// Particle hot path on gfx950: tile-binned SoA particles, the fused
// gather-Boris-move-absorb-count-deposit-bin kernel, arrivals, re-binning and
// particle reductions.
//
// Reference behaviour followed (hsimyu/fopis):
//   position/weights  include/position.hpp:13-21,59-69
//   velocity (Boris)  src/Grid/GridParticleTreatmentES.cpp:13-88
//   position + absorb src/Grid/GridParticleTreatmentES.cpp:90-113, include/grid.hpp:309-395
//   density           src/Grid/RootGrid/RootGridField.cpp:1129-1150
//   CIC deposit       src/Grid/RootGrid/RootGridField.cpp:1161-1192
//   conductor capture src/Spacecraft/Spacecraft.cpp:512-613, src/Particle/particle.cpp:46-204
//   energy            src/Grid/GridParticleTreatment.cpp:104-118
//
// FP64 arithmetic is written in the reference's operation order and compiled
// without contraction, so an FP64 push reproduces the reference bit for bit
// on identical inputs.
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "fopis_internal.hpp"
#include "radix.hpp"

#pragma clang fp contract(off)

namespace fopis {

enum : int { M_VEL = 1, M_POS = 2, M_DEP = 4, M_BIN = 8, M_DENS = 16 };

struct ObjArgs {
  const int32_t* node_cmat;      // [(nx+2)(ny+2)(nz+2)] -> cmat index or -1
  double* tmp;                   // temporary charge map of this species
  unsigned long long* absorbed;  // absorbed superparticles of this species
  unsigned long long* lost;      // of those, the ones whose entry face was not found
};

template <class R>
struct TileArgs {
  int N[3];        // cells per axis
  int T[3];        // tiles per axis
  int cx, cy, cz;  // node array dims with glue
  double xmax[3];  // largest valid coordinate, n - 1 (slx - dx)
  double qm, q, bx, by, bz, boris, size;
  // the same constants in the particles' precision (the push and the face
  // test read these: FP32 kernels then hold them as single scalars)
  R rqm, rb[3], rboris, rxmax[3];
  const double* ex;  // nodal E: component c at ex + c * e_s
  int64_t e_s;
  SoaPtr<R> reg;
  SoaPtr<R> dst;   // regions the binning pass writes (the other buffer of a double-buffered set, or reg)
  const int64_t* start;
  const int32_t* capa;
  int32_t* cnt;
  SoaPtr<R> out;
  uint8_t* odir;
  const int64_t* ostart;
  const int32_t* ocapa;
  int32_t* ocnt;
  const double* in;  // staged new particles: component c at in + c * in_s
  int64_t in_s;
  const int32_t* ibeg;
  SoaPtr<R> stray;
  int64_t stray_cap;
  int32_t* status;  // kStat words of this species (fopis_internal.hpp)
  double* rho_part;
  int32_t* dens_part;
  double* rho_g;
  double* dens_g;
  const uint8_t* mask;
  const uint8_t* tile_near;  // [ntiles] a conductor cell within one tile of the tile
  int nobj;
  ObjArgs obj[FOPIS_MAX_OBJECTS];
  double* cap_rec;  // captured particles: 8 doubles each (state, object)
  int64_t cap_cap;
  R* sink;  // kSink elements nobody reads (stores of lanes with nothing to keep)
};

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ int fdiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

struct DPos {  // include/position.hpp in FP64
  int i, j, k;
  double x, y, z, dx1, dy1, dz1, dx2, dy2, dz2;
};

__device__ __forceinline__ void set_xyz(DPos& p, double x, double y, double z) {
  p.x = x;
  p.y = y;
  p.z = z;
  p.i = (int)floor(x) + 1;
  p.j = (int)floor(y) + 1;
  p.k = (int)floor(z) + 1;
  p.dx1 = x - (double)(p.i - 1);
  p.dy1 = y - (double)(p.j - 1);
  p.dz1 = z - (double)(p.k - 1);
  p.dx2 = 1.0 - p.dx1;
  p.dy2 = 1.0 - p.dy1;
  p.dz2 = 1.0 - p.dz1;
}

// Conductor geometry of one object for the capture path (copied by value so
// the kernel arguments never need an address).
struct Cond {
  int cx, cy, cz, o;
  double q;
  const uint8_t* mask;
  const int32_t* node_cmat;
  double* tmp;
  __device__ bool valid_cell(int i, int j, int k) const {
    return i >= 0 && i < cx - 1 && j >= 0 && j < cy - 1 && k >= 0 && k < cz - 1;
  }
  __device__ bool cell(int i, int j, int k) const {
    if (!valid_cell(i, j, k)) return false;
    return mask[((size_t)i * (cy - 1) + j) * (cz - 1) + k] == (uint8_t)(o + 1);
  }
  __device__ int cmat(int i, int j, int k) const {
    if (i < 0 || j < 0 || k < 0 || i >= cx || j >= cy || k >= cz) return -1;
    return node_cmat[((size_t)i * cy + j) * cz + k];
  }
  // Spacecraft.cpp:461-504
  __device__ bool xpoint(const DPos& p, int sign) const {
    if (p.dx1 >= 1e-10) return false;
    if (valid_cell(p.i, p.j, p.k) && valid_cell(p.i - 1, p.j, p.k))
      return sign > 0 ? (cell(p.i, p.j, p.k) && !cell(p.i - 1, p.j, p.k))
                      : (!cell(p.i, p.j, p.k) && cell(p.i - 1, p.j, p.k));
    return cmat(p.i, p.j, p.k) >= 0 && cmat(p.i, p.j + 1, p.k) >= 0 && cmat(p.i, p.j, p.k + 1) >= 0 &&
           cmat(p.i, p.j + 1, p.k + 1) >= 0;
  }
  __device__ bool ypoint(const DPos& p, int sign) const {
    if (p.dy1 >= 1e-10) return false;
    if (valid_cell(p.i, p.j, p.k) && valid_cell(p.i, p.j - 1, p.k))
      return sign > 0 ? (cell(p.i, p.j, p.k) && !cell(p.i, p.j - 1, p.k))
                      : (!cell(p.i, p.j, p.k) && cell(p.i, p.j - 1, p.k));
    return cmat(p.i, p.j, p.k) >= 0 && cmat(p.i + 1, p.j, p.k) >= 0 && cmat(p.i, p.j, p.k + 1) >= 0 &&
           cmat(p.i + 1, p.j, p.k + 1) >= 0;
  }
  __device__ bool zpoint(const DPos& p, int sign) const {
    if (p.dz1 >= 1e-10) return false;
    if (valid_cell(p.i, p.j, p.k) && valid_cell(p.i, p.j, p.k - 1))
      return sign > 0 ? (cell(p.i, p.j, p.k) && !cell(p.i, p.j, p.k - 1))
                      : (!cell(p.i, p.j, p.k) && cell(p.i, p.j, p.k - 1));
    return cmat(p.i, p.j, p.k) >= 0 && cmat(p.i + 1, p.j, p.k) >= 0 && cmat(p.i, p.j + 1, p.k) >= 0 &&
           cmat(p.i + 1, p.j + 1, p.k) >= 0;
  }
  __device__ void add(int c, double v) const {
    if (c >= 0) atomicAdd(&tmp[c], v);
  }
  // Spacecraft.cpp:594-613
  __device__ void addx(const DPos& p) const {
    add(cmat(p.i, p.j, p.k), q * p.dy2 * p.dz2);
    add(cmat(p.i, p.j + 1, p.k), q * p.dy1 * p.dz2);
    add(cmat(p.i, p.j, p.k + 1), q * p.dy2 * p.dz1);
    add(cmat(p.i, p.j + 1, p.k + 1), q * p.dy1 * p.dz1);
  }
  __device__ void addy(const DPos& p) const {
    add(cmat(p.i, p.j, p.k), q * p.dx2 * p.dz2);
    add(cmat(p.i + 1, p.j, p.k), q * p.dx1 * p.dz2);
    add(cmat(p.i, p.j, p.k + 1), q * p.dx2 * p.dz1);
    add(cmat(p.i + 1, p.j, p.k + 1), q * p.dx1 * p.dz1);
  }
  __device__ void addz(const DPos& p) const {
    add(cmat(p.i, p.j, p.k), q * p.dx2 * p.dy2);
    add(cmat(p.i + 1, p.j, p.k), q * p.dx1 * p.dy2);
    add(cmat(p.i, p.j + 1, p.k), q * p.dx2 * p.dy1);
    add(cmat(p.i + 1, p.j + 1, p.k), q * p.dx1 * p.dy1);
  }
};

__device__ __forceinline__ double move_ratio(double d1, double d2, double v) {
  if (v == 0.0) return 0.0;
  return v > 0.0 ? d1 / fabs(v) : d2 / fabs(v);
}

// Spacecraft::distributeInnerParticleCharge with Particle::computeCrossPoints
// (particle.cpp:82-204): the charge goes to the four cmat nodes of the first
// crossed face whose orientation matches the velocity sign. Returns whether
// such a face was found.
__device__ bool absorb(const Cond& cd, double x, double y, double z, double vx, double vy, double vz) {
  DPos pos, old;
  set_xyz(pos, x, y, z);
  set_xyz(old, x - vx, y - vy, z - vz);
  const double mvx = 1.0 - move_ratio(pos.dx1, pos.dx2, vx);
  const double mvy = 1.0 - move_ratio(pos.dy1, pos.dy2, vy);
  const double mvz = 1.0 - move_ratio(pos.dz1, pos.dz2, vz);
  const bool mx = pos.i != old.i, my = pos.j != old.j, mz = pos.k != old.k;
  DPos xc{}, yc{}, zc{};
  if (mx) set_xyz(xc, vx > 0.0 ? ceil(old.x) : floor(old.x), old.y + vy * mvx, old.z + vz * mvx);
  if (my) set_xyz(yc, old.x + vx * mvy, vy > 0.0 ? ceil(old.y) : floor(old.y), old.z + vz * mvy);
  if (mz) set_xyz(zc, old.x + vx * mvz, old.y + vy * mvz, vz > 0.0 ? ceil(old.z) : floor(old.z));
  // order codes: 0 = x, 1 = y, 2 = z
  int ord[3];
  int n = 0;
  const int count = (int)mx + (int)my + (int)mz;
  if (count == 3) {
    if (mvx <= mvy) {
      if (mvx <= mvz) {
        ord[0] = 0;
        if (mvy <= mvz) {
          ord[1] = 1;
          ord[2] = 2;
        } else {
          ord[1] = 2;
          ord[2] = 1;
        }
      } else {
        ord[0] = 2;
        ord[1] = 0;
        ord[2] = 1;
      }
    } else {
      if (mvy <= mvz) {
        ord[0] = 1;
        if (mvx <= mvz) {
          ord[1] = 0;
          ord[2] = 2;
        } else {
          ord[1] = 2;
          ord[2] = 0;
        }
      } else {
        ord[0] = 2;
        ord[1] = 1;
        ord[2] = 0;
      }
    }
    n = 3;
  } else if (count == 2) {
    if (mx) {
      if (my) {
        ord[0] = mvx <= mvy ? 0 : 1;
        ord[1] = mvx <= mvy ? 1 : 0;
      } else {
        ord[0] = 0;  // particle.cpp:175 compares mvx with itself
        ord[1] = 2;
      }
    } else {
      ord[0] = mvy <= mvz ? 1 : 2;
      ord[1] = mvy <= mvz ? 2 : 1;
    }
    n = 2;
  } else if (count == 1) {
    ord[0] = mx ? 0 : (my ? 1 : 2);
    n = 1;
  }
  const int is = vx > 0.0 ? 1 : -1, js = vy > 0.0 ? 1 : -1, ks = vz > 0.0 ? 1 : -1;
  bool found = false;
  for (int q = 0; q < n && !found; ++q) {
    const DPos& c = ord[q] == 0 ? xc : (ord[q] == 1 ? yc : zc);
    if (cd.xpoint(c, is)) {
      cd.addx(c);
      found = true;
    } else if (cd.ypoint(c, js)) {
      cd.addy(c);
      found = true;
    } else if (cd.zpoint(c, ks)) {
      cd.addz(c);
      found = true;
    }
  }
  return found;
}

// Surface charge of the particles captured in this step (one thread each).
// The reference counts every captured particle in its object's current and
// reports those whose face it cannot find ("Surface cannot detect",
// Spacecraft.cpp:573-589); both counts are added once per wavefront and
// object (an atomic per particle queued on the one counter).
__global__ void k_capture(const double* rec, int n, int cx, int cy, int cz, double q, const uint8_t* mask,
                          ObjArgs o0, ObjArgs o1, ObjArgs o2, ObjArgs o3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < n;
  int o = -1;
  bool found = false;
  if (act) {
    const double* r = rec + (size_t)8 * i;
    o = (int)r[6];
    const ObjArgs oa = o == 0 ? o0 : (o == 1 ? o1 : (o == 2 ? o2 : o3));
    Cond cd{cx, cy, cz, o, q, mask, oa.node_cmat, oa.tmp};
    found = absorb(cd, r[0], r[1], r[2], r[3], r[4], r[5]);
  }
  unsigned long long todo = __ballot(act);
  while (todo) {
    const int ol = __builtin_amdgcn_readlane(o, __ffsll((long long)todo) - 1);
    const unsigned long long g = __ballot(act && o == ol) & todo;
    const unsigned long long gl = __ballot(act && o == ol && !found) & todo;
    if ((threadIdx.x & 63) == __ffsll((long long)g) - 1) {
      const ObjArgs oa = ol == 0 ? o0 : (ol == 1 ? o1 : (ol == 2 ? o2 : o3));
      atomicAdd(oa.absorbed, (unsigned long long)__popcll(g));
      if (gl) atomicAdd(oa.lost, (unsigned long long)__popcll(gl));
    }
    todo &= ~g;
  }
}

constexpr int kWaves = kBlock / 64;  // wavefronts per block

// A uniform pointer as a value held in scalar registers, in the global address
// space. Selecting among such values keeps a per-lane choice of destination
// array a register select (a select among the kernel-argument fields
// themselves can be compiled as a per-lane load of the chosen field), and
// stores through it are global stores (a generic pointer's would be flat
// stores, which also count against the LDS wait counter).
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* sgpr_ptr(T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<__attribute__((address_space(1))) T*>(((uint64_t)hi << 32) | lo);
}

// Ranks of two predicates over the block, in thread order.
__device__ __forceinline__ void block_rank2(bool pa, bool pb, int& ra, int& rb, int& ta, int& tb, int* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const unsigned long long ma = __ballot(pa), mb = __ballot(pb);
  if (lane == 0) {
    sh[w] = __popcll(ma);
    sh[kWaves + w] = __popcll(mb);
  }
  __syncthreads();
  int oa = 0, ob = 0;
  ta = 0;
  tb = 0;
#pragma unroll
  for (int q = 0; q < kWaves; ++q) {
    const int ca = sh[q], cb = sh[kWaves + q];
    if (q < w) {
      oa += ca;
      ob += cb;
    }
    ta += ca;
    tb += cb;
  }
  ra = oa + __popcll(ma & lt);
  rb = ob + __popcll(mb & lt);
  __syncthreads();
}

// Ranks of four predicates over the block: a0/b0 of every thread first, then
// a1/b1 (offsets of the second pair start after the first pair's totals are
// added by the caller).
__device__ __forceinline__ void block_rank4(bool a0, bool b0, bool a1, bool b1, int* ra, int* rb, int* ta, int* tb,
                                            int* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const unsigned long long m[4] = {__ballot(a0), __ballot(b0), __ballot(a1), __ballot(b1)};
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) sh[kWaves * k + w] = __popcll(m[k]);
  __syncthreads();
  int off[4] = {0, 0, 0, 0}, tot[4] = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < kWaves; ++q)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int cnt = sh[kWaves * k + q];
      if (q < w) off[k] += cnt;
      tot[k] += cnt;
    }
  ra[0] = off[0] + __popcll(m[0] & lt);
  rb[0] = off[1] + __popcll(m[1] & lt);
  ra[1] = off[2] + __popcll(m[2] & lt);
  rb[1] = off[3] + __popcll(m[3] & lt);
  ta[0] = tot[0];
  tb[0] = tot[1];
  ta[1] = tot[2];
  tb[1] = tot[3];
  __syncthreads();
}

template <class R>
__device__ __forceinline__ void put_stray(const TileArgs<R>& a, const R v[6]) {
  const int slot = atomicAdd(&a.status[0], 1);
  if (slot < a.stray_cap) {
#pragma unroll
    for (int c = 0; c < 6; ++c) a.stray.c[c][slot] = v[c];
  } else {
    atomicAdd(&a.status[1], 1);
  }
}

template <class R>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

template <class R, int MODE>
struct TileSmem {
  static constexpr bool kE = (MODE & M_VEL) != 0;
  static constexpr bool kR = (MODE & M_DEP) != 0;
  static constexpr bool kD = (MODE & (M_DEP | M_DENS)) != 0;
  // every segment starts on a 16-byte boundary: the FP64 charge window must be
  // 8-byte aligned in LDS whatever the size of the E window before it
  static constexpr size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }
  // E window interleaved per node as (x, y, z, unused): one 16-byte LDS read per corner (FP32)
  static constexpr size_t e_bytes = kE ? align16(sizeof(R) * 4 * kEW * kEW * kEW) : 0;
  // charge windows accumulate 64-bit fixed-point weights for both particle
  // precisions (ds_add_u64: as fast as FP64 LDS atomics and order-independent;
  // FP32 LDS atomics are 3x slower on gfx950)
  static constexpr size_t r_bytes = kR ? align16(sizeof(unsigned long long) * kRW * kRW * kRW) : 0;
  static constexpr size_t bytes = e_bytes + r_bytes + 16;
  static_assert(e_bytes % 16 == 0 && r_bytes % 16 == 0, "LDS segments must stay 16-byte aligned");
};

// Fixed-point scale of the charge windows: 2^44 units per unit CIC weight. A
// node of a tile window sums at most 2^19 particles' worth of weight before
// the signed 64-bit sum could overflow (the mean is 8 cells x particles per
// cell); a tile that processes kFixMaxParticles or more particles in one
// launch deposits through the global FP64 path instead. The quantum 2^-44 is
// 5.7e-14 of a particle's charge.
constexpr double kFix = 17592186044416.0;
constexpr int kFixMaxParticles = 1 << 19;
// FP32 particles use 2^31 units per unit weight: a weight's scaled value fits
// 32 bits, and a node's signed 64-bit sum holds 2^32 particles' worth (more
// than a tile launch can see), so FP32 tiles always deposit in LDS. The
// quantum 2^-31 is below the FP32 weights' own rounding.
constexpr float kFix32 = 2147483648.0f;

// round(v * kFix) for a CIC weight v in [0, 1] (the value of an exact scaling
// rounded to the nearest integer, ties to even). v * kFix + 2^52 lies in
// [2^52, 2^53), where a double's unit in the last place is 1: the one rounding
// of the fused multiply-add puts the rounded integer in the low 52 bits, and
// masking off the exponent leaves it. Two instructions instead of a general
// double-to-64-bit-integer conversion.
__device__ __forceinline__ unsigned long long fix_weight(double v) {
  const double m = __builtin_fma(v, kFix, 4503599627370496.0);
  return (unsigned long long)__double_as_longlong(m) & 0x000FFFFFFFFFFFFFull;
}

// One workgroup per tile: stage the tile's E nodes in LDS, stream its
// particles (and the staged new ones) in chunks, apply the requested stages,
// accumulate charge and counts in LDS windows, and re-bin: particles staying
// in the tile are compacted in place, the others go to the tile's outbox.
// DB: the binning pass writes the tile's particles to the other buffer
// (a.dst) and every wavefront reserves its slots with one LDS atomic per
// iteration, so the waves of a block never wait for each other until the
// end; otherwise particles are compacted in place (a.dst == a.reg), which
// needs a block-wide scan with two barriers per iteration so that no slot is
// written before it has been read.
template <class R, int MODE, bool DB = false>
__global__ __launch_bounds__(kBlock) void k_tile(TileArgs<R> a) {
  using SM = TileSmem<R, MODE>;
  __shared__ __attribute__((aligned(16))) char smem[SM::bytes];
  __shared__ int sh_rank[4 * kWaves];
  __shared__ int s_res[2];  // DB: slots reserved so far in the region and in the outbox
  // where the binning pass stores a particle of each kind (0 tile region, 1
  // outbox, 2 stray list, 3 sink): the set's base address and its component
  // stride in bytes; one 16-byte LDS read per particle replaces a four-way
  // choice among eight 64-bit scalars
  __shared__ __attribute__((aligned(16))) unsigned long long s_dst[(MODE & M_BIN) != 0 ? 4 : 1][2];
  R* sE = reinterpret_cast<R*>(smem);
  unsigned long long* sR = reinterpret_cast<unsigned long long*>(smem + SM::e_bytes);
  // the count window is its own LDS array so that its atomics stay LDS atomics (ds_add_u32)
  __shared__ int sD[SM::kD ? kDW * kDW * kDW : 1];

  const int t = blockIdx.x;
  const int tz = t % a.T[2], ty = (t / a.T[2]) % a.T[1], tx = t / (a.T[2] * a.T[1]);
  const int o0 = tx * kTile, o1 = ty * kTile, o2 = tz * kTile;
  const int ok[3] = {o0, o1, o2};
  // per axis: the tile's origin minus one, and the cells from the origin to the grid's end
  const int om1[3] = {o0 - 1, o1 - 1, o2 - 1};
  const int nmo[3] = {a.N[0] - o0, a.N[1] - o1, a.N[2] - o2};
  // FP32 particles are stored relative to their tile's origin (a coordinate
  // then resolves 2^-21 of a cell or better across the tile, where an absolute
  // one resolves 2^-16 near x = 255); FP64 particles keep absolute
  // coordinates. Per axis, in stored coordinates: wo turns the floor of a
  // coordinate into the cell counted from the tile's origin minus one, cl is
  // the floor of the last cell, lo and hi bound the domain.
  constexpr bool kLocal = sizeof(R) == 4;
  int wo[3];
  R cl[3], lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    wo[k] = kLocal ? -1 : om1[k];
    cl[k] = (R)(kLocal ? nmo[k] - 1 : a.N[k] - 1);
    lo[k] = kLocal ? (R)(-ok[k]) : (R)0;
    hi[k] = kLocal ? a.rxmax[k] - (R)ok[k] : a.rxmax[k];
  }

  if constexpr (SM::kE) {
    constexpr int ne = kEW * kEW * kEW;
    for (int q = threadIdx.x; q < ne; q += kBlock) {
      const int aa = q / (kEW * kEW), bb = (q / kEW) % kEW, cc = q % kEW;
      const int gi = o0 + 1 + aa, gj = o1 + 1 + bb, gk = o2 + 1 + cc;
      R vx = 0, vy = 0, vz = 0;
      if (gi < a.cx && gj < a.cy && gk < a.cz) {
        const size_t g = ((size_t)gi * a.cy + gj) * a.cz + gk;
        vx = (R)a.ex[g];
        vy = (R)a.ex[g + a.e_s];
        vz = (R)a.ex[g + 2 * a.e_s];
      }
      sE[4 * q] = vx;
      sE[4 * q + 1] = vy;
      sE[4 * q + 2] = vz;
      sE[4 * q + 3] = (R)0;
    }
  }
  if constexpr (SM::kR) {
    for (int q = threadIdx.x; q < kRW * kRW * kRW; q += kBlock) sR[q] = 0ull;
  }
  if constexpr (SM::kD) {
    for (int q = threadIdx.x; q < kDW * kDW * kDW; q += kBlock) sD[q] = 0;
  }
  if constexpr ((MODE & M_BIN) != 0) {
    if (threadIdx.x < 2) s_res[threadIdx.x] = 0;
    if (threadIdx.x < 4) {
      const int k = threadIdx.x;
      const SoaPtr<R>& set = k == 0 ? a.dst : (k == 1 ? a.out : a.stray);
      const uint64_t snk = reinterpret_cast<uint64_t>(a.sink);
      s_dst[k][0] = k == 3 ? snk : reinterpret_cast<uint64_t>(set.c[0]);
      s_dst[k][1] = k == 3 ? 0ull : reinterpret_cast<uint64_t>(set.c[1]) - reinterpret_cast<uint64_t>(set.c[0]);
    }
  }
  __syncthreads();

  const int64_t base = a.start[t];
  const int nreg = a.cnt[t];
  int nin = 0, ib0 = 0;
  if constexpr ((MODE & (M_DEP | M_DENS)) != 0) {
    if (a.ibeg) {
      ib0 = a.ibeg[t];
      nin = a.ibeg[t + 1] - ib0;
    }
  }
  const int ntot = nreg + nin;
  if constexpr ((MODE & M_BIN) != 0) {
    if (threadIdx.x == 0 && nin > 0) atomicAdd(&a.status[5], nin);  // staged particles joining the regions
  }
  // every weight this tile's window sums comes from these ntot particles
  const bool fix_ok = ntot < kFixMaxParticles;
  const bool near_obj = a.nobj > 0 && a.tile_near[t] != 0;
  int wptr = 0, optr = 0;
  const int ocap = ((MODE & M_BIN) != 0) ? a.ocapa[t] : 0;
  const int64_t ost = ((MODE & M_BIN) != 0) ? a.ostart[t] : 0;
  const int rcap = a.capa[t];
  const int64_t sink_at = ((int64_t)t * kBlock + threadIdx.x) & (kSink - 1);
  const uint64_t odir_b = reinterpret_cast<uint64_t>(a.odir), snk_b = reinterpret_cast<uint64_t>(a.sink);

  // PPT particles per thread and iteration (FP32: two chunks of kBlock, whose
  // loads are in flight together; one block scan ranks both), the next
  // iteration's particles loaded while the current ones are processed
  constexpr int PPT = sizeof(R) == 4 ? 2 : 1;
  // Every lane loads from the tile region (a lane past the region's particles
  // reads its first slot, which exists: a region holds at least 64), so each
  // iteration issues the same loads whatever its lanes hold.
  // the tile region's first element of each component as a uniform address
  // and a 32-bit element index: loads take the scalar base plus a vector offset
  const R* const rb = a.reg.c[0] + base;
  const int64_t rs = a.reg.c[1] - a.reg.c[0];
  const __attribute__((address_space(1))) R* rbc[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) rbc[c] = sgpr_ptr(rb + c * rs);
  auto fetch = [&](int idx, R* q) {
    // a 32-bit byte offset (a tile region holds far fewer than 2^29 particles),
    // so that each load is the scalar base plus one vector offset register
    const uint32_t jb = (idx < nreg ? (uint32_t)idx : 0u) * (uint32_t)sizeof(R);
#pragma unroll
    for (int c = 0; c < 6; ++c)
      q[c] = *reinterpret_cast<const __attribute__((address_space(1))) R*>(
          reinterpret_cast<const __attribute__((address_space(1))) char*>(rbc[c]) + jb);
    if (idx >= nreg) {
      if (idx < ntot) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
          q[c] = (R)(a.in[c * a.in_s + ib0 + idx - nreg] - (kLocal && c < 3 ? (double)ok[c] : 0.0));
      } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) q[c] = (R)0;
      }
    }
  };
  R nxt[PPT][6];
#pragma unroll
  for (int u = 0; u < PPT; ++u) fetch(u * kBlock + threadIdx.x, nxt[u]);

  // one particle through velocity, position, cell, density, capture and
  // deposit; returns whether it is still valid (and in dt the tile it bins
  // into, as steps 0, 1, 2 per axis from the neighbour below)
  auto work = [&](R (&p)[6], int idx, int (&dt)[3], bool& lost) -> bool {
    const bool act = idx < ntot;
    const bool from_in = idx >= nreg;
    bool valid = act && !isnan(p[0]);

      // ---- velocity: trilinear gather of nodal E, Boris rotation with t = b
      if constexpr ((MODE & M_VEL) != 0 && sizeof(R) == 4) {
        // FP32 particles: the same gather and rotation with fused multiply-adds
        // and factored weights (FP32 makes no claim of bit equality with the
        // reference, so the operation order is free)
        if (valid && !from_in) {
          const float fx = fminf(floorf(p[0]), cl[0]);
          const float fy = fminf(floorf(p[1]), cl[1]);
          const float fz = fminf(floorf(p[2]), cl[2]);
          // the cell counted from the tile's origin (the coordinates are)
          const int li = (int)fx, lj = (int)fy, lk = (int)fz;
          const float dx1 = p[0] - fx, dy1 = p[1] - fy, dz1 = p[2] - fz;
          const float dx2 = 1.0f - dx1, dy2 = 1.0f - dy1, dz2 = 1.0f - dz1;
          const float qm = a.rqm;
          const float qz1 = qm * dz1, qz2 = qm * dz2;
          const float a22 = dx2 * dy2, a12 = dx1 * dy2, a21 = dx2 * dy1, a11 = dx1 * dy1;
          const float v1 = a22 * qz2, v2 = a12 * qz2, v3 = a21 * qz2, v4 = a11 * qz2;
          const float v5 = a22 * qz1, v6 = a12 * qz1, v7 = a21 * qz1, v8 = a11 * qz1;
          float e[3];
          if (li >= 0 && li < kTile && lj >= 0 && lj < kTile && lk >= 0 && lk < kTile) {
            const int b0 = (li * kEW + lj) * kEW + lk;
            // volatile keeps each corner one 16-byte read (ds_read_b128, pad
            // included) rather than the 12 bytes used (ds_read_b96), which the
            // LDS serves more slowly on gfx950: 4% of this kernel at C3
            typedef float f4v __attribute__((ext_vector_type(4)));
            const volatile __attribute__((address_space(3))) f4v* E4 =
                (const volatile __attribute__((address_space(3))) f4v*)sE + b0;
            const f4v n1 = E4[0], n2 = E4[kEW * kEW], n3 = E4[kEW], n4 = E4[kEW * kEW + kEW];
            const f4v n5 = E4[1], n6 = E4[kEW * kEW + 1], n7 = E4[kEW + 1], n8 = E4[kEW * kEW + kEW + 1];
            e[0] = __builtin_fmaf(v8, n8.x, __builtin_fmaf(v7, n7.x, __builtin_fmaf(v6, n6.x, __builtin_fmaf(v5, n5.x,
                   __builtin_fmaf(v4, n4.x, __builtin_fmaf(v3, n3.x, __builtin_fmaf(v2, n2.x, v1 * n1.x)))))));
            e[1] = __builtin_fmaf(v8, n8.y, __builtin_fmaf(v7, n7.y, __builtin_fmaf(v6, n6.y, __builtin_fmaf(v5, n5.y,
                   __builtin_fmaf(v4, n4.y, __builtin_fmaf(v3, n3.y, __builtin_fmaf(v2, n2.y, v1 * n1.y)))))));
            e[2] = __builtin_fmaf(v8, n8.z, __builtin_fmaf(v7, n7.z, __builtin_fmaf(v6, n6.z, __builtin_fmaf(v5, n5.z,
                   __builtin_fmaf(v4, n4.z, __builtin_fmaf(v3, n3.z, __builtin_fmaf(v2, n2.z, v1 * n1.z)))))));
          } else {  // outside the tile (a bin left stale by an interrupted step): global gather
            const int ci = li + o0, cj = lj + o1, ck = lk + o2;
            const size_t sI = (size_t)a.cy * a.cz, sJ = a.cz;
            const size_t g = (size_t)(ci + 1) * sI + (size_t)(cj + 1) * sJ + (ck + 1);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const double* E = a.ex + c * a.e_s + g;
              e[c] = v1 * (float)E[0] + v2 * (float)E[sI] + v3 * (float)E[sJ] + v4 * (float)E[sI + sJ] +
                     v5 * (float)E[1] + v6 * (float)E[sI + 1] + v7 * (float)E[sJ + 1] + v8 * (float)E[sI + sJ + 1];
            }
          }
          const float bx = a.rb[0], by = a.rb[1], bz = a.rb[2], boris = a.rboris;
          const float vx1 = p[3] + e[0], vy1 = p[4] + e[1], vz1 = p[5] + e[2];
          const float vxt = __builtin_fmaf(vy1, bz, __builtin_fmaf(-vz1, by, vx1));
          const float vyt = __builtin_fmaf(vz1, bx, __builtin_fmaf(-vx1, bz, vy1));
          const float vzt = __builtin_fmaf(vx1, by, __builtin_fmaf(-vy1, bx, vz1));
          p[3] = vx1 + __builtin_fmaf(boris, __builtin_fmaf(vyt, bz, -vzt * by), e[0]);
          p[4] = vy1 + __builtin_fmaf(boris, __builtin_fmaf(vzt, bx, -vxt * bz), e[1]);
          p[5] = vz1 + __builtin_fmaf(boris, __builtin_fmaf(vxt, by, -vyt * bx), e[2]);
        }
      } else if constexpr ((MODE & M_VEL) != 0) {
        if (valid && !from_in) {
          const R x = p[0], y = p[1], z = p[2];
          int ci = (int)floor(x), cj = (int)floor(y), ck = (int)floor(z);
          ci = min(ci, a.N[0] - 1);
          cj = min(cj, a.N[1] - 1);
          ck = min(ck, a.N[2] - 1);
          const R dx1 = x - (R)ci, dy1 = y - (R)cj, dz1 = z - (R)ck;
          const R dx2 = (R)1 - dx1, dy2 = (R)1 - dy1, dz2 = (R)1 - dz1;
          const R qm = a.rqm;
          const R v1 = qm * dx2 * dy2 * dz2;
          const R v2 = qm * dx1 * dy2 * dz2;
          const R v3 = qm * dx2 * dy1 * dz2;
          const R v4 = qm * dx1 * dy1 * dz2;
          const R v5 = qm * dx2 * dy2 * dz1;
          const R v6 = qm * dx1 * dy2 * dz1;
          const R v7 = qm * dx2 * dy1 * dz1;
          const R v8 = qm * dx1 * dy1 * dz1;
          const int li = ci - o0, lj = cj - o1, lk = ck - o2;
          R e[3];
          if (li >= 0 && li < kTile && lj >= 0 && lj < kTile && lk >= 0 && lk < kTile) {
            const int b0 = (li * kEW + lj) * kEW + lk;
            // the eight corner nodes, each (x, y, z) in one vector read; same sums as the reference order
            using V4 = typename Vec4<R>::type;
            const V4* E4 = reinterpret_cast<const V4*>(sE) + b0;
            const V4 n1 = E4[0], n2 = E4[kEW * kEW], n3 = E4[kEW], n4 = E4[kEW * kEW + kEW];
            const V4 n5 = E4[1], n6 = E4[kEW * kEW + 1], n7 = E4[kEW + 1], n8 = E4[kEW * kEW + kEW + 1];
            e[0] = v1 * n1.x + v2 * n2.x + v3 * n3.x + v4 * n4.x + v5 * n5.x + v6 * n6.x + v7 * n7.x + v8 * n8.x;
            e[1] = v1 * n1.y + v2 * n2.y + v3 * n3.y + v4 * n4.y + v5 * n5.y + v6 * n6.y + v7 * n7.y + v8 * n8.y;
            e[2] = v1 * n1.z + v2 * n2.z + v3 * n3.z + v4 * n4.z + v5 * n5.z + v6 * n6.z + v7 * n7.z + v8 * n8.z;
          } else {  // outside the tile (a bin left stale by an interrupted step): global gather
            const size_t sI = (size_t)a.cy * a.cz, sJ = a.cz;
            const size_t g = (size_t)(ci + 1) * sI + (size_t)(cj + 1) * sJ + (ck + 1);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const double* E = a.ex + c * a.e_s + g;
              e[c] = v1 * (R)E[0] + v2 * (R)E[sI] + v3 * (R)E[sJ] + v4 * (R)E[sI + sJ] + v5 * (R)E[1] +
                     v6 * (R)E[sI + 1] + v7 * (R)E[sJ + 1] + v8 * (R)E[sI + sJ + 1];
            }
          }
          const R bx = a.rb[0], by = a.rb[1], bz = a.rb[2], boris = a.rboris;
          const R vx1 = p[3] + e[0], vy1 = p[4] + e[1], vz1 = p[5] + e[2];
          const R vxt = vx1 + vy1 * bz - vz1 * by;
          const R vyt = vy1 + vz1 * bx - vx1 * bz;
          const R vzt = vz1 + vx1 * by - vy1 * bx;
          p[3] = vx1 + (e[0] + boris * (vyt * bz - vzt * by)) * (R)1;
          p[4] = vy1 + (e[1] + boris * (vzt * bx - vxt * bz)) * (R)1;
          p[5] = vz1 + (e[2] + boris * (vxt * by - vyt * bx)) * (R)1;
        }
      }

      // ---- position: x += v; leaving a physical face invalidates
      if constexpr ((MODE & M_POS) != 0) {
        if (valid && !from_in) {
          p[0] = p[0] + p[3];
          p[1] = p[1] + p[4];
          p[2] = p[2] + p[5];
        }
      }
      // outside the domain (a pushed particle is absorbed there, a staged one is dropped)
      const bool out = (p[0] < lo[0]) | (p[0] > hi[0]) | (p[1] < lo[1]) | (p[1] > hi[1]) | (p[2] < lo[2]) |
                       (p[2] > hi[2]);
      if constexpr ((MODE & M_POS) != 0) {
        lost = valid & !from_in & out;
        valid = valid & !out;
      } else {
        valid = valid & !(from_in & out);
      }

      // the cell (floor of the position), and the same cell counted from the
      // tile's origin minus one (w); the deposit's cell d = min(cell, N - 1)
      // keeps the upper face's particles in the last cell (as the reference's
      // weights do), wd being w clamped the same way. Density window, charge
      // window and conductor test share one predicate: every axis of w in
      // [0, 10), the tile's cells and one more on each side (for a valid
      // particle, w in [0, 10) if and only if wd is).
      R fl[3];
      int w[3], wd[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        fl[k] = floor(p[k]);
        w[k] = (int)fl[k] - wo[k];
        wd[k] = min(w[k], nmo[k]);
      }
      const bool near = max(max((unsigned)w[0], (unsigned)w[1]), (unsigned)w[2]) < (unsigned)kDW;
      static_assert(kDW == kRW - 1, "the density and charge windows cover the same cells");
      // this tile's neighbour the particle bins into, as steps -1, 0, +1 per
      // axis plus one (floor division of the clamped cell by the tile width)
#pragma unroll
      for (int k = 0; k < 3; ++k) dt[k] = (wd[k] + kTile - 1) >> kTileShift;

      // ---- density: count per reference cell (RootGridField.cpp:1129-1150)
      if constexpr (SM::kD) {
        if (valid) {
          if (near) {
            atomicAdd(&sD[(w[0] * kDW + w[1]) * kDW + w[2]], 1);
          } else {
            const int ci = w[0] + o0 - 1, cj = w[1] + o1 - 1, ck = w[2] + o2 - 1;
            const size_t g = ((size_t)(ci + 1) * (a.cy - 1) + (cj + 1)) * (a.cz - 1) + (ck + 1);
            atomicAdd(&a.dens_g[g], a.size);
            atomicAdd(&a.status[2], 1);
          }
        }
      }

      // ---- conductor capture, then CIC deposit
      if constexpr ((MODE & M_DEP) != 0) {
        // a tile with no conductor cell within one tile of it reads no mask for
        // particles that end next to it
        if (valid && a.nobj > 0 && (near_obj || !near)) {
          const int mi = w[0] + o0, mj = w[1] + o1, mk = w[2] + o2;
          const uint8_t m = a.mask[((size_t)mi * (a.cy - 1) + mj) * (a.cz - 1) + mk];
          if (m != 0) {
            // captured: recorded for the capture kernel (surface charge on the device)
            const int slot = atomicAdd(&a.status[3], 1);
            if (slot < a.cap_cap) {
              double* r = a.cap_rec + (size_t)8 * slot;
#pragma unroll
              for (int c = 0; c < 6; ++c) r[c] = (double)p[c] + (kLocal && c < 3 ? (double)ok[c] : 0.0);
              r[6] = (double)(m - 1);
            } else {
              atomicAdd(&a.status[6], 1);
            }
            valid = false;
          }
        }
        if (valid) {
          const R dx1 = p[0] - fmin(fl[0], cl[0]), dy1 = p[1] - fmin(fl[1], cl[1]), dz1 = p[2] - fmin(fl[2], cl[2]);
          const R dx2 = (R)1 - dx1, dy2 = (R)1 - dy1, dz2 = (R)1 - dz1;
          // CIC weights of the eight nodes (the particle's charge is applied when
          // the window is written out)
          const R w8[8] = {dx2 * dy2 * dz2, dx1 * dy2 * dz2, dx2 * dy1 * dz2, dx1 * dy1 * dz2,
                           dx2 * dy2 * dz1, dx1 * dy2 * dz1, dx2 * dy1 * dz1, dx1 * dy1 * dz1};
          const int wi = wd[0], wj = wd[1], wk = wd[2];
          if (sizeof(R) == 4 && near) {
            // FP32 particles: kFix32 units per weight, the factored weights
            // scaled and rounded in one multiply-add each (+0.5, then the
            // conversion truncates), 32-bit values into the 64-bit window
            const float s1 = (float)dz1 * kFix32, s2 = (float)dz2 * kFix32;
            const float a22 = (float)(dx2 * dy2), a12 = (float)(dx1 * dy2), a21 = (float)(dx2 * dy1),
                        a11 = (float)(dx1 * dy1);
            unsigned long long* W = sR + (wi * kRW + wj) * kRW + wk;
            auto fx = [](float v) -> unsigned long long { return (unsigned long long)(unsigned)v; };
            atomicAdd(W, fx(__builtin_fmaf(a22, s2, 0.5f)));
            atomicAdd(W + kRW * kRW, fx(__builtin_fmaf(a12, s2, 0.5f)));
            atomicAdd(W + kRW, fx(__builtin_fmaf(a21, s2, 0.5f)));
            atomicAdd(W + kRW * kRW + kRW, fx(__builtin_fmaf(a11, s2, 0.5f)));
            atomicAdd(W + 1, fx(__builtin_fmaf(a22, s1, 0.5f)));
            atomicAdd(W + kRW * kRW + 1, fx(__builtin_fmaf(a12, s1, 0.5f)));
            atomicAdd(W + kRW + 1, fx(__builtin_fmaf(a21, s1, 0.5f)));
            atomicAdd(W + kRW * kRW + kRW + 1, fx(__builtin_fmaf(a11, s1, 0.5f)));
          } else if (sizeof(R) == 8 && fix_ok && near) {
            // 64-bit fixed point (kFix units per weight 1): integer sums do not
            // depend on the order the atomics land in, so the window, and rho,
            // are the same bit for bit from run to run
            unsigned long long* W = sR + (wi * kRW + wj) * kRW + wk;
            auto fx = [](R v) -> unsigned long long { return fix_weight((double)v); };
            atomicAdd(W, fx(w8[0]));
            atomicAdd(W + kRW * kRW, fx(w8[1]));
            atomicAdd(W + kRW, fx(w8[2]));
            atomicAdd(W + kRW * kRW + kRW, fx(w8[3]));
            atomicAdd(W + 1, fx(w8[4]));
            atomicAdd(W + kRW * kRW + 1, fx(w8[5]));
            atomicAdd(W + kRW + 1, fx(w8[6]));
            atomicAdd(W + kRW * kRW + kRW + 1, fx(w8[7]));
          } else {
            const int di = wi + o0 - 1, dj = wj + o1 - 1, dk = wk + o2 - 1;
            const double q = a.q;
            const size_t sI = (size_t)a.cy * a.cz, sJ = a.cz;
            double* G = a.rho_g + (size_t)(di + 1) * sI + (size_t)(dj + 1) * sJ + (dk + 1);
            atomicAdd(G, q * (double)w8[0]);
            atomicAdd(G + sI, q * (double)w8[1]);
            atomicAdd(G + sJ, q * (double)w8[2]);
            atomicAdd(G + sI + sJ, q * (double)w8[3]);
            atomicAdd(G + 1, q * (double)w8[4]);
            atomicAdd(G + sI + 1, q * (double)w8[5]);
            atomicAdd(G + sJ + 1, q * (double)w8[6]);
            atomicAdd(G + sI + sJ + 1, q * (double)w8[7]);
            atomicAdd(&a.status[2], 1);
          }
        }
      }

    return valid;
  };

  // Particles absorbed at a domain face: one count per wavefront, added to the
  // launch's total once at the tile's end. (An atomic per iteration stalled
  // the face tiles: all of them run at once at a launch's start, and their
  // wavefronts queued on the one counter.)
  int nlost = 0;
  for (int c0 = 0; c0 < ntot; c0 += PPT * kBlock) {
    R p[PPT][6];
#pragma unroll
    for (int u = 0; u < PPT; ++u)
#pragma unroll
      for (int c = 0; c < 6; ++c) p[u][c] = nxt[u][c];
#pragma unroll
    for (int u = 0; u < PPT; ++u) fetch(c0 + (PPT + u) * kBlock + threadIdx.x, nxt[u]);
    bool valid[PPT], lost[PPT];
    int dt[PPT][3];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      lost[u] = false;
      valid[u] = work(p[u], c0 + u * kBlock + threadIdx.x, dt[u], lost[u]);
    }
    if constexpr ((MODE & M_POS) != 0) {
      // particles absorbed at a domain face, counted per wavefront
#pragma unroll
      for (int u = 0; u < PPT; ++u) nlost += __popcll(__ballot(lost[u]));
    }

    if constexpr ((MODE & M_BIN) != 0) {
      int dir[PPT];
      bool far[PPT], stay[PPT], move[PPT];
#pragma unroll
      for (int u = 0; u < PPT; ++u) {
        dir[u] = 13;
        far[u] = false;
        if (valid[u]) {
          // one of the 27 tiles around this one (13: this tile), or farther
          if (max(max((unsigned)dt[u][0], (unsigned)dt[u][1]), (unsigned)dt[u][2]) > 2u)
            far[u] = true;
          else
            dir[u] = (dt[u][0] * 3 + dt[u][1]) * 3 + dt[u][2];
        }
        stay[u] = valid[u] && !far[u] && dir[u] == 13;
        move[u] = valid[u] && !far[u] && dir[u] != 13;
      }
      int rs[PPT], rm[PPT], ts[PPT], tm[PPT];
      if constexpr (DB) {
        // slots of this wavefront: its stayers and movers in (u, lane) order
        // after one reservation each in the region and the outbox; below(m)
        // counts the lanes of mask m below this one (v_mbcnt)
        auto below = [](unsigned long long m) -> int {
          return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        };
        const int lane = threadIdx.x & 63;
        unsigned long long ms[PPT], mm[PPT];
        int ns = 0, nm = 0;
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
          ms[u] = __ballot(stay[u]);
          mm[u] = __ballot(move[u]);
          rs[u] = ns + below(ms[u]);
          rm[u] = nm + below(mm[u]);
          ns += __popcll(ms[u]);
          nm += __popcll(mm[u]);
        }
        int bs = 0, bm = 0;
        if (lane == 0) {
          if (ns) bs = atomicAdd(&s_res[0], ns);
          if (nm) bm = atomicAdd(&s_res[1], nm);
        }
        bs = __builtin_amdgcn_readfirstlane(bs);
        bm = __builtin_amdgcn_readfirstlane(bm);
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
          rs[u] += bs;
          rm[u] += bm;
          ts[u] = 0;
          tm[u] = 0;
        }
      } else {
        // ranks in (u, thread) order; the scan's barriers also order all reads before the writes
        if constexpr (PPT == 1)
          block_rank2(stay[0], move[0], rs[0], rm[0], ts[0], tm[0], sh_rank);
        else
          block_rank4(stay[0], move[0], stay[1], move[1], rs, rm, ts, tm, sh_rank);
      }
      // Every lane issues the same seven stores per particle slot (a lane with
      // nothing to keep writes to the sink): with a fixed count of memory
      // operations per iteration, the next particles' loads are waited for
      // without waiting for this iteration's stores.
      int wb = DB ? 0 : wptr, ob0 = DB ? 0 : optr;
#pragma unroll
      for (int u = 0; u < PPT; ++u) {
        int kind = 3;  // 0 tile region, 1 outbox, 2 stray list, 3 sink
        int64_t at = sink_at;
        bool stray = far[u];
        if (stay[u]) {
          const int pos = wb + rs[u];
          if (pos < rcap) {
            kind = 0;
            at = base + pos;
          } else {
            stray = true;
          }
        } else if (move[u]) {
          const int pos = ob0 + rm[u];
          if (pos < ocap) {
            kind = 1;
            at = ost + pos;
          } else {
            stray = true;
          }
        }
        if (stray) {
          const int slot = atomicAdd(&a.status[0], 1);
          if (slot < a.stray_cap) {
            kind = 2;
            at = slot;
            if constexpr (kLocal) {  // the stray list holds absolute coordinates
#pragma unroll
              for (int k = 0; k < 3; ++k) p[u][k] += (R)ok[k];
            }
          } else {
            atomicAdd(&a.status[1], 1);
          }
        }
        // one base address and one component stride per destination: each
        // particle set is one allocation with its components a fixed stride
        // apart, and the sink takes all six components in one slot
        const uint64_t db = s_dst[kind][0], ds = s_dst[kind][1];
        uint64_t ad = db + (uint64_t)at * sizeof(R);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          *reinterpret_cast<__attribute__((address_space(1))) R*>(ad) = p[u][c];
          ad += ds;
        }
        // the direction byte: the outbox's for a mover, the sink's otherwise
        const uint64_t od = kind == 1 ? odir_b + (uint64_t)at : snk_b + (uint64_t)sink_at;
        *reinterpret_cast<__attribute__((address_space(1))) uint8_t*>(od) = (uint8_t)dir[u];
        wb += ts[u];
        ob0 += tm[u];
      }
      wptr = min(wb, rcap);
      optr = min(ob0, ocap);
    } else if constexpr ((MODE & (M_VEL | M_POS)) != 0) {
#pragma unroll
      for (int u = 0; u < PPT; ++u) {
        const int idx = c0 + u * kBlock + threadIdx.x;
        if (idx < ntot && idx < nreg) {
          if (!valid[u]) p[u][0] = (R)NAN;  // invalid until the next re-bin drops it
#pragma unroll
          for (int c = 0; c < 6; ++c) a.reg.c[c][base + idx] = p[u][c];
        }
      }
    }
  }

  if constexpr ((MODE & M_POS) != 0) {
    if ((threadIdx.x & 63) == 0 && nlost > 0) atomicAdd(&a.status[4], nlost);
  }
  if constexpr (DB) __syncthreads();  // every wavefront's reservations are in
  if constexpr ((MODE & M_BIN) != 0) {
    if (threadIdx.x == 0) {
      a.cnt[t] = DB ? min(s_res[0], rcap) : wptr;
      a.ocnt[t] = DB ? min(s_res[1], ocap) : optr;
    }
  }
  if constexpr (SM::kR || SM::kD || (MODE & M_BIN) != 0) __syncthreads();
  if constexpr (SM::kR) {
    double* P = a.rho_part + (size_t)t * (kRW * kRW * kRW);
    const double scale = a.q / (sizeof(R) == 4 ? (double)kFix32 : kFix);  // exact: powers of two
    for (int q = threadIdx.x; q < kRW * kRW * kRW; q += kBlock) P[q] = (double)(long long)sR[q] * scale;
  }
  if constexpr (SM::kD) {
    int32_t* P = a.dens_part + (size_t)t * (kDW * kDW * kDW);
    for (int q = threadIdx.x; q < kDW * kDW * kDW; q += kBlock) P[q] = sD[q];
  }
}

// Arrivals, from the sources: the workgroup of tile s sends every entry of
// its outbox to the tile in the entry's direction, after that tile's
// stayers. A wavefront reserves the slots of all its entries bound for one
// tile with one atomic on that tile's arrival count (one per distinct
// direction among its 64 entries, issued together), so each outbox is read once. Arrivals
// past a tile's capacity go to the stray list. k_arrive_done then adds the
// arrivals to the tiles' counts and clears them for the next step.
template <class R>
__global__ __launch_bounds__(kBlock) void k_scatter(TileArgs<R> a, int32_t* __restrict__ arr) {
  const int s = blockIdx.x;
  const int sz = s % a.T[2], sy = (s / a.T[2]) % a.T[1], sx = s / (a.T[2] * a.T[1]);
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int m = a.ocnt[s];
  const int64_t ob = a.ostart[s];
  for (int i0 = 0; i0 < m; i0 += kBlock) {
    const int i = i0 + threadIdx.x;
    const bool act = i < m;
    const int d = act ? (int)a.odir[ob + i] : 13;
    // the destination tile (the binning pass only files directions of existing neighbours)
    const int t = ((sx + d / 9 - 1) * a.T[1] + (sy + (d / 3) % 3 - 1)) * a.T[2] + (sz + d % 3 - 1);
    // the entry, loaded before the reservation so that its loads are in flight meanwhile
    R p[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) p[c] = act ? a.out.c[c][ob + i] : (R)0;
    const int ct = a.cnt[t], cp = a.capa[t];
    // grp: the lanes filing into the same direction as this lane (ballots
    // only, one per distinct direction); then the lowest lane of each group
    // reserves the group's slots, all groups' atomics in flight together
    unsigned long long grp = 0;
    unsigned long long todo = __ballot(act);
    while (todo) {
      const int dl = __builtin_amdgcn_readlane(d, __ffsll((long long)todo) - 1);
      const unsigned long long g = __ballot(act && d == dl) & todo;
      if (act && d == dl) grp = g;
      todo &= ~g;
    }
    int b = 0;
    if (act && (grp & lt) == 0) b = atomicAdd(&arr[t], __popcll(grp));
    b = __shfl(b, act ? __ffsll((long long)grp) - 1 : lane, 64);
    const int pos = b + __popcll(grp & lt);
    if (act) {
      if constexpr (sizeof(R) == 4) {
        // FP32 outbox entries are relative to the source tile's origin; the
        // destination's origin lies one tile width away per step of d
        p[0] += (R)(kTile * (1 - d / 9));
        p[1] += (R)(kTile * (1 - (d / 3) % 3));
        p[2] += (R)(kTile * (1 - d % 3));
      }
      const int at = ct + pos;
      if (at < cp) {
        const int64_t g = a.start[t] + at;
#pragma unroll
        for (int c = 0; c < 6; ++c) a.dst.c[c][g] = p[c];
      } else {
        if constexpr (sizeof(R) == 4) {  // the stray list holds absolute coordinates
          p[0] += (R)(kTile * (sx + d / 9 - 1));
          p[1] += (R)(kTile * (sy + (d / 3) % 3 - 1));
          p[2] += (R)(kTile * (sz + d % 3 - 1));
        }
        put_stray(a, p);
      }
    }
  }
}

__global__ void k_arrive_done(int32_t* __restrict__ cnt, const int32_t* __restrict__ capa, int32_t* __restrict__ arr,
                              int nt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  cnt[t] = min(cnt[t] + arr[t], capa[t]);
  arr[t] = 0;
}

// ------------------------------------------------------------- re-binning

template <class R>
__global__ void k_keys_tile(SoaPtr<R> f, int64_t n, TileArgs<R> a, uint32_t* keys, int with_cell, uint32_t invalid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const R x = f.c[0][i], y = f.c[1][i], z = f.c[2][i];
  uint32_t key = invalid;
  const bool ok = !isnan(x) && !isnan(y) && !isnan(z) && x >= (R)0 && y >= (R)0 && z >= (R)0 &&
                  x <= (R)a.xmax[0] && y <= (R)a.xmax[1] && z <= (R)a.xmax[2];
  if (ok) {
    const int ci = min((int)floor(x), a.N[0] - 1), cj = min((int)floor(y), a.N[1] - 1),
              ck = min((int)floor(z), a.N[2] - 1);
    const uint32_t tile = (uint32_t)(((ci / kTile) * a.T[1] + (cj / kTile)) * a.T[2] + (ck / kTile));
    if (with_cell) {
      const uint32_t lc = (uint32_t)(((ci % kTile) * kTile + (cj % kTile)) * kTile + (ck % kTile));
      key = tile * (kTile * kTile * kTile) + lc;
    } else {
      key = tile;
    }
  }
  keys[i] = key;
}

// Tile starts in a sorted key array: first[u] = index of the first key whose
// tile (key >> shift, capped at ntiles for invalid particles) is >= u, for
// u = 0..ntiles. One thread per tile, by binary search: the cost does not
// depend on how the keys spread over the tiles (a few staged particles over
// many empty tiles, or every particle of the species).
__global__ void k_tile_starts(const uint32_t* keys, int64_t n, int shift, uint32_t ntiles, int64_t* first) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u > (int64_t)ntiles) return;
  int64_t lo = 0, hi = n;  // first index in [lo, hi) whose tile is >= u
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if ((int64_t)min(keys[mid] >> shift, ntiles) < u)
      lo = mid + 1;
    else
      hi = mid;
  }
  first[u] = lo;
}

__global__ void k_i32_from_i64(const int64_t* in, int32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int32_t)in[i];
}

__global__ void k_counts_from_starts(const int64_t* first, uint32_t ntiles, uint32_t* cnt) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u > ntiles) return;
  cnt[u] = u < ntiles ? (uint32_t)(first[u + 1] - first[u]) : 0u;
}

// Per-tile counts of n sorted keys (cnt[ntiles] = 0); `first` is scratch of ntiles+1.
void count_sorted_tiles(hipStream_t st, const uint32_t* keys, int64_t n, int shift, int nt, int64_t* first,
                        uint32_t* cnt) {
  if (n <= 0) {
    FOPIS_HIP(hipMemsetAsync(cnt, 0, (nt + 1) * sizeof(uint32_t), st));
    return;
  }
  k_tile_starts<<<div_up((int64_t)nt + 1, 256), 256, 0, st>>>(keys, n, shift, (uint32_t)nt, first);
  k_counts_from_starts<<<div_up(nt + 1, 256), 256, 0, st>>>(first, (uint32_t)nt, cnt);
}

__global__ void k_capacity(const uint32_t* cnt, int nt, double slack, int minc, uint32_t* capa, uint32_t* ocapa) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const double want = (double)cnt[t] * slack + (double)minc;
  uint32_t c = (uint32_t)ceil(want);
  c = (c + 63u) & ~63u;
  capa[t] = c;
  uint32_t oc = c / 4u;
  oc = (oc + 63u) & ~63u;
  if (oc < 64u) oc = 64u;
  ocapa[t] = oc;
}

// The tile's origin (cells) in the coordinates of its particles' storage:
// FP32 tile regions hold coordinates relative to it.
__device__ __forceinline__ void tile_origin(uint32_t t, int t1, int t2, int o[3]) {
  o[2] = (int)(t % (uint32_t)t2) * kTile;
  o[1] = (int)((t / (uint32_t)t2) % (uint32_t)t1) * kTile;
  o[0] = (int)(t / ((uint32_t)t2 * (uint32_t)t1)) * kTile;
}

// Particles of a flat set (absolute coordinates) into their tile regions;
// t1, t2: tiles along y and z (FP32 regions are relative to the tile's origin).
template <class R>
__global__ void k_fill_regions(SoaPtr<R> dst, const int64_t* start, SoaPtr<R> src, const uint32_t* keys,
                               const uint32_t* perm, int64_t n, int shift, const int64_t* first, uint32_t ntiles,
                               int t1, int t2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = keys[i] >> shift;
  if (t >= ntiles) return;
  const int64_t pos = start[t] + (i - first[t]);
  const uint32_t j = perm[i];
  int o[3] = {0, 0, 0};
  if (sizeof(R) == 4 && t1 > 0) tile_origin(t, t1, t2, o);
#pragma unroll
  for (int c = 0; c < 6; ++c) dst.c[c][pos] = src.c[c][j] - (c < 3 ? (R)o[c] : (R)0);
}

// Tile regions into one flat set in tile order, in absolute coordinates of
// type O (the origin of an FP32 region added in O: exact into FP64).
template <class R, class O>
__global__ __launch_bounds__(kBlock) void k_gather_regions(SoaPtr<O> dst, const int64_t* first, SoaPtr<R> src,
                                                           const int64_t* start, const int32_t* cnt, int t1,
                                                           int t2) {
  const int t = blockIdx.x;
  const int64_t s0 = start[t], d0 = first[t];
  const int n = cnt[t];
  int o[3] = {0, 0, 0};
  if (sizeof(R) == 4) tile_origin((uint32_t)t, t1, t2, o);
  for (int i = threadIdx.x; i < n; i += kBlock) {
#pragma unroll
    for (int c = 0; c < 6; ++c) dst.c[c][d0 + i] = (O)src.c[c][s0 + i] + (c < 3 ? (O)o[c] : (O)0);
  }
}

template <class R, class O>
__global__ void k_copy_soa(SoaPtr<O> dst, int64_t doff, SoaPtr<R> src, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) dst.c[c][doff + i] = (O)src.c[c][i];
}

template <class R>
__global__ void k_from_f64(SoaPtr<R> dst, SoaPtr<double> src, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) dst.c[c][i] = (R)src.c[c][i];
}

template <class R>
__global__ void k_to_f64(SoaPtr<double> dst, SoaPtr<R> src, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) dst.c[c][i] = (double)src.c[c][i];
}

__global__ void k_u32_from_i32(const int32_t* in, uint32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)in[i];
}

__global__ void k_i32_from_u32(const uint32_t* in, int32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int32_t)in[i];
}

template <class R>
__global__ void k_cell_keys(SoaPtr<R> f, int64_t n, int ny1, int nz1, int32_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ci = (int)floor(f.c[0][i]) + 1, cj = (int)floor(f.c[1][i]) + 1, ck = (int)floor(f.c[2][i]) + 1;
  keys[i] = (ci * ny1 + cj) * nz1 + ck;
}

// sum of v^2 over n particles into per-block partials
// Sum of vx^2 + vy^2 + vz^2 (FP64) over the live particles of the tile
// regions, read in place: a fixed grid of workgroups strides over the tiles
// and each writes one partial. A particle that a standalone position update
// absorbed keeps its slot with x = NaN until the next re-bin;
// Grid::getParticleEnergy skips it (GridParticleTreatment.cpp:104-118).
template <class R>
__global__ __launch_bounds__(kBlock) void k_v2_tiles(SoaPtr<R> f, const int64_t* start, const int32_t* cnt,
                                                     int ntiles, double* part) {
  __shared__ double s[kBlock];
  double acc = 0.0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t b0 = start[t];
    const int n = cnt[t];
    for (int i = threadIdx.x; i < n; i += kBlock) {
      if (isnan(f.c[0][b0 + i])) continue;
      const double vx = f.c[3][b0 + i], vy = f.c[4][b0 + i], vz = f.c[5][b0 + i];
      acc += vx * vx + vy * vy + vz * vz;
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Per-tile sums of x, y, z, vx, vy, vz and vx^2, vy^2, vz^2 (FP64) over the
// tile's region in place (one workgroup per tile).
template <class R>
__global__ __launch_bounds__(kBlock) void k_moments(SoaPtr<R> f, const int64_t* start, const int32_t* cnt,
                                                    double* part) {
  __shared__ double s[9][kBlock];
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t b0 = start[blockIdx.x], n = b0 + cnt[blockIdx.x];
  for (int64_t i = b0 + threadIdx.x; i < n; i += kBlock) {
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] += (double)f.c[q][i];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double v = (double)f.c[3 + q][i];
      acc[6 + q] += v * v;
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) s[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
#pragma unroll
      for (int q = 0; q < 9; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 9) part[(size_t)blockIdx.x * 9 + threadIdx.x] = s[threadIdx.x][0];
}

// Live particles of each tile region (x not NaN), one workgroup per tile.
template <class R>
__global__ __launch_bounds__(kBlock) void k_count_valid(const R* x, const int64_t* start, const int32_t* cnt,
                                                        unsigned long long* out) {
  const int64_t b0 = start[blockIdx.x];
  const int n = cnt[blockIdx.x];
  int k = 0;
  for (int i = threadIdx.x; i < n; i += kBlock) k += isnan(x[b0 + i]) ? 0 : 1;
  for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o);
  if ((threadIdx.x & 63) == 0 && k) atomicAdd(out, (unsigned long long)k);
}

__global__ void k_sum_serial(const double* part, int n, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *out = s;
  }
}

// ------------------------------------------------------- device generator
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
// as easy as 1, 2, 3"): 10 rounds of two 32x32->64 multiplies and a Weyl key.
struct U4 {
  uint32_t v[4];
};

__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.v[2];
    U4 n;
    n.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
    n.v[1] = (uint32_t)p1;
    n.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
    n.v[3] = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 53-bit uniform in [0, 1) from two 32-bit words
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

template <class R>
__global__ void k_load_maxwellian(SoaPtr<R> f, int64_t n, uint64_t seed, double3 lo, double3 hi, double3 drift,
                                  double sigma, const uint8_t* mask, int cy1, int cz1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  U4 c{{(uint32_t)i, (uint32_t)(i >> 32), 0u, 0u}};
  const U4 a = philox(c, k0, k1);
  c.v[2] = 1u;
  const U4 b = philox(c, k0, k1);
  c.v[2] = 2u;
  const U4 d = philox(c, k0, k1);
  const double x = lo.x + (hi.x - lo.x) * u01(a.v[0], a.v[1]);
  const double y = lo.y + (hi.y - lo.y) * u01(a.v[2], a.v[3]);
  const double z = lo.z + (hi.z - lo.z) * u01(b.v[0], b.v[1]);
  const double r1 = sqrt(-2.0 * log(1.0 - u01(b.v[2], b.v[3])));
  const double t1 = 6.283185307179586 * u01(d.v[0], d.v[1]);
  c.v[2] = 3u;
  const U4 e = philox(c, k0, k1);
  const double r2 = sqrt(-2.0 * log(1.0 - u01(d.v[2], d.v[3])));
  const double t2 = 6.283185307179586 * u01(e.v[0], e.v[1]);
  R px = (R)x, py = (R)y, pz = (R)z;
  if (mask) {
    const int ci = (int)floor(x) + 1, cj = (int)floor(y) + 1, ck = (int)floor(z) + 1;
    if (mask[((size_t)ci * cy1 + cj) * cz1 + ck] != 0) px = (R)NAN;
  }
  f.c[0][i] = px;
  f.c[1][i] = py;
  f.c[2][i] = pz;
  f.c[3][i] = (R)(drift.x + sigma * (r1 * cos(t1)));
  f.c[4][i] = (R)(drift.y + sigma * (r1 * sin(t1)));
  f.c[5][i] = (R)(drift.z + sigma * (r2 * cos(t2)));
}

// Boundary injection (RootGrid::injectParticlesFromBoundary,
// RootGridParticle.cpp:3-130, AmbientParticleType.cpp:101-157) drawn on the
// device: particle i of this call belongs to the face whose range of the
// prefix `first` holds it. Normal velocity: N(u_a, vth) conditioned on
// pointing into the domain (the reference redraws until it does; here the
// same truncated normal is drawn exactly by its inverse CDF); tangential
// velocities N(u, sigma); position uniform over the face and, along the
// normal, within v of it. Written as FP64 into the device staging.
struct InjArgs {
  int64_t first[7];      // prefix of the per-face counts (faces x-, x+, y-, y+, z-, z+)
  double drift[3];
  double vth, sigma;
  double mx[3];          // n - 1 per axis
  uint64_t seed;
};

__global__ void k_inject(SoaPtr<double> f, int64_t off, InjArgs A) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.first[6]) return;
  int face = 0;
  while (face < 5 && i >= A.first[face + 1]) ++face;
  const int axis = face >> 1, side = face & 1;
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  U4 c{{(uint32_t)i, (uint32_t)(i >> 32), 0u, 0u}};
  const U4 a = philox(c, k0, k1);
  c.v[2] = 1u;
  const U4 b = philox(c, k0, k1);
  c.v[2] = 2u;
  const U4 d = philox(c, k0, k1);
  // normal component: truncated normal on the inward half-line
  const double ua = A.drift[axis];
  const double p0 = normcdf((0.0 - ua) / A.vth);  // P(v < 0)
  const double uu = (u01(a.v[0], a.v[1]) + 0.5 / 9007199254740992.0);  // in (0, 1)
  double pn = side == 0 ? p0 + uu * (1.0 - p0) : uu * p0;
  pn = fmin(fmax(pn, 1e-300), 1.0 - 1.1102230246251565e-16);
  double vn = ua + A.vth * normcdfinv(pn);
  if (side == 0 && !(vn > 0.0)) vn = 1e-300;
  if (side == 1 && !(vn < 0.0)) vn = -1e-300;
  // tangential components (Box-Muller)
  const double r = sqrt(-2.0 * log(1.0 - u01(a.v[2], a.v[3])));
  const double th = 6.283185307179586 * u01(b.v[0], b.v[1]);
  double v[3];
  const int t1 = axis == 0 ? 1 : 0, t2 = axis == 2 ? 1 : 2;
  v[axis] = vn;
  v[t1] = A.drift[t1] + A.sigma * (r * cos(th));
  v[t2] = A.drift[t2] + A.sigma * (r * sin(th));
  // position
  double x[3];
  x[t1] = A.mx[t1] * u01(b.v[2], b.v[3]);
  x[t2] = A.mx[t2] * u01(d.v[0], d.v[1]);
  const double un = u01(d.v[2], d.v[3]);
  x[axis] = side == 0 ? un * vn : A.mx[axis] + (1.0 - un) * vn;
  const int64_t o = off + i;
  f.c[0][o] = x[0];
  f.c[1][o] = x[1];
  f.c[2][o] = x[2];
  f.c[3][o] = v[0];
  f.c[4][o] = v[1];
  f.c[5][o] = v[2];
}

// Photoemission from the lit face of an object (Spacecraft::emitParticles,
// Spacecraft.cpp:829-900, photoelectron branch) drawn on the device. Thread i
// emits particle i % per_node of lit node i / per_node:
// - velocity (PhotoElectronParticleType.cpp:9-106): speed vth, polar angle
//   to the outward normal asin(2u - 1), azimuth 2 pi u';
// - birth point: the node plus a uniform offset in the face square, moved
//   back along v by (1 - rt) of a step, rt uniform (inside the object);
// - its charge leaves the surface at the point where the path crosses the
//   face (Particle::getNext{X,Y,Z}CrossPoint, particle.cpp:206-255), shared
//   bilinearly over the face square's four cmat nodes
//   (subtractChargeOfParticleFrom{X,Y,Z}surface, Spacecraft.cpp:945-970);
// - this step's position update then moves it by v (the reference emits
//   between the velocity and the position update of mainLoopES), so the
//   staged particle is at birth point + v; one that ends outside the domain
//   is dropped when the staging is binned, as the position update's
//   absorbing faces would drop it.
struct EmitArgs {
  int64_t n;        // lit nodes x per_node
  int per_node, axis;
  double ev;        // outward normal along axis: -shine[axis]
  double vth, q;    // speed, superparticle charge
  int cx, cy, cz;   // glue node extents
  uint64_t seed;
};

__global__ void k_emit(SoaPtr<double> f, int64_t off, EmitArgs A, const int32_t* __restrict__ lit,
                       const int32_t* __restrict__ node_cmat, double* dq) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const int64_t node = i / A.per_node;
  const int ax = A.axis, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
  const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  U4 c{{(uint32_t)i, (uint32_t)(i >> 32), 0u, 0u}};
  const U4 a = philox(c, k0, k1);
  c.v[2] = 1u;
  const U4 b = philox(c, k0, k1);
  c.v[2] = 2u;
  const U4 d = philox(c, k0, k1);
  const double dep = asin(2.0 * u01(a.v[0], a.v[1]) - 1.0);
  const double azi = 6.283185307179586 * u01(a.v[2], a.v[3]);
  const double sd = sin(dep), ca = cos(azi), sa = sin(azi);
  double v[3];
  v[ax] = A.vth * cos(dep) * A.ev;
  // tangential pair in the reference's assignment (y: cos, z: sin for an x face, and so on)
  v[t1] = A.vth * sd * (ax == 1 ? sa : ca);
  v[t2] = A.vth * sd * (ax == 1 ? ca : sa);
  const double rt = u01(b.v[0], b.v[1]);
  const double r[2] = {u01(b.v[2], b.v[3]), u01(d.v[0], d.v[1])};
  const int ijk[3] = {lit[3 * node], lit[3 * node + 1], lit[3 * node + 2]};
  double p[3];
  for (int e = 0; e < 3; ++e) p[e] = (ijk[e] - 1.0) + (rt - 1.0) * v[e];
  p[t1] += r[0];
  p[t2] += r[1];
  // crossing point with the face plane
  double x[3] = {p[0], p[1], p[2]};
  const double frac = p[ax] - floor(p[ax]);
  if (v[ax] > 0.0) {
    const double m = (1.0 - frac) / v[ax];
    x[ax] = ceil(p[ax]);
    x[t1] = p[t1] + v[t1] * m;
    x[t2] = p[t2] + v[t2] * m;
  } else if (v[ax] < 0.0) {
    const double m = frac / (-v[ax]);
    x[ax] = floor(p[ax]);
    x[t1] = p[t1] + v[t1] * m;
    x[t2] = p[t2] + v[t2] * m;
  }
  int ci[3];
  double w1[3];
  for (int e = 0; e < 3; ++e) {
    ci[e] = (int)floor(x[e]) + 1;
    w1[e] = x[e] - (ci[e] - 1);
  }
  const int ext[3] = {A.cx, A.cy, A.cz};
#pragma unroll
  for (int du = 0; du < 2; ++du)
#pragma unroll
    for (int dw = 0; dw < 2; ++dw) {
      int nn[3] = {ci[0], ci[1], ci[2]};
      nn[t1] += du;
      nn[t2] += dw;
      bool ok = true;
      for (int e = 0; e < 3; ++e) ok = ok && nn[e] >= 0 && nn[e] < ext[e];
      if (!ok) continue;
      const int cm = node_cmat[((size_t)nn[0] * A.cy + nn[1]) * A.cz + nn[2]];
      if (cm < 0) continue;
      const double wu = du ? w1[t1] : 1.0 - w1[t1], ww = dw ? w1[t2] : 1.0 - w1[t2];
      atomicAdd(dq + cm, -(A.q * wu * ww));
    }
  const int64_t o = off + i;
  f.c[0][o] = p[0] + v[0];
  f.c[1][o] = p[1] + v[1];
  f.c[2][o] = p[2] + v[2];
  f.c[3][o] = v[0];
  f.c[4][o] = v[1];
  f.c[5][o] = v[2];
}

// Number of particles in [0, m] per axis, one atomic per wavefront.
__global__ void k_count_inside(SoaPtr<double> f, int64_t n, double mx, double my, double mz, unsigned long long* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = false;
  if (i < n) {
    const double x = f.c[0][i], y = f.c[1][i], z = f.c[2][i];
    ok = x >= 0.0 && y >= 0.0 && z >= 0.0 && x <= mx && y <= my && z <= mz;
  }
  const unsigned long long m = __ballot(ok);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(out, (unsigned long long)__popcll(m));
}

// ------------------------------------------------------------------ host side

namespace {

struct Scratch {
  RadixSorter sorter;
  DBuf<uint32_t> keys, perm, cnt, capa, ocapa;
  DBuf<int64_t> first, tmp;
  DBuf<unsigned long long> ucount;
};
Scratch& scratch(Ctx& c) {
  if (!c.particle_scratch) c.particle_scratch = std::make_shared<Scratch>();
  return *static_cast<Scratch*>(c.particle_scratch.get());
}

template <class R>
SoaPtr<R> ptrs(Soa<R>& s) {
  SoaPtr<R> p;
  for (int i = 0; i < 6; ++i) p.c[i] = s.c[i].p;
  return p;
}

template <class R>
std::vector<SpeciesBins<R>>& bins(Ctx& c);
template <>
std::vector<SpeciesBins<double>>& bins<double>(Ctx& c) {
  return c.sp64;
}
template <>
std::vector<SpeciesBins<float>>& bins<float>(Ctx& c) {
  return c.sp32;
}

template <class R>
TileArgs<R> tile_args(Ctx& c, int s) {
  auto& b = bins<R>(c)[s];
  TileArgs<R> a{};
  for (int q = 0; q < 3; ++q) {
    a.N[q] = c.Nc[q];
    a.T[q] = c.nt[q];
  }
  a.cx = c.cx;
  a.cy = c.cy;
  a.cz = c.cz;
  a.xmax[0] = (double)(c.nx - 1);
  a.xmax[1] = (double)(c.ny - 1);
  a.xmax[2] = (double)(c.nz - 1);
  const fopis_species_desc& sd = c.d.species[s];
  a.qm = 0.5 * sd.charge / sd.mass;
  a.q = sd.charge * sd.size;
  a.size = sd.size;
  a.bx = c.d.static_b[0];
  a.by = c.d.static_b[1];
  a.bz = c.d.static_b[2];
  a.boris = 2.0 / (1.0 + (a.bx * a.bx + a.by * a.by + a.bz * a.bz));
  a.ex = c.eref.p;
  a.e_s = (int64_t)c.nnode;
  a.rqm = (R)a.qm;
  a.rb[0] = (R)a.bx;
  a.rb[1] = (R)a.by;
  a.rb[2] = (R)a.bz;
  a.rboris = (R)a.boris;
  for (int q = 0; q < 3; ++q) a.rxmax[q] = (R)a.xmax[q];
  a.reg = ptrs(b.reg);
  a.dst = a.reg;
  a.start = b.start.p;
  a.capa = b.capa.p;
  a.cnt = b.cnt.p;
  a.out = ptrs(b.out);
  a.odir = b.odir.p;
  a.ostart = b.ostart.p;
  a.ocapa = b.ocapa.p;
  a.ocnt = b.ocnt.p;
  a.in = b.inbox.c[0].p;
  a.in_s = b.inbox.c[1].p - b.inbox.c[0].p;
  a.ibeg = b.inbox_n > 0 ? b.ibeg.p : nullptr;
  a.stray = ptrs(b.stray);
  a.stray_cap = b.stray_cap;
  a.status = c.status.p + kStat * s;
  a.rho_part = c.rho_part.p + (size_t)s * c.ntiles * (kRW * kRW * kRW);
  a.dens_part = c.dens_part.p + (size_t)s * c.ntiles * (kDW * kDW * kDW);
  a.rho_g = c.rho_far.p + (size_t)s * c.nnode;
  a.dens_g = c.dens_far.p + (size_t)s * c.ncell_d;
  a.mask = c.mask.p;
  a.tile_near = c.tile_near.p;
  a.sink = reinterpret_cast<R*>(c.sink.p);
  a.nobj = c.have_mask ? c.nobj : 0;
  for (int o = 0; o < c.nobj; ++o) {
    a.obj[o].node_cmat = c.obj[o].node_cmat.p;
    a.obj[o].tmp = c.obj[o].tmp_charge.p ? c.obj[o].tmp_charge.p + (size_t)s * c.obj[o].ncmat : nullptr;
    a.obj[o].absorbed = c.obj[o].absorbed.p ? c.obj[o].absorbed.p + s : nullptr;
    a.obj[o].lost = c.obj[o].absorbed.p ? c.obj[o].absorbed.p + c.S + s : nullptr;
  }
  a.cap_rec = c.cap_rec.p + (size_t)8 * c.cap_cap * s;
  a.cap_cap = c.cap_cap;
  return a;
}

// Bins n particles of a flat set into the species' tile regions (stable by
// tile), dropping invalid ones. key_cell additionally orders by cell inside a
// tile and keeps the current per-tile counts.
template <class R>
void bin_flat(Ctx& c, int s, Soa<R>& flat, int64_t n, bool key_cell) {
  auto& b = bins<R>(c)[s];
  Scratch& sc = scratch(c);
  hipStream_t st = c.stream;
  const int nt = c.ntiles;
  TileArgs<R> a = tile_args<R>(c, s);
  sc.keys.alloc(std::max<size_t>(sc.keys.n, std::max<int64_t>(n, 1)));
  sc.perm.alloc(std::max<size_t>(sc.perm.n, std::max<int64_t>(n, 1)));
  sc.cnt.alloc(std::max<size_t>(sc.cnt.n, nt + 1));
  sc.capa.alloc(std::max<size_t>(sc.capa.n, nt + 1));
  sc.ocapa.alloc(std::max<size_t>(sc.ocapa.n, nt + 1));
  sc.first.alloc(std::max<size_t>(sc.first.n, nt + 1));
  const int shift = key_cell ? 9 : 0;
  int bits = 1;
  while (((int64_t)1 << bits) < ((int64_t)nt << shift) + 1) ++bits;
  if (n > 0) {
    // invalid particles get the key one past the last tile and sort last
    k_keys_tile<R><<<div_up(n, 256), 256, 0, st>>>(ptrs(flat), n, a, sc.keys.p, key_cell ? 1 : 0,
                                                   (uint32_t)nt << shift);
    sc.sorter.sort(st, sc.keys.p, sc.perm.p, n, bits, true);
  }
  count_sorted_tiles(st, sc.keys.p, n, shift, nt, sc.first.p, sc.cnt.p);
  exclusive_scan_u32(st, sc.cnt.p, nt + 1, sc.first.p, sc.tmp);
  if (!key_cell) {
    k_capacity<<<div_up(nt, 256), 256, 0, st>>>(sc.cnt.p, nt, c.d.capacity_slack, 64, sc.capa.p, sc.ocapa.p);
    b.start.alloc(nt + 1);
    b.capa.alloc(nt + 1);
    b.cnt.alloc(nt + 1);
    b.ostart.alloc(nt + 1);
    b.ocapa.alloc(nt + 1);
    b.ocnt.alloc(nt + 1);
    b.arr.alloc(nt + 1);
    FOPIS_HIP(hipMemsetAsync(sc.capa.p + nt, 0, sizeof(uint32_t), st));
    FOPIS_HIP(hipMemsetAsync(sc.ocapa.p + nt, 0, sizeof(uint32_t), st));
    exclusive_scan_u32(st, sc.capa.p, nt + 1, b.start.p, sc.tmp);
    exclusive_scan_u32(st, sc.ocapa.p, nt + 1, b.ostart.p, sc.tmp);
    k_i32_from_u32<<<div_up(nt + 1, 256), 256, 0, st>>>(sc.capa.p, b.capa.p, nt + 1);
    k_i32_from_u32<<<div_up(nt + 1, 256), 256, 0, st>>>(sc.ocapa.p, b.ocapa.p, nt + 1);
    k_i32_from_u32<<<div_up(nt + 1, 256), 256, 0, st>>>(sc.cnt.p, b.cnt.p, nt + 1);
    FOPIS_HIP(hipMemsetAsync(b.ocnt.p, 0, (nt + 1) * sizeof(int32_t), st));
    FOPIS_HIP(hipMemsetAsync(b.arr.p, 0, (nt + 1) * sizeof(int32_t), st));
    int64_t tot[2];
    FOPIS_HIP(hipMemcpyAsync(&tot[0], b.start.p + nt, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FOPIS_HIP(hipMemcpyAsync(&tot[1], b.ostart.p + nt, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FOPIS_HIP(hipStreamSynchronize(st));
    if ((size_t)tot[0] != b.reg.size()) {
      b.reg.alloc(std::max<int64_t>(tot[0], 1));
      b.reg2.release();  // sized again by the next binning pass that needs it
    }
    if ((size_t)tot[1] != b.out.size()) {
      b.out.alloc(std::max<int64_t>(tot[1], 1));
      b.odir.alloc(std::max<int64_t>(tot[1], 1));
    }
    b.total_cap = tot[0];
    const int64_t want_stray = std::max<int64_t>(1 << 16, tot[0] / 32);
    if (b.stray_cap < want_stray) {
      b.stray.alloc(want_stray);
      b.stray_cap = want_stray;
    }
  }
  if (!key_cell) b.has_invalid = false;
  if (n > 0) {
    SoaPtr<R> dst = ptrs(b.reg);
    k_fill_regions<R><<<div_up(n, 256), 256, 0, st>>>(dst, b.start.p, ptrs(flat), sc.keys.p, sc.perm.p, n, shift,
                                                      sc.first.p, (uint32_t)nt, c.nt[1], c.nt[2]);
  }
  FOPIS_HIP(hipGetLastError());
}

// Copies every live particle of species s into `flat` (tile order, then the
// strays) and returns the count.
template <class R, class O = R>
int64_t gather_all(Ctx& c, int s, Soa<O>& flat, bool with_strays) {
  auto& b = bins<R>(c)[s];
  Scratch& sc = scratch(c);
  hipStream_t st = c.stream;
  const int nt = c.ntiles;
  if (!b.cnt.p) return 0;
  sc.cnt.alloc(std::max<size_t>(sc.cnt.n, nt + 1));
  sc.first.alloc(std::max<size_t>(sc.first.n, nt + 1));
  k_u32_from_i32<<<div_up(nt, 256), 256, 0, st>>>(b.cnt.p, sc.cnt.p, nt);
  FOPIS_HIP(hipMemsetAsync(sc.cnt.p + nt, 0, sizeof(uint32_t), st));
  exclusive_scan_u32(st, sc.cnt.p, nt + 1, sc.first.p, sc.tmp);
  int64_t nreg = 0;
  int32_t stat[4] = {0, 0, 0, 0};
  FOPIS_HIP(hipMemcpyAsync(&nreg, sc.first.p + nt, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  FOPIS_HIP(hipMemcpyAsync(stat, c.status.p + kStat * s, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  FOPIS_HIP(hipStreamSynchronize(st));
  if (stat[1] > 0) throw Error(FOPIS_E_NOMEM, "stray buffer overflow: particles were lost");
  const int64_t nstray = with_strays ? std::min<int64_t>(stat[0], b.stray_cap) : 0;
  const int64_t n = nreg + nstray;
  if ((int64_t)flat.size() < n) flat.alloc(n);
  if (nreg > 0)
    k_gather_regions<R, O><<<nt, kBlock, 0, st>>>(ptrs(flat), sc.first.p, ptrs(b.reg), b.start.p, b.cnt.p, c.nt[1],
                                                  c.nt[2]);
  if (nstray > 0) k_copy_soa<R, O><<<div_up(nstray, 256), 256, 0, st>>>(ptrs(flat), nreg, ptrs(b.stray), nstray);
  FOPIS_HIP(hipGetLastError());
  return n;
}

template <class R>
void rebin(Ctx& c, int s) {
  Soa<R> flat;
  const int64_t n = gather_all<R>(c, s, flat, true);
  FOPIS_HIP(hipMemsetAsync(c.status.p + kStat * s, 0, kStat * sizeof(int32_t), c.stream));
  bin_flat<R>(c, s, flat, n, false);
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

// Uploads the host-staged new particles and sorts them by tile.
template <class R>
void prepare_inbox(Ctx& c, int s, bool consume) {
  auto& b = bins<R>(c)[s];
  const int64_t nh = (int64_t)b.host_inbox.n;
  const int64_t n = nh + b.dev_n;
  b.inbox_n = n;
  if (n == 0) return;
  hipStream_t st = c.stream;
  Scratch& sc = scratch(c);
  const int nt = c.ntiles;
  Soa<double>& raw = b.inbox_raw;
  raw.reserve(n);
  // host-staged particles first, then the ones drawn on the device
  if (nh > 0) {
    for (int q = 0; q < 6; ++q)
      FOPIS_HIP(hipMemcpyAsync(raw.c[q].p, b.host_inbox.c[q], nh * sizeof(double), hipMemcpyHostToDevice, st));
    b.upload_pending = true;
  }
  if (b.dev_n > 0)
    for (int q = 0; q < 6; ++q)
      FOPIS_HIP(hipMemcpyAsync(raw.c[q].p + nh, b.dev_inbox.c[q].p, b.dev_n * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
  TileArgs<double> a{};
  for (int q = 0; q < 3; ++q) {
    a.N[q] = c.Nc[q];
    a.T[q] = c.nt[q];
  }
  a.xmax[0] = c.nx - 1;
  a.xmax[1] = c.ny - 1;
  a.xmax[2] = c.nz - 1;
  sc.keys.alloc(std::max<size_t>(sc.keys.n, n));
  sc.perm.alloc(std::max<size_t>(sc.perm.n, n));
  int bits = 1;
  while (((int64_t)1 << bits) < (int64_t)nt + 1) ++bits;
  k_keys_tile<double><<<div_up(n, 256), 256, 0, st>>>(ptrs(raw), n, a, sc.keys.p, 0, (uint32_t)nt);
  sc.sorter.sort(st, sc.keys.p, sc.perm.p, n, bits, true);
  sc.cnt.alloc(std::max<size_t>(sc.cnt.n, nt + 1));
  sc.first.alloc(std::max<size_t>(sc.first.n, nt + 1));
  count_sorted_tiles(st, sc.keys.p, n, 0, nt, sc.first.p, sc.cnt.p);
  exclusive_scan_u32(st, sc.cnt.p, nt + 1, sc.first.p, sc.tmp);
  b.ibeg.alloc(nt + 1);
  b.inbox.reserve(n);
  // inbox in tile order (invalid ones sort past ibeg[nt] and are not copied)
  k_fill_regions<double><<<div_up(n, 256), 256, 0, st>>>(ptrs(b.inbox), sc.first.p, ptrs(raw), sc.keys.p,
                                                         sc.perm.p, n, 0, sc.first.p, (uint32_t)nt, 0, 0);
  k_i32_from_i64<<<div_up(nt + 1, 256), 256, 0, st>>>(sc.first.p, b.ibeg.p, nt + 1);
  FOPIS_HIP(hipGetLastError());
  // the pinned staging may be refilled once the upload has been consumed by the
  // stream (the step synchronises on its status read before the next stage)
  if (consume) {
    b.host_inbox.n = 0;
    b.dev_n = 0;
  }
}

template <class R, int MODE>
void launch_tiles(Ctx& c, int s) {
  TileArgs<R> a = tile_args<R>(c, s);
  k_tile<R, MODE><<<c.ntiles, kBlock, 0, c.stream>>>(a);
  FOPIS_HIP(hipGetLastError());
}

// The binning pass of species s into its second region buffer (allocated on
// first use, the size of the first); the buffers swap once the arrivals are in.
// Whether species s can bin into a second region buffer: it is allocated on
// first use (the size of the first) when the device has room for it twice
// over (a later full re-bin gathers a flat copy of the species); otherwise
// the species compacts in place.
template <class R>
bool reg2_ready(Ctx& c, int s) {
  auto& b = bins<R>(c)[s];
  if (b.reg2.blk.p && b.reg2.size() == b.reg.size()) return true;
  b.reg2.release();
  size_t fr = 0, tot = 0;
  FOPIS_HIP(hipMemGetInfo(&fr, &tot));
  const size_t need = 6 * b.reg.size() * sizeof(R);
  if (fr < 2 * need + ((size_t)4 << 30)) return false;
  b.reg2.alloc(b.reg.size());
  return true;
}

template <class R, int MODE>
void launch_tiles_db(Ctx& c, int s) {
  auto& b = bins<R>(c)[s];
  TileArgs<R> a = tile_args<R>(c, s);
  a.dst = ptrs(b.reg2);
  k_tile<R, MODE, true><<<c.ntiles, kBlock, 0, c.stream>>>(a);
  FOPIS_HIP(hipGetLastError());
}

// Returns the species that deposited or counted past their tile windows.
template <class R>
uint32_t after_bin(Ctx& c) {
  std::vector<int32_t> stat(kStat * c.S);
  FOPIS_HIP(hipMemcpyAsync(stat.data(), c.status.p, kStat * c.S * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  for (auto& b : bins<R>(c)) b.upload_pending = false;
  for (int s = 0; s < c.S; ++s) {
    if (stat[kStat * s + 1] > 0)
      throw Error(FOPIS_E_NOMEM, "stray buffer overflow (" + std::to_string(stat[kStat * s + 1]) +
                                     " particles past its capacity): particles were lost");
    if (stat[kStat * s + 6] > 0)
      throw Error(FOPIS_E_NOMEM, "capture buffer overflow (" + std::to_string(stat[kStat * s + 3]) +
                                     " captured in one deposit, room for " + std::to_string(c.cap_cap) +
                                     "): particles were lost");
    auto& b = bins<R>(c)[s];
    b.captured += stat[kStat * s + 3];
    b.absorbed += stat[kStat * s + 4];
    b.joined += stat[kStat * s + 5];
  }
  // surface charge of the captured particles (Spacecraft::distributeInnerParticleCharge)
  for (int s = 0; s < c.S; ++s) {
    const int n = stat[kStat * s + 3];
    if (n <= 0 || c.nobj == 0) continue;
    ObjArgs oa[FOPIS_MAX_OBJECTS] = {};
    for (int o = 0; o < c.nobj; ++o) {
      oa[o].node_cmat = c.obj[o].node_cmat.p;
      oa[o].tmp = c.obj[o].tmp_charge.p + (size_t)s * c.obj[o].ncmat;
      oa[o].absorbed = c.obj[o].absorbed.p + s;
      oa[o].lost = c.obj[o].absorbed.p + c.S + s;
    }
    const fopis_species_desc& sd = c.d.species[s];
    k_capture<<<div_up(n, 128), 128, 0, c.stream>>>(c.cap_rec.p + (size_t)8 * c.cap_cap * s, n, c.cx, c.cy, c.cz,
                                                    sd.charge * sd.size, c.mask.p, oa[0], oa[1], oa[2], oa[3]);
    FOPIS_HIP(hipGetLastError());
  }
  // strays -> full re-bin of that species
  for (int s = 0; s < c.S; ++s)
    if (stat[kStat * s + 0] > 0) rebin<R>(c, s);
  uint32_t far = 0;
  for (int s = 0; s < c.S; ++s)
    if (stat[kStat * s + 2] > 0) far |= 1u << s;
  return far;
}

template <class R>
void deposit_bin(Ctx& c, bool push) {
  hipStream_t st = c.stream;
  Phase whole(c, "particles");
  for (int s = 0; s < c.S; ++s) prepare_inbox<R>(c, s, true);
  // (rho and density are written whole by the window reduction)
  FOPIS_HIP(hipMemsetAsync(c.status.p, 0, c.status.n * sizeof(int32_t), st));
  for (int o = 0; o < c.nobj; ++o)
    if (c.obj[o].absorbed.p)
      FOPIS_HIP(hipMemsetAsync(c.obj[o].absorbed.p, 0, 2 * c.S * sizeof(unsigned long long), st));
  std::vector<char> dbs(c.S, 0);
  for (int s = 0; s < c.S; ++s) dbs[s] = c.tile_double_buffer && reg2_ready<R>(c, s);
  {
    Phase ph(c, "tile");
    for (int s = 0; s < c.S; ++s) {
      if (dbs[s]) {
        if (push)
          launch_tiles_db<R, M_VEL | M_POS | M_DEP | M_BIN>(c, s);
        else
          launch_tiles_db<R, M_DEP | M_BIN>(c, s);
      } else {
        if (push)
          launch_tiles<R, M_VEL | M_POS | M_DEP | M_BIN>(c, s);
        else
          launch_tiles<R, M_DEP | M_BIN>(c, s);
      }
    }
  }
  {
    Phase ph(c, "bin");
    for (int s = 0; s < c.S; ++s) {
      auto& b = bins<R>(c)[s];
      TileArgs<R> a = tile_args<R>(c, s);
      if (dbs[s]) a.dst = ptrs(b.reg2);
      k_scatter<R><<<c.ntiles, kBlock, 0, st>>>(a, b.arr.p);
      k_arrive_done<<<div_up(c.ntiles, 256), 256, 0, st>>>(b.cnt.p, b.capa.p, b.arr.p, c.ntiles);
      if (dbs[s]) b.reg.swap(b.reg2);
    }
    FOPIS_HIP(hipGetLastError());
  }
  for (int s = 0; s < c.S; ++s) {
    bins<R>(c)[s].inbox_n = 0;
    bins<R>(c)[s].has_invalid = false;  // re-binning dropped every invalid particle
  }
  const uint32_t far_mask = after_bin<R>(c);
  {
    Phase ph(c, "reduce");
    fields_reduce_windows(c, true, true, far_mask);
  }
}

}  // namespace

void particles_init(Ctx& c) {
  const char* ip = std::getenv("FOPIS_TILE_INPLACE");
  c.tile_double_buffer = !(ip && ip[0] == '1');
  if (c.fp32)
    c.sp32.resize(c.S);
  else
    c.sp64.resize(c.S);
  // rho_part holds S*ntiles FP64 charge windows (both particle precisions)
  c.rho_part.alloc((size_t)c.S * c.ntiles * (kRW * kRW * kRW));
  c.rho_far.alloc((size_t)c.S * c.nnode);
  c.dens_far.alloc((size_t)c.S * c.ncell_d);
  FOPIS_HIP(hipMemsetAsync(c.rho_far.p, 0, c.rho_far.n * sizeof(double), c.stream));
  FOPIS_HIP(hipMemsetAsync(c.dens_far.p, 0, c.dens_far.n * sizeof(double), c.stream));
  c.dens_part.alloc((size_t)c.S * c.ntiles * (kDW * kDW * kDW));
  c.cap_cap = 1 << 20;  // captured particles per species and step
  c.cap_rec.alloc((size_t)8 * c.cap_cap * c.S);
  c.sink.alloc(kSink);  // doubles: kSink elements of either precision, and kSink bytes
  c.status.alloc(kStat * FOPIS_MAX_SPECIES);
  FOPIS_HIP(hipMemsetAsync(c.status.p, 0, c.status.n * sizeof(int32_t), c.stream));
  noalloc_setup(c);
  for (int s = 0; s < c.S; ++s) {
    if (c.fp32)
      bin_flat<float>(c, s, c.sp32[s].reg, 0, false);
    else
      bin_flat<double>(c, s, c.sp64[s].reg, 0, false);
  }
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

template <class R>
static void set_impl(Ctx& c, int s, int64_t n, const double* const* h) {
  hipStream_t st = c.stream;
  // the species is replaced as a whole: particles staged for the next deposit go too
  // (after any queued upload from the host staging has run)
  auto& b = bins<R>(c)[s];
  FOPIS_HIP(hipStreamSynchronize(st));
  b.host_inbox.n = 0;
  b.upload_pending = false;
  b.dev_n = 0;
  Soa<R> flat;
  if (n > 0) {
    flat.alloc(n);
    Soa<double> raw;
    raw.alloc(n);
    for (int q = 0; q < 6; ++q)
      FOPIS_HIP(hipMemcpyAsync(raw.c[q].p, h[q], n * sizeof(double), hipMemcpyHostToDevice, st));
    k_from_f64<R><<<div_up(n, 256), 256, 0, st>>>(ptrs(flat), ptrs(raw), n);
    FOPIS_HIP(hipStreamSynchronize(st));
  }
  FOPIS_HIP(hipMemsetAsync(c.status.p + kStat * s, 0, kStat * sizeof(int32_t), st));
  bin_flat<R>(c, s, flat, n, false);
  FOPIS_HIP(hipStreamSynchronize(st));
}

template <class R>
static void load_impl(Ctx& c, int s, int64_t n, uint64_t seed, const double* lo, const double* hi,
                      const double* drift, double sigma) {
  // the species is replaced as a whole, as in set_impl: particles staged for
  // the next deposit go too (after any queued upload from the host staging)
  auto& b = bins<R>(c)[s];
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  b.host_inbox.n = 0;
  b.upload_pending = false;
  b.dev_n = 0;
  Soa<R> flat;
  if (n > 0) {
    flat.alloc(n);
    k_load_maxwellian<R><<<div_up(n, 256), 256, 0, c.stream>>>(
        ptrs(flat), n, seed, make_double3(lo[0], lo[1], lo[2]), make_double3(hi[0], hi[1], hi[2]),
        make_double3(drift[0], drift[1], drift[2]), sigma, c.have_mask ? c.mask.p : nullptr, c.cy - 1, c.cz - 1);
    FOPIS_HIP(hipGetLastError());
  }
  FOPIS_HIP(hipMemsetAsync(c.status.p + kStat * s, 0, kStat * sizeof(int32_t), c.stream));
  bin_flat<R>(c, s, flat, n, false);
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

void particles_load_maxwellian(Ctx& c, int s, int64_t n, uint64_t seed, const double* lo, const double* hi,
                               const double* drift, double sigma) {
  for (int a = 0; a < 3; ++a) {
    if (!(lo[a] >= 0.0) || !(hi[a] <= (double)(a == 0 ? c.nx : (a == 1 ? c.ny : c.nz)) - 1.0) || !(hi[a] > lo[a]))
      throw Error(FOPIS_E_ARG, "load box must lie inside [0, n-1] with lo < hi");
  }
  if (c.fp32)
    load_impl<float>(c, s, n, seed, lo, hi, drift, sigma);
  else
    load_impl<double>(c, s, n, seed, lo, hi, drift, sigma);
}

void particles_set(Ctx& c, int s, int64_t n, const double* const* h) {
  if (c.fp32)
    set_impl<float>(c, s, n, h);
  else
    set_impl<double>(c, s, n, h);
}

// Device-staged particles inside the domain (binning drops the others).
template <class R>
static int64_t staged_inside(Ctx& c, SpeciesBins<R>& b) {
  if (b.dev_n == 0) return 0;
  Scratch& sc = scratch(c);
  sc.ucount.alloc(1);
  FOPIS_HIP(hipMemsetAsync(sc.ucount.p, 0, sizeof(unsigned long long), c.stream));
  k_count_inside<<<div_up(b.dev_n, 256), 256, 0, c.stream>>>(ptrs(b.dev_inbox), b.dev_n, c.nx - 1.0, c.ny - 1.0,
                                                              c.nz - 1.0, sc.ucount.p);
  FOPIS_HIP(hipGetLastError());
  unsigned long long k = 0;
  FOPIS_HIP(hipMemcpyAsync(&k, sc.ucount.p, sizeof(k), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  return (int64_t)k;
}

template <class R>
static int64_t count_impl(Ctx& c, int s) {
  auto& b = bins<R>(c)[s];
  int64_t n = 0;
  if (b.has_invalid) {
    // Grid::getValidParticleNumber counts valid particles only: a standalone
    // position update left the absorbed ones in their slots (x = NaN)
    Scratch& sc = scratch(c);
    sc.ucount.alloc(1);
    FOPIS_HIP(hipMemsetAsync(sc.ucount.p, 0, sizeof(unsigned long long), c.stream));
    k_count_valid<R><<<c.ntiles, kBlock, 0, c.stream>>>(b.reg.c[0].p, b.start.p, b.cnt.p, sc.ucount.p);
    FOPIS_HIP(hipGetLastError());
    unsigned long long k = 0;
    FOPIS_HIP(hipMemcpyAsync(&k, sc.ucount.p, sizeof(k), hipMemcpyDeviceToHost, c.stream));
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    n = (int64_t)k;
  } else {
    std::vector<int32_t> cnt(c.ntiles);
    FOPIS_HIP(hipMemcpyAsync(cnt.data(), b.cnt.p, c.ntiles * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    for (int v : cnt) n += v;
  }
  return n + (int64_t)b.host_inbox.n + staged_inside(c, b);
}

int64_t particles_count(Ctx& c, int s) { return c.fp32 ? count_impl<float>(c, s) : count_impl<double>(c, s); }

template <class R>
static void get_impl(Ctx& c, int s, int64_t cap, double* const* h, int64_t* nout) {
  auto& b = bins<R>(c)[s];
  // gathered straight into FP64 (an FP32 region's origin is added exactly)
  Soa<double> out;
  const int64_t n = gather_all<R, double>(c, s, out, false);
  // particles staged for the next step follow the resident ones: host-staged
  // first, then the ones drawn on the device (the order they are binned in);
  // staged particles outside the domain, which binning drops, are left out
  const int64_t nh = (int64_t)b.host_inbox.n, nd = b.dev_n;
  std::vector<double> dv(6 * (size_t)nd);
  for (int q = 0; q < 6 && nd > 0; ++q)
    FOPIS_HIP(hipMemcpyAsync(dv.data() + q * (size_t)nd, b.dev_inbox.c[q].p, nd * sizeof(double),
                             hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  const double lim[3] = {c.nx - 1.0, c.ny - 1.0, c.nz - 1.0};
  std::vector<int64_t> keep;
  keep.reserve(nd);
  for (int64_t i = 0; i < nd; ++i) {
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
      const double x = dv[a * (size_t)nd + i];
      ok = ok && x >= 0.0 && x <= lim[a];
    }
    if (ok) keep.push_back(i);
  }
  // resident particles; a standalone position update may have left absorbed
  // ones (x = NaN) in their slots until the next re-bin: they are not copied
  int64_t nv = n;
  std::vector<double> tmp;
  if (b.has_invalid && n > 0) {
    tmp.resize(6 * (size_t)n);
    for (int q = 0; q < 6; ++q)
      FOPIS_HIP(hipMemcpyAsync(tmp.data() + q * (size_t)n, out.c[q].p, n * sizeof(double), hipMemcpyDeviceToHost,
                               c.stream));
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    nv = 0;
    for (int64_t i = 0; i < n; ++i)
      if (!std::isnan(tmp[i])) {
        for (int q = 0; q < 6; ++q) tmp[q * (size_t)n + nv] = tmp[q * (size_t)n + i];
        ++nv;
      }
  }
  *nout = nv + nh + (int64_t)keep.size();
  if (*nout > cap) throw Error(FOPIS_E_ARG, "output buffer too small: need " + std::to_string(*nout));
  if (b.has_invalid) {
    for (int q = 0; q < 6 && nv > 0; ++q) std::memcpy(h[q], tmp.data() + q * (size_t)n, nv * sizeof(double));
  } else {
    for (int q = 0; q < 6 && n > 0; ++q)
      FOPIS_HIP(hipMemcpyAsync(h[q], out.c[q].p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FOPIS_HIP(hipStreamSynchronize(c.stream));
  }
  for (int q = 0; q < 6 && nh > 0; ++q) std::memcpy(h[q] + nv, b.host_inbox.c[q], nh * sizeof(double));
  for (int q = 0; q < 6; ++q)
    for (size_t k = 0; k < keep.size(); ++k) h[q][nv + nh + (int64_t)k] = dv[q * (size_t)nd + keep[k]];
}

void particles_get(Ctx& c, int s, int64_t cap, double* const* h, int64_t* n) {
  if (c.fp32)
    get_impl<float>(c, s, cap, h, n);
  else
    get_impl<double>(c, s, cap, h, n);
}

void particles_stage(Ctx& c, int s, int64_t n, const double* const* h) {
  PinSoa& hb = c.fp32 ? c.sp32[s].host_inbox : c.sp64[s].host_inbox;
  bool& pending = c.fp32 ? c.sp32[s].upload_pending : c.sp64[s].upload_pending;
  // an upload from the staging may still be queued: wait for it before refilling
  if (pending) {
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    pending = false;
  }
  hb.append(h, (size_t)n, c.stream);
}

// Grows the device staging to hold n more particles, keeping what earlier
// calls of this step drew.
template <class R>
static void grow_dev_inbox(Ctx& c, SpeciesBins<R>& b, int64_t n) {
  if ((int64_t)b.dev_inbox.size() >= b.dev_n + n) return;
  Soa<double> bigger;
  bigger.alloc(std::max<int64_t>(2 * (b.dev_n + n), 1 << 16));
  for (int q = 0; q < 6 && b.dev_n > 0; ++q)
    FOPIS_HIP(hipMemcpyAsync(bigger.c[q].p, b.dev_inbox.c[q].p, b.dev_n * sizeof(double), hipMemcpyDeviceToDevice,
                             c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  b.dev_inbox.swap(bigger);
}

template <class R>
static void inject_impl(Ctx& c, int s, const int64_t* count, uint64_t seed, const double* drift, double vth,
                        double sigma) {
  auto& b = bins<R>(c)[s];
  InjArgs A{};
  A.first[0] = 0;
  for (int f = 0; f < 6; ++f) {
    if (count[f] < 0) throw Error(FOPIS_E_ARG, "negative injection count");
    A.first[f + 1] = A.first[f] + count[f];
  }
  const int64_t n = A.first[6];
  if (n == 0) return;
  if (!(vth > 0.0) || !(sigma > 0.0)) throw Error(FOPIS_E_ARG, "vth and sigma must be positive");
  for (int a = 0; a < 3; ++a) A.drift[a] = drift[a];
  A.vth = vth;
  A.sigma = sigma;
  A.mx[0] = c.nx - 1;
  A.mx[1] = c.ny - 1;
  A.mx[2] = c.nz - 1;
  A.seed = seed;
  grow_dev_inbox(c, b, n);
  k_inject<<<div_up(n, 256), 256, 0, c.stream>>>(ptrs(b.dev_inbox), b.dev_n, A);
  FOPIS_HIP(hipGetLastError());
  b.dev_n += n;
}

void particles_inject(Ctx& c, int s, const int64_t* count, uint64_t seed, const double* drift, double vth,
                      double sigma) {
  if (c.fp32)
    inject_impl<float>(c, s, count, seed, drift, vth, sigma);
  else
    inject_impl<double>(c, s, count, seed, drift, vth, sigma);
}

template <class R>
static void emit_impl(Ctx& c, int o, int s, int axis, int sign, int per_node, double vth, uint64_t seed) {
  ObjectDev& ob = c.obj[o];
  const int32_t* lit = nullptr;
  const int nlit = objects_lit_nodes(c, o, axis, sign, &lit);
  const int64_t n = (int64_t)nlit * per_node;
  const double q = c.d.species[s].charge * c.d.species[s].size;
  // every emitted superparticle leaves the object (Spacecraft.cpp:951)
  ob.host_current_pending[s] -= q * (double)n;
  if (n == 0) return;
  auto& b = bins<R>(c)[s];
  EmitArgs A{};
  A.n = n;
  A.per_node = per_node;
  A.axis = axis;
  A.ev = sign > 0 ? -1.0 : 1.0;
  A.vth = vth;
  A.q = q;
  A.cx = c.cx;
  A.cy = c.cy;
  A.cz = c.cz;
  A.seed = seed;
  grow_dev_inbox(c, b, n);
  k_emit<<<div_up(n, 256), 256, 0, c.stream>>>(ptrs(b.dev_inbox), b.dev_n, A, lit, ob.node_cmat.p,
                                               ob.tmp_charge.p + (size_t)s * ob.ncmat);
  FOPIS_HIP(hipGetLastError());
  b.dev_n += n;
}

void particles_emit(Ctx& c, int o, int s, int axis, int sign, int per_node, double vth, uint64_t seed) {
  if (c.fp32)
    emit_impl<float>(c, o, s, axis, sign, per_node, vth, seed);
  else
    emit_impl<double>(c, o, s, axis, sign, per_node, vth, seed);
}

void particles_velocity(Ctx& c) {
  Phase ph(c, "velocity");
  for (int s = 0; s < c.S; ++s) {
    if (c.fp32)
      launch_tiles<float, M_VEL>(c, s);
    else
      launch_tiles<double, M_VEL>(c, s);
  }
}

void particles_position(Ctx& c) {
  {
    Phase ph(c, "position");
    FOPIS_HIP(hipMemsetAsync(c.status.p, 0, c.status.n * sizeof(int32_t), c.stream));
    for (int s = 0; s < c.S; ++s) {
      if (c.fp32) {
        launch_tiles<float, M_POS>(c, s);
        c.sp32[s].has_invalid = true;
      } else {
        launch_tiles<double, M_POS>(c, s);
        c.sp64[s].has_invalid = true;
      }
    }
  }
  std::vector<int32_t> stat(kStat * c.S);
  FOPIS_HIP(hipMemcpyAsync(stat.data(), c.status.p, stat.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  for (int s = 0; s < c.S; ++s) {
    if (c.fp32)
      c.sp32[s].absorbed += stat[kStat * s + 4];
    else
      c.sp64[s].absorbed += stat[kStat * s + 4];
  }
}

void particles_ledger(Ctx& c, int s, int64_t* out) {
  if (c.fp32) {
    out[0] = c.sp32[s].joined;
    out[1] = c.sp32[s].absorbed;
    out[2] = c.sp32[s].captured;
  } else {
    out[0] = c.sp64[s].joined;
    out[1] = c.sp64[s].absorbed;
    out[2] = c.sp64[s].captured;
  }
}

void particles_density_only(Ctx& c) {
  for (int s = 0; s < c.S; ++s) {
    if (c.fp32)
      prepare_inbox<float>(c, s, false);
    else
      prepare_inbox<double>(c, s, false);
  }
  FOPIS_HIP(hipMemsetAsync(c.density.p, 0, c.density.n * sizeof(double), c.stream));
  for (int s = 0; s < c.S; ++s) {
    if (c.fp32)
      launch_tiles<float, M_DENS>(c, s);
    else
      launch_tiles<double, M_DENS>(c, s);
  }
  for (int s = 0; s < c.S; ++s) {
    if (c.fp32)
      c.sp32[s].inbox_n = 0;
    else
      c.sp64[s].inbox_n = 0;
  }
  fields_reduce_windows(c, false, true);
}

void particles_deposit_bin(Ctx& c, bool push) {
  if (c.fp32)
    deposit_bin<float>(c, push);
  else
    deposit_bin<double>(c, push);
}

template <class R>
static void sort_cell_impl(Ctx& c) {
  for (int s = 0; s < c.S; ++s) {
    Soa<R> flat;
    rebin<R>(c, s);  // fold strays in first so every tile holds exactly its particles
    const int64_t n = gather_all<R>(c, s, flat, false);
    bin_flat<R>(c, s, flat, n, true);
    FOPIS_HIP(hipStreamSynchronize(c.stream));
  }
}

void particles_sort_by_cell(Ctx& c) {
  if (c.fp32)
    sort_cell_impl<float>(c);
  else
    sort_cell_impl<double>(c);
}

template <class R>
static void keys_impl(Ctx& c, int s, int64_t cap, int32_t* keys, int64_t* nout) {
  Soa<R> flat;
  const int64_t n = gather_all<R>(c, s, flat, false);
  *nout = n;
  if (n > cap) throw Error(FOPIS_E_ARG, "output buffer too small");
  if (n == 0) return;
  DBuf<int32_t> k;
  k.alloc(n);
  k_cell_keys<R><<<div_up(n, 256), 256, 0, c.stream>>>(ptrs(flat), n, c.ny + 1, c.nz + 1, k.p);
  FOPIS_HIP(hipMemcpyAsync(keys, k.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

void particles_cell_keys(Ctx& c, int s, int64_t cap, int32_t* keys, int64_t* n) {
  if (c.fp32)
    keys_impl<float>(c, s, cap, keys, n);
  else
    keys_impl<double>(c, s, cap, keys, n);
}

// Queues, for every species s, sum v^2 over its live particles into
// sums[s]: part holds kEnergyBlocks partials per species (fixed-order sums,
// so the result is the same from run to run).
template <class R>
static void queue_v2(Ctx& c, double* part, double* sums) {
  for (int s = 0; s < c.S; ++s) {
    auto& b = bins<R>(c)[s];
    double* ps = part + (size_t)s * kEnergyBlocks;
    if (b.cnt.p && c.ntiles > 0) {
      k_v2_tiles<R><<<kEnergyBlocks, kBlock, 0, c.stream>>>(ptrs(b.reg), b.start.p, b.cnt.p, c.ntiles, ps);
      k_sum_serial<<<1, 1, 0, c.stream>>>(ps, kEnergyBlocks, sums + s);
    } else {
      k_sum_serial<<<1, 1, 0, c.stream>>>(ps, 0, sums + s);
    }
  }
  FOPIS_HIP(hipGetLastError());
}

// 1/2 sum_s m_s size_s sum v^2 from the per-species sums copied to the host.
static double kinetic(const Ctx& c, const double* v2) {
  double res = 0.0;
  for (int s = 0; s < c.S; ++s) {
    const fopis_species_desc& sd = c.d.species[s];
    res += 0.5 * sd.mass * v2[s] * sd.size;
  }
  return res;
}

void particles_energy_queue(Ctx& c, double* part, double* sums) {
  if (c.fp32)
    queue_v2<float>(c, part, sums);
  else
    queue_v2<double>(c, part, sums);
}

double particles_energy_from(const Ctx& c, const double* v2) { return kinetic(c, v2); }

double particles_energy(Ctx& c) {
  const size_t np = (size_t)c.S * kEnergyBlocks;
  c.red.alloc(std::max<size_t>(c.red.n, np + c.S));
  particles_energy_queue(c, c.red.p, c.red.p + np);
  std::vector<double> v2(c.S);
  FOPIS_HIP(hipMemcpyAsync(v2.data(), c.red.p + np, c.S * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  return kinetic(c, v2.data());
}

template <class R>
static void moments_impl(Ctx& c, int s, double* out) {
  auto& b = bins<R>(c)[s];
  for (int q = 0; q < 10; ++q) out[q] = 0.0;
  if (!b.cnt.p) return;
  const int nb = c.ntiles;
  std::vector<int32_t> cnt(nb);
  DBuf<double> part;
  part.alloc((size_t)nb * 9);
  k_moments<R><<<nb, kBlock, 0, c.stream>>>(ptrs(b.reg), b.start.p, b.cnt.p, part.p);
  std::vector<double> h((size_t)nb * 9);
  FOPIS_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipMemcpyAsync(cnt.data(), b.cnt.p, nb * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  int64_t n = 0;
  for (int t = 0; t < nb; ++t) n += cnt[t];
  out[0] = (double)n;
  if (n == 0) return;
  double sum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = 0; t < nb; ++t) {
    for (int q = 0; q < 9; ++q) sum[q] += h[(size_t)t * 9 + q];
    if (sizeof(R) == 4) {  // FP32 regions are relative to the tile's origin
      const int o[3] = {t / (c.nt[2] * c.nt[1]) * kTile, (t / c.nt[2]) % c.nt[1] * kTile, t % c.nt[2] * kTile};
      for (int q = 0; q < 3; ++q) sum[q] += (double)cnt[t] * o[q];
    }
  }
  for (int q = 0; q < 6; ++q) out[1 + q] = sum[q] / (double)n;
  for (int q = 0; q < 3; ++q) out[7 + q] = sum[6 + q] / (double)n - out[4 + q] * out[4 + q];
}

// count, mean x, y, z, mean vx, vy, vz, variance of vx, vy, vz of species s
void particles_moments(Ctx& c, int s, double* out) {
  if (c.fp32)
    moments_impl<float>(c, s, out);
  else
    moments_impl<double>(c, s, out);
}

// ------------------------------------------- allocation-free energy reduction
//
// fopis_energy_noalloc keeps its partial sums in this file-scope array instead
// of library-allocated memory: kEnergyBlocks partials per species and for the
// field, then the S + 1 sums. Every call overwrites all of them (nothing has to
// start at zero). Each device holds its own copy for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device, never through the array's name.
constexpr size_t kNoAllocDoubles = (size_t)(FOPIS_MAX_SPECIES + 1) * kEnergyBlocks + FOPIS_MAX_SPECIES + 1;
static __device__ __align__(256) double noalloc_scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_setup(Ctx& c) {
  int dev = 0;
  FOPIS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(FOPIS_E_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess || !p)
      throw Error(FOPIS_E_DEVICE, std::string("scratch array address: ") + hipGetErrorString(e));
    g_noalloc_ptr[dev] = static_cast<double*>(p);
  }
  c.noalloc = g_noalloc_ptr[dev];
}

void energy_noalloc(Ctx& c, double* out) {
  if (!c.noalloc) throw Error(FOPIS_E_STATE, "scratch array not set up for this context");
  const size_t need = (size_t)(c.S + 1) * kEnergyBlocks + c.S + 1;
  if (c.S < 0 || c.S > FOPIS_MAX_SPECIES || need > kNoAllocDoubles)
    throw Error(FOPIS_E_ARG, "energy partials of " + std::to_string(c.S) + " species exceed the scratch array (" +
                                 std::to_string(FOPIS_MAX_SPECIES) + " species at most)");
  double* part = c.noalloc;
  double* fpart = part + (size_t)c.S * kEnergyBlocks;
  double* sums = fpart + kEnergyBlocks;
  particles_energy_queue(c, part, sums);
  fields_energy_queue(c, fpart, sums + c.S);
  double h[FOPIS_MAX_SPECIES + 1];
  FOPIS_HIP(hipMemcpyAsync(h, sums, (size_t)(c.S + 1) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  out[0] = particles_energy_from(c, h);
  out[1] = 0.5 * c.d.eps0 * h[c.S];
}

}  // namespace fopis
