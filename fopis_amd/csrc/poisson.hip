// Poisson solve for the ES step: FP64 defect correction around geometric
// multigrid V-cycles (red-black SOR smoother, full-weighting restriction,
// trilinear prolongation) on the vertex-centred node grid.
//
// It solves the reference's discrete problem (RootGridField.cpp:105-264):
// the 7-point operator 6 phi - sum(neighbours) = rho_coeff * rho[0] at nodes
// 2..n-1 of each axis (glue indexing), with the face nodes held at their
// current values (0 on Dirichlet faces), then applies the face conditions
// (Field.cpp:10-96). The reference reaches it by zebra SOR; this build stops
// on ||b - A phi||_2 / ||b||_2 <= tol in FP64.
//
// Launch structure. The host queues cycles in batches without waiting: every
// kernel of a cycle first reads a device flag that the norm kernel sets once
// the tolerance (or the cycle cap) is reached, and returns at once when it is
// set, so a batch that overshoots costs only empty launches. The host reads
// the flag once per batch. The coarse levels whose arrays fit in LDS run as
// one single-workgroup kernel per cycle.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "fopis_internal.hpp"

#pragma clang fp contract(off)

namespace fopis {

namespace {

struct Lv {
  int n0, n1, n2;
  int64_t sx, sy;
};

// Workgroups are handed to the 8 XCDs round-robin by linear id, and each XCD
// has its own L2. Renumber them so that XCD x runs a contiguous run of the
// logical ids (whole planes of the level), keeping a stencil's neighbour rows
// in the same L2. The map is a bijection for any grid size.
__device__ inline int xcd_block(int p, int n) {
  const int per = n >> 3, rem = n & 7, x = p & 7, slot = p >> 3;
  return x * per + min(x, rem) + slot;
}

// (chunk, row, plane) of the calling workgroup in a 3-D launch, XCD-renumbered.
__device__ inline void block3(int& bx, int& by, int& bz) {
  const int gx = gridDim.x, gy = gridDim.y;
  const int n = gx * gy * gridDim.z;
  const int L = xcd_block(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), n);
  bx = L % gx;
  by = (L / gx) % gy;
  bz = L / (gx * gy);
}

// The level-0 passes below (residuals, update) walk the interior by rows:
// each workgroup takes one contiguous range of (a, b) rows (XCD-renumbered, so
// that neighbouring planes stay in one XCD's L2) and spreads its 256 threads
// over 1, 2 or 4 rows at a time, a row's points over 256, 128 or 64 threads.
// f(g, a, b, c) gets the node index and the coordinates of every interior point once.
template <class F>
__device__ inline void for_interior_rows(const int n0, const int n1, const int n2, const int64_t sx,
                                         const int64_t sy, F&& f) {
  const int nrows = n0 * n1;
  const int per = (nrows + (int)gridDim.x - 1) / (int)gridDim.x;
  const int r0 = xcd_block(blockIdx.x, gridDim.x) * per;
  const int r1 = min(r0 + per, nrows);
  const int tpr = n2 <= 64 ? 64 : (n2 <= 128 ? 128 : 256);  // threads per row
  const int sub = threadIdx.x / tpr, lane = threadIdx.x % tpr;
  for (int row = r0 + sub; row < r1; row += 256 / tpr) {
    const int a = row / n1 + 1, b = row % n1 + 1;
    const int64_t base = a * sx + b * sy;
    for (int c = lane + 1; c <= n2; c += tpr) f(base + c, a, b, c);
  }
}

// b = rc * rho at the level-0 interior; r = b - A phi written as P;
// per-block partial sums of r^2 and b^2 in FP64.
// (r64, when given, receives the FP64 residual that the cycle updates keep;
// with r and r64 both null only the norms are formed: the check at a solve's end.)
template <class P>
__global__ __launch_bounds__(256) void k_resid0(const double* __restrict__ phi, const double* __restrict__ rho,
                                                double rc, P* __restrict__ r, Lv L, double* part,
                                                const int* __restrict__ done, double* __restrict__ r64) {
  if (*done) return;
  __shared__ double s1[256], s2[256];
  double rr = 0.0, bb = 0.0;
  for_interior_rows(L.n0, L.n1, L.n2, L.sx, L.sy, [&](const int64_t g, int, int, int) {
    const double bv = rc * rho[g];
    const double ap = 6.0 * phi[g] - (phi[g - L.sx] + phi[g + L.sx] + phi[g - L.sy] + phi[g + L.sy] + phi[g - 1] +
                                      phi[g + 1]);
    const double rv = bv - ap;
    if (r) r[g] = (P)rv;
    if (r64) r64[g] = rv;
    rr += rv * rv;
    bb += bv * bv;
  });
  s1[threadIdx.x] = rr;
  s2[threadIdx.x] = bb;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

// Solve state on the device: st[0] done flag, st[1] cycles applied, st[3]
// stays 0 (the done flag of launches that always run), st[4] the current
// pair's skip flag, st[5] FP64 updates applied, st[6] pair residual passes
// run; dst[0..2] = ||r||^2, ||b||^2, relative residual of the
// last check, dst[3] the relative residual the solve started from.
__global__ void k_solve_begin(int* st) {
  st[0] = 0;
  st[1] = 0;
  st[5] = 0;
  st[6] = 0;
}

// After a V-cycle's correction e: phi += e and the FP64 residual
// r = b - A phi is updated as r -= A e (exact up to FP64 rounding, so no
// stencil over phi is needed), its P copy becomes the next cycle's right-hand
// side, and per-block partial sums of r^2 are written (the b^2 slot is 0).
// Reads e at neighbours and writes phi, r only at the point itself.
template <class P>
__global__ __launch_bounds__(256) void k_update(double* __restrict__ phi, const P* __restrict__ e, double* r64,
                                                P* rp, Lv L, double* part, const int* __restrict__ done) {
  if (*done) return;
  __shared__ double s1[256];
  const bool same = (const void*)rp == (const void*)r64;
  double rr = 0.0;
  for_interior_rows(L.n0, L.n1, L.n2, L.sx, L.sy, [&](const int64_t g, int, int, int) {
    const double eg = (double)e[g];
    const double ae = 6.0 * eg - ((double)e[g - L.sx] + (double)e[g + L.sx] + (double)e[g - L.sy] +
                                  (double)e[g + L.sy] + (double)e[g - 1] + (double)e[g + 1]);
    const double rv = r64[g] - ae;
    r64[g] = rv;
    if (!same) rp[g] = (P)rv;
    phi[g] += eg;
    rr += rv * rv;
  });
  s1[threadIdx.x] = rr;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s1[threadIdx.x] += s1[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// Between the two V-cycles of a pair: per-block partial sums of t^2, t = r - A e
// at the level-0 interior (the cycles' right-hand side and the first cycle's
// correction; t formed in P as the cycles form it, its squares summed in
// FP64; the b^2 slot is 0).
template <class P>
__global__ __launch_bounds__(256) void k_pair_resid(const P* __restrict__ e, const P* __restrict__ r, Lv L,
                                                    double* part, const int* __restrict__ done) {
  if (*done) return;
  __shared__ double s1[256];
  double rr = 0.0;
  for_interior_rows(L.n0, L.n1, L.n2, L.sx, L.sy, [&](const int64_t g, int, int, int) {
    const P t = r[g] - ((P)6 * e[g] - (e[g - L.sx] + e[g + L.sx] + e[g - L.sy] + e[g + L.sy] + e[g - 1] + e[g + 1]));
    rr += (double)t * (double)t;
  });
  s1[threadIdx.x] = rr;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s1[threadIdx.x] += s1[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// The pair's decision (one block, fixed tree): skip[0] = 1, so that every
// kernel of the second V-cycle returns at once, when the solve is already
// done or the first cycle's correction alone reaches the tolerance (against
// ||b||^2 of the solve's start, dst[1]). The residual after the first cycle is
// the cycles' right-hand side minus A e; that right-hand side is the FP64
// residual rounded to P, whose rounding (~1e-7 of it in FP32) is far below
// the ~30x a V-cycle removes, so the test agrees with the FP64 check that
// follows the update except within rounding of the tolerance itself.
__global__ __launch_bounds__(256) void k_pair_check(const double* part, int nb, const double* dst, int* st,
                                                    double tol, int* skip) {
  __shared__ double s1[256];
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) a += part[2 * i];
  s1[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s1[threadIdx.x] += s1[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rr = s1[0], bb = dst[1];
    const double rel = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
    skip[0] = (st[0] || rel <= tol || (bb == 0.0 && rr < 1e-300)) ? 1 : 0;
    if (!st[0]) st[6] += 1;  // (a pass that ran; k_pair_resid returns at once once done)
  }
}

// Final sum of the residual partials (one block, fixed tree: deterministic),
// cycle count and the stopping test of the solve loop. keep_bb: ||b||^2 is
// the one already in dst[1] (the partials carry r^2 only). count_cycle V-cycles
// were applied, less one when *skipped (a pair whose second cycle was skipped).
__global__ __launch_bounds__(256) void k_norm(const double* part, int nb, double* dst, int* st, double tol,
                                              int max_cycles, int count_cycle, int keep_bb,
                                              const int* skipped = nullptr) {
  if (st[0]) return;
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rr = s1[0], bb = keep_bb ? dst[1] : s2[0];
    const int cyc = st[1] + count_cycle - (skipped ? skipped[0] : 0);
    if (count_cycle > 0) st[5] += 1;  // an FP64 update preceded this check
    const double rel = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
    dst[0] = rr;
    dst[1] = bb;
    dst[2] = rel;
    st[1] = cyc;
    st[0] = (rel <= tol || (bb == 0.0 && rr < 1e-300) || cyc >= max_cycles) ? 1 : 0;
  }
}

// The true residual's norms at the end of a solve (flag untouched).
__global__ __launch_bounds__(256) void k_norm_final(const double* part, int nb, double* dst) {
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    dst[0] = s1[0];
    dst[1] = s2[0];
    dst[2] = s2[0] > 0.0 ? sqrt(s1[0] / s2[0]) : sqrt(s1[0]);
  }
}

__global__ void k_reopen(int* st) { st[0] = 0; }

// The relative residual the solve starts from, kept beside the running norms.
__global__ void k_keep_start(double* dst) { dst[3] = dst[2]; }

// Kernels over a level's interior give each thread the same column position
// in kRows consecutive rows (b), issue every load of those rows before the
// first store, and so keep several independent loads in flight per thread.
constexpr int kRows = 4;

// Red-black SOR half-sweep over points with (a+b+c) % 2 == color:
// e = (1-w) e + (w/6) (sum of neighbours + r). A half-sweep reads only points
// of the other colour and writes only its own, so the loads of all rows may
// precede the stores.
template <class P>
__global__ void k_smooth(P* __restrict__ e, const P* __restrict__ r, Lv L, int color, P w1, P w6,
                         const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int m = bx * blockDim.x + threadIdx.x;
  const int a = bz + 1;
  P val[kRows];
  int64_t gg[kRows];
  bool ok[kRows];
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    const int b = by * kRows + q + 1;
    const int c = (((a + b + 1) & 1) == color ? 1 : 2) + 2 * m;
    ok[q] = b <= L.n1 && c <= L.n2;
    const int64_t g = a * L.sx + (ok[q] ? b : 1) * L.sy + (ok[q] ? c : 1);
    gg[q] = g;
    const P nb = e[g - L.sx] + e[g + L.sx] + e[g - L.sy] + e[g + L.sy] + e[g - 1] + e[g + 1];
    val[q] = w1 * e[g] + w6 * (nb + r[g]);
  }
#pragma unroll
  for (int q = 0; q < kRows; ++q)
    if (ok[q]) e[gg[q]] = val[q];
}

// First full sweep of a cycle from e = 0, both colours in one pass: red
// points get (w/6) r, black points the black half-sweep over those red values
// (formed here from r): the operations of a red half-sweep from zero (red
// points (w/6) r, black points cleared) followed by the black half-sweep, in
// the same order (the compiler's fused multiply-adds may round them
// differently), but r is read once and e written once. Points outside the
// interior have r = 0 and so read as red values of 0, as the untouched level
// borders do.
template <class P>
__global__ void k_smooth_first2(P* __restrict__ e, const P* __restrict__ r, Lv L, P w1, P w6,
                                const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  // two neighbouring points of a row per thread, one of each colour, so the
  // lanes of a wavefront all take the same path; every load of the kRows rows
  // is issued before the first store (rows past the level read row 1 instead)
  const int c0 = 2 * (bx * blockDim.x + threadIdx.x) + 1;
  const int a = bz + 1;
  if (c0 > L.n2) return;
  const bool two = c0 + 1 <= L.n2;
  P vb[kRows], vr[kRows];
  int64_t gbq[kRows], grq[kRows];
  bool okq[kRows], fbq[kRows];
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    const int b = by * kRows + q + 1;
    okq[q] = b <= L.n1;
    const int bb = okq[q] ? b : 1;
    const int64_t g0 = a * L.sx + bb * L.sy + c0;
    // the black point of the pair is c0 when (a + b + c0) is odd, else c0 + 1
    fbq[q] = ((a + bb + c0) & 1) != 0;
    const int64_t gb = fbq[q] ? g0 : g0 + 1;
    const int64_t gr = fbq[q] ? g0 + 1 : g0;
    gbq[q] = gb;
    grq[q] = gr;
    const P nb = w6 * r[gb - L.sx] + w6 * r[gb + L.sx] + w6 * r[gb - L.sy] + w6 * r[gb + L.sy] + w6 * r[gb - 1] +
                 w6 * r[gb + 1];
    vb[q] = w1 * (P)0 + w6 * (nb + r[gb]);
    vr[q] = w6 * r[gr];
  }
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    if (!okq[q]) continue;
    if (fbq[q] || two) e[gbq[q]] = vb[q];
    if (!fbq[q] || two) e[grq[q]] = vr[q];
  }
}

template <class P>
__device__ inline P fw_weight(int da, int db, int dc) {
  const int nz = (da != 0) + (db != 0) + (dc != 0);
  return nz == 0 ? (P)1 : (nz == 1 ? (P)0.5 : (nz == 2 ? (P)0.25 : (P)0.125));
}

// Full weighting of the fine residual r - A e to the coarse grid, scaled by 4
// for the unscaled 2h operator; the residual of each of the 27 fine points is
// formed on the fly (no residual array is written).
template <class P>
__global__ void k_restrict_resid(P* __restrict__ rc, Lv C, const P* __restrict__ e, const P* __restrict__ r, Lv F,
                                 const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int c = bx * blockDim.x + threadIdx.x + 1;
  const int b = by + 1, a = bz + 1;
  if (c > C.n2) return;
  const int64_t f = (2 * a) * F.sx + (2 * b) * F.sy + 2 * c;
  P acc = 0;
#pragma unroll
  for (int da = -1; da <= 1; ++da)
#pragma unroll
    for (int db = -1; db <= 1; ++db)
#pragma unroll
      for (int dc = -1; dc <= 1; ++dc) {
        const int64_t g = f + da * F.sx + db * F.sy + dc;
        const P t = r[g] - ((P)6 * e[g] - (e[g - F.sx] + e[g + F.sx] + e[g - F.sy] + e[g + F.sy] + e[g - 1] + e[g + 1]));
        acc += fw_weight<P>(da, db, dc) * t;
      }
  rc[a * C.sx + b * C.sy + c] = acc * (P)0.5;  // (4/8) * sum
}

// The residual t = r - A e at every interior point of a level, kRows rows per
// thread (the same expression as the restriction forms on the fly).
template <class P>
__global__ void k_resid_level(P* __restrict__ t, const P* __restrict__ e, const P* __restrict__ r, Lv L,
                              const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int c = bx * blockDim.x + threadIdx.x + 1;
  const int a = bz + 1;
  if (c > L.n2) return;
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    const int b = by * kRows + q + 1;
    if (b > L.n1) break;
    const int64_t g = a * L.sx + b * L.sy + c;
    t[g] = r[g] - ((P)6 * e[g] - (e[g - L.sx] + e[g + L.sx] + e[g - L.sy] + e[g + L.sy] + e[g - 1] + e[g + 1]));
  }
}

// Full weighting of a residual already formed (k_resid_level), in the same
// order as k_restrict_resid: on large levels every fine residual is then
// computed once instead of once per coarse point that weights it.
template <class P>
__global__ void k_restrict_t(P* __restrict__ rc, Lv C, const P* __restrict__ t, Lv F, const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int c = bx * blockDim.x + threadIdx.x + 1;
  const int b = by + 1, a = bz + 1;
  if (c > C.n2) return;
  const int64_t f = (2 * a) * F.sx + (2 * b) * F.sy + 2 * c;
  P acc = 0;
#pragma unroll
  for (int da = -1; da <= 1; ++da)
#pragma unroll
    for (int db = -1; db <= 1; ++db)
#pragma unroll
      for (int dc = -1; dc <= 1; ++dc) acc += fw_weight<P>(da, db, dc) * t[f + da * F.sx + db * F.sy + dc];
  rc[a * C.sx + b * C.sy + c] = acc * (P)0.5;  // (4/8) * sum
}

// e_fine += trilinear interpolation of e_coarse (coarse node A sits at fine 2A).
template <class P>
__global__ void k_prolong(P* __restrict__ e, Lv F, const P* __restrict__ ec, Lv C, const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int c = bx * blockDim.x + threadIdx.x + 1;
  const int a = bz + 1;
  if (c > F.n2) return;
  const int a0 = a >> 1, a1 = (a + 1) >> 1, c0 = c >> 1, c1 = (c + 1) >> 1;
  auto E = [&](int x, int y, int z) { return ec[x * C.sx + y * C.sy + z]; };
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    const int b = by * kRows + q + 1;
    if (b > F.n1) break;
    const int b0 = b >> 1, b1 = (b + 1) >> 1;
    const P v = (E(a0, b0, c0) + E(a1, b0, c0) + E(a0, b1, c0) + E(a1, b1, c0) + E(a0, b0, c1) + E(a1, b0, c1) +
                 E(a0, b1, c1) + E(a1, b1, c1)) *
                (P)0.125;
    e[a * F.sx + b * F.sy + c] += v;
  }
}

// ---- the finest FP32 level stored by colour ------------------------------
//
// A red-black half-sweep updates the points of one colour from those of the
// other. With both colours in one array every cache line holds both, so a
// half-sweep reads all of e and r and writes all of e: 201 MB per half-sweep
// at 255^3 where the data it needs is 136 MB. Stored by colour, it reads the
// other colour's e and its own e and r and writes its own e, each half-array
// once. Point (a, b, c) of the padded level (interior 1..n) has colour
// (a + b + c) & 1 and sits at a * sx + b * sy + c / 2 of its colour's array
// (rows of (n2 + 3) / 2); its x neighbours c - 1 and c + 1 are at (c - 1) / 2
// and (c + 1) / 2 of the other colour's row, its y and z neighbours at c / 2
// of the other colour's neighbouring rows. Every kernel here forms the same
// sums in the same order as its one-array counterpart, so a solve gives the
// same result bit for bit either way.
struct Rb {
  float* e;  // colour 0; colour 1 at + cs
  float* r;
  int64_t cs, sx, sy;
  int n0, n1, n2;
};

__device__ inline int64_t rb_at(const Rb& R, int a, int b, int c) {
  return a * R.sx + b * R.sy + (c >> 1) + (((a + b + c) & 1) ? R.cs : 0);
}

// A point's slot i in its colour's array, that array (own) and the other
// colour's (oth), where all six neighbours are: oth[i -+ sx], oth[i -+ sy],
// and oth[i + xm], oth[i + xp] for c - 1 and c + 1.
struct RbPt {
  int64_t i;
  int64_t own, oth;  // offsets of the two colours' arrays
  int xm, xp;
};

__device__ inline RbPt rb_pt(const Rb& R, int a, int b, int c) {
  const bool k = ((a + b + c) & 1) != 0, codd = (c & 1) != 0;
  return RbPt{a * R.sx + b * R.sy + (c >> 1), k ? R.cs : 0, k ? 0 : R.cs, codd ? 0 : -1, codd ? 1 : 0};
}

// A e at the point (6 e - the neighbours, summed in the one-array order:
// -x, +x, -y, +y, -z, +z in the level's a, b, c axes)
template <class T>
__device__ inline T rb_ae(const Rb& R, const RbPt& p) {
  const float* eo = R.e + p.own;
  const float* en = R.e + p.oth;
  return (T)6 * (T)eo[p.i] - ((T)en[p.i - R.sx] + (T)en[p.i + R.sx] + (T)en[p.i - R.sy] + (T)en[p.i + R.sy] +
                              (T)en[p.i + p.xm] + (T)en[p.i + p.xp]);
}

// Red-black half-sweep of colour `color` (k_smooth): thread j takes the j-th
// point of the colour in each of its rows (c = 2j + 1 or 2j + 2).
__global__ void k_smooth_rb(Rb R, int color, float w1, float w6, const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int j = bx * blockDim.x + threadIdx.x;
  const int a = bz + 1;
  float* own = R.e + (color ? R.cs : 0);
  const float* oth = R.e + (color ? 0 : R.cs);
  const float* ro = R.r + (color ? R.cs : 0);
  float val[kRows];
  int64_t ii[kRows];
  bool ok[kRows];
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    const int b = by * kRows + q + 1;
    const int odd = (a + b + color) & 1;  // the row's points of this colour have odd c
    const int c = 2 * j + (odd ? 1 : 2), h = c >> 1;
    ok[q] = b <= R.n1 && c <= R.n2;
    // (a slot outside the level reads slot 1 of row 1 instead)
    const int64_t i = ok[q] ? a * R.sx + b * R.sy + h : a * R.sx + R.sy + 1;
    const bool codd = ok[q] && (c & 1) != 0;
    ii[q] = i;
    const float nb = oth[i - R.sx] + oth[i + R.sx] + oth[i - R.sy] + oth[i + R.sy] + oth[i + (codd ? 0 : -1)] +
                     oth[i + (codd ? 1 : 0)];
    val[q] = w1 * own[i] + w6 * (nb + ro[i]);
  }
#pragma unroll
  for (int q = 0; q < kRows; ++q)
    if (ok[q]) own[ii[q]] = val[q];
}

// First sweep of a cycle from e = 0 (k_smooth_first2): slot h of a row holds
// one red and one black point.
__global__ void k_smooth_first2_rb(Rb R, float w1, float w6, const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int h = bx * blockDim.x + threadIdx.x;
  const int a = bz + 1;
  const float* r0 = R.r;
  const float* r1 = R.r + R.cs;
  float vr[kRows], vb[kRows];
  int64_t ir[kRows], ib[kRows];
  bool okr[kRows], okb[kRows];
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    const int b = by * kRows + q + 1;
    const int cr = 2 * h + ((a + b) & 1), cb = 2 * h + ((a + b + 1) & 1);
    okr[q] = b <= R.n1 && cr >= 1 && cr <= R.n2;
    okb[q] = b <= R.n1 && cb >= 1 && cb <= R.n2;
    ir[q] = okr[q] ? a * R.sx + b * R.sy + h : a * R.sx + R.sy + 1;
    ib[q] = okb[q] ? a * R.sx + b * R.sy + h : a * R.sx + R.sy + 1;
    vr[q] = w6 * r0[ir[q]];
    const int64_t i = ib[q];
    const bool codd = okb[q] && (cb & 1) != 0;
    const float nb = w6 * r0[i - R.sx] + w6 * r0[i + R.sx] + w6 * r0[i - R.sy] + w6 * r0[i + R.sy] +
                     w6 * r0[i + (codd ? 0 : -1)] + w6 * r0[i + (codd ? 1 : 0)];
    vb[q] = w1 * 0.0f + w6 * (nb + r1[i]);
  }
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    if (okr[q]) R.e[ir[q]] = vr[q];
    if (okb[q]) R.e[R.cs + ib[q]] = vb[q];
  }
}

// t = r - A e at every interior point, t in the one-array layout (k_resid_level).
// (one thread per row slot h, both colours of the slot: every load of a
// wavefront reads one contiguous run of a colour's array)
__global__ void k_resid_level_rb(float* __restrict__ t, Rb R, Lv L, const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int h = bx * blockDim.x + threadIdx.x;
  const int a = bz + 1;
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    const int b = by * kRows + q + 1;
    if (b > L.n1) break;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int c = 2 * h + ((a + b + k) & 1);
      if (c < 1 || c > L.n2) continue;
      const RbPt pt = rb_pt(R, a, b, c);
      t[a * L.sx + b * L.sy + c] = R.r[pt.own + pt.i] - rb_ae<float>(R, pt);
    }
  }
}

// e += trilinear interpolation of the coarse correction (k_prolong).
__global__ void k_prolong_rb(Rb R, const float* __restrict__ ec, Lv C, const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int c = bx * blockDim.x + threadIdx.x + 1;
  const int a = bz + 1;
  if (c > R.n2) return;
  const int a0 = a >> 1, a1 = (a + 1) >> 1, c0 = c >> 1, c1 = (c + 1) >> 1;
  auto E = [&](int x, int y, int z) { return ec[x * C.sx + y * C.sy + z]; };
#pragma unroll
  for (int q = 0; q < kRows; ++q) {
    const int b = by * kRows + q + 1;
    if (b > R.n1) break;
    const int b0 = b >> 1, b1 = (b + 1) >> 1;
    const float v = (E(a0, b0, c0) + E(a1, b0, c0) + E(a0, b1, c0) + E(a1, b1, c0) + E(a0, b0, c1) + E(a1, b0, c1) +
                     E(a0, b1, c1) + E(a1, b1, c1)) *
                    0.125f;
    R.e[rb_at(R, a, b, c)] += v;
  }
}

// The colour-split analogue of for_interior_rows: each workgroup takes one
// contiguous range of rows and walks their slots h (both colours of a slot
// in one thread); f(g, a, b, c) gets every interior point once.
template <class F>
__device__ inline void for_interior_slots(const Rb& R, const int64_t sx, const int64_t sy, F&& f) {
  const int nrows = R.n0 * R.n1, nh = (R.n2 >> 1) + 1;
  const int per = (nrows + (int)gridDim.x - 1) / (int)gridDim.x;
  const int r0 = xcd_block(blockIdx.x, gridDim.x) * per;
  const int r1 = min(r0 + per, nrows);
  const int tpr = nh <= 64 ? 64 : (nh <= 128 ? 128 : 256);  // threads per row
  const int sub = threadIdx.x / tpr, lane = threadIdx.x % tpr;
  for (int row = r0 + sub; row < r1; row += 256 / tpr) {
    const int a = row / R.n1 + 1, b = row % R.n1 + 1;
    const int64_t base = a * sx + b * sy;
    for (int h = lane; h < nh; h += tpr) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int c = 2 * h + ((a + b + k) & 1);
        if (c >= 1 && c <= R.n2) f(base + c, a, b, c);
      }
    }
  }
}

// The level-0 passes of the defect correction with e and r stored by colour
// (k_resid0, k_update, k_pair_resid); phi, rho and the FP64 residual keep the
// glue layout. The two that read and write FP64 arrays walk the points in
// row order (a slot walk would touch those arrays with a stride of two);
// the residual pass between the cycles of a pair walks the slots.
__global__ __launch_bounds__(256) void k_resid0_rb(const double* __restrict__ phi, const double* __restrict__ rho,
                                                   double rc, Rb R, bool store, Lv L, double* part,
                                                   const int* __restrict__ done, double* __restrict__ r64) {
  if (*done) return;
  __shared__ double s1[256], s2[256];
  double rr = 0.0, bb = 0.0;
  for_interior_rows(L.n0, L.n1, L.n2, L.sx, L.sy, [&](const int64_t g, int a, int b, int c) {
    const double bv = rc * rho[g];
    const double ap = 6.0 * phi[g] - (phi[g - L.sx] + phi[g + L.sx] + phi[g - L.sy] + phi[g + L.sy] + phi[g - 1] +
                                      phi[g + 1]);
    const double rv = bv - ap;
    if (store) {
      const RbPt pt = rb_pt(R, a, b, c);
      R.r[pt.own + pt.i] = (float)rv;
      r64[g] = rv;
    }
    rr += rv * rv;
    bb += bv * bv;
  });
  s1[threadIdx.x] = rr;
  s2[threadIdx.x] = bb;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

__global__ __launch_bounds__(256) void k_update_rb(double* __restrict__ phi, Rb R, double* r64, Lv L, double* part,
                                                   const int* __restrict__ done) {
  if (*done) return;
  __shared__ double s1[256];
  double rr = 0.0;
  const float* e = R.e;
  for_interior_rows(L.n0, L.n1, L.n2, L.sx, L.sy, [&](const int64_t g, int a, int b, int c) {
    const RbPt pt = rb_pt(R, a, b, c);
    const double eg = (double)e[pt.own + pt.i];
    const double ae = rb_ae<double>(R, pt);
    const double rv = r64[g] - ae;
    r64[g] = rv;
    R.r[pt.own + pt.i] = (float)rv;
    phi[g] += eg;
    rr += rv * rv;
  });
  s1[threadIdx.x] = rr;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s1[threadIdx.x] += s1[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

__global__ __launch_bounds__(256) void k_pair_resid_rb(Rb R, Lv L, double* part, const int* __restrict__ done) {
  if (*done) return;
  __shared__ double s1[256];
  double rr = 0.0;
  const float* e = R.e;
  for_interior_slots(R, L.sx, L.sy, [&](const int64_t, int a, int b, int c) {
    const RbPt pt = rb_pt(R, a, b, c);
    const float t = R.r[pt.own + pt.i] - rb_ae<float>(R, pt);
    rr += (double)t * (double)t;
  });
  s1[threadIdx.x] = rr;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s1[threadIdx.x] += s1[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

template <class P>
__global__ void k_zero_level(P* __restrict__ e, Lv L, const int* __restrict__ done) {
  if (*done) return;
  int bx, by, bz;
  block3(bx, by, bz);
  const int c = bx * blockDim.x + threadIdx.x + 1;
  const int b = by + 1, a = bz + 1;
  if (c > L.n2) return;
  e[a * L.sx + b * L.sy + c] = (P)0;
}

// One unknown: A = 6.
template <class P>
__global__ void k_coarse_single(P* e, const P* r, Lv L, const int* __restrict__ done) {
  if (*done) return;
  const int64_t g = L.sx + L.sy + 1;
  e[g] = r[g] / (P)6;
}

// ---- coarse tail: the V-cycle below level Lc in one workgroup, arrays in LDS

constexpr int kTailMaxLevels = 8;
constexpr int kTailThreads = 1024;

struct TailLv {
  int n0, n1, n2;
  int sx, sy;    // strides of the padded (n+2)^3 array
  int size;      // elements
  int oe, orr, ot;  // element offsets of e, r, t in LDS
};

template <class P>
struct TailArgs {
  int nl;
  TailLv lv[kTailMaxLevels];
  P* ge;                // level-Lc e in global memory
  const P* gr;          // level-Lc r in global memory
  int64_t gsx, gsy;     // its strides
  P w1, w6;
  int pre, post, coarse_sweeps;
  int total;            // LDS elements
  const int* done;
};

// Tail loops give each thread whole (b, c) columns of a level and walk a.
template <class P>
__device__ void tail_smooth(P* s, const TailLv& L, int sweeps, P w1, P w6) {
  const int ncol = L.n1 * L.n2;
  P* e = s + L.oe;
  const P* r = s + L.orr;
  for (int it = 0; it < sweeps; ++it)
    for (int col = 0; col < 2; ++col) {
      for (int t = threadIdx.x; t < ncol; t += kTailThreads) {
        const int b = t / L.n2 + 1, c = t % L.n2 + 1;
        const int a0 = (col ^ ((b + c) & 1)) ? 1 : 2;  // (a + b + c) % 2 == col
        for (int a = a0; a <= L.n0; a += 2) {
          const int g = a * L.sx + b * L.sy + c;
          const P nb = e[g - L.sx] + e[g + L.sx] + e[g - L.sy] + e[g + L.sy] + e[g - 1] + e[g + 1];
          e[g] = w1 * e[g] + w6 * (nb + r[g]);
        }
      }
      __syncthreads();
    }
}

template <class P>
__global__ __launch_bounds__(kTailThreads) void k_coarse_tail(TailArgs<P> A) {
  if (*A.done) return;
  extern __shared__ __attribute__((aligned(16))) char tail_lds[];
  P* s = reinterpret_cast<P*>(tail_lds);
  for (int q = threadIdx.x; q < A.total; q += kTailThreads) s[q] = (P)0;
  __syncthreads();
  {
    const TailLv& L = A.lv[0];
    for (int t = threadIdx.x; t < L.n1 * L.n2; t += kTailThreads) {
      const int b = t / L.n2 + 1, c = t % L.n2 + 1;
      for (int a = 1; a <= L.n0; ++a) s[L.orr + a * L.sx + b * L.sy + c] = A.gr[a * A.gsx + b * A.gsy + c];
    }
    __syncthreads();
  }
  // descend
  for (int l = 0; l + 1 < A.nl; ++l) {
    const TailLv& F = A.lv[l];
    const TailLv& C = A.lv[l + 1];
    tail_smooth<P>(s, F, A.pre, A.w1, A.w6);
    const P* e = s + F.oe;
    for (int t = threadIdx.x; t < F.n1 * F.n2; t += kTailThreads) {
      const int b = t / F.n2 + 1, c = t % F.n2 + 1;
      for (int a = 1; a <= F.n0; ++a) {
        const int g = a * F.sx + b * F.sy + c;
        s[F.ot + g] = s[F.orr + g] - ((P)6 * e[g] - (e[g - F.sx] + e[g + F.sx] + e[g - F.sy] + e[g + F.sy] +
                                                      e[g - 1] + e[g + 1]));
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < C.n1 * C.n2; t += kTailThreads) {
      const int b = t / C.n2 + 1, c = t % C.n2 + 1;
      for (int a = 1; a <= C.n0; ++a) {
        const int f = (2 * a) * F.sx + (2 * b) * F.sy + 2 * c;
        P acc = 0;
        for (int da = -1; da <= 1; ++da)
          for (int db = -1; db <= 1; ++db)
            for (int dc = -1; dc <= 1; ++dc)
              acc += fw_weight<P>(da, db, dc) * s[F.ot + f + da * F.sx + db * F.sy + dc];
        const int gc = a * C.sx + b * C.sy + c;
        s[C.orr + gc] = acc * (P)0.5;
        s[C.oe + gc] = (P)0;
      }
    }
    __syncthreads();
  }
  // coarsest level
  {
    const TailLv& L = A.lv[A.nl - 1];
    if (L.n0 == 1 && L.n1 == 1 && L.n2 == 1) {
      if (threadIdx.x == 0) {
        const int g = L.sx + L.sy + 1;
        s[L.oe + g] = s[L.orr + g] / (P)6;
      }
      __syncthreads();
    } else {
      tail_smooth<P>(s, L, A.coarse_sweeps, (P)0, (P)1 / (P)6);
    }
  }
  // ascend
  for (int l = A.nl - 2; l >= 0; --l) {
    const TailLv& F = A.lv[l];
    const TailLv& C = A.lv[l + 1];
    const P* ec = s + C.oe;
    for (int t = threadIdx.x; t < F.n1 * F.n2; t += kTailThreads) {
      const int b = t / F.n2 + 1, c = t % F.n2 + 1;
      const int b0 = b >> 1, b1 = (b + 1) >> 1, c0 = c >> 1, c1 = (c + 1) >> 1;
      for (int a = 1; a <= F.n0; ++a) {
        const int a0 = a >> 1, a1 = (a + 1) >> 1;
        auto E = [&](int x, int y, int z) { return ec[x * C.sx + y * C.sy + z]; };
        const P v = (E(a0, b0, c0) + E(a1, b0, c0) + E(a0, b1, c0) + E(a1, b1, c0) + E(a0, b0, c1) +
                     E(a1, b0, c1) + E(a0, b1, c1) + E(a1, b1, c1)) *
                    (P)0.125;
        s[F.oe + a * F.sx + b * F.sy + c] += v;
      }
    }
    __syncthreads();
    tail_smooth<P>(s, F, A.post, A.w1, A.w6);
  }
  {
    const TailLv& L = A.lv[0];
    for (int t = threadIdx.x; t < L.n1 * L.n2; t += kTailThreads) {
      const int b = t / L.n2 + 1, c = t % L.n2 + 1;
      for (int a = 1; a <= L.n0; ++a) A.ge[a * A.gsx + b * A.gsy + c] = s[L.oe + a * L.sx + b * L.sy + c];
    }
  }
}

// The same V-cycle for the cubic chain 15 -> 7 -> 3 -> 1 (every grid of
// 2^k + 1 nodes per axis ends in it), sizes fixed at compile time: index maps
// become multiplies and each half-sweep visits only its own colour.
template <int N>
struct TailDim {
  static constexpr int S = N + 2, SX = S * S, SY = S, SIZE = S * S * S, H = (N + 1) / 2;
};

template <class P, int N>
__device__ void c_smooth(P* e, const P* r, int sweeps, P w1, P w6) {
  using D = TailDim<N>;
  for (int it = 0; it < sweeps; ++it)
    for (int col = 0; col < 2; ++col) {
      for (int q = threadIdx.x; q < N * N * D::H; q += kTailThreads) {
        const int a = q / (N * D::H) + 1, b = (q / D::H) % N + 1, m = q % D::H;
        const int c = (((col + a + b) & 1) ? 1 : 2) + 2 * m;
        if (c <= N) {
          const int g = a * D::SX + b * D::SY + c;
          const P nb = e[g - D::SX] + e[g + D::SX] + e[g - D::SY] + e[g + D::SY] + e[g - 1] + e[g + 1];
          e[g] = w1 * e[g] + w6 * (nb + r[g]);
        }
      }
      __syncthreads();
    }
}

template <class P, int N>
__device__ void c_cycle(P* s, P w1, P w6, int pre, int post) {
  using D = TailDim<N>;
  P* e = s;
  P* r = s + D::SIZE;
  if constexpr (N == 1) {
    if (threadIdx.x == 0) e[D::SX + D::SY + 1] = r[D::SX + D::SY + 1] / (P)6;
    __syncthreads();
  } else {
    constexpr int NC = (N - 1) / 2;
    using C = TailDim<NC>;
    P* t = s + 2 * D::SIZE;
    P* ec = s + 3 * D::SIZE;
    P* rc = ec + C::SIZE;
    c_smooth<P, N>(e, r, pre, w1, w6);
    for (int q = threadIdx.x; q < N * N * N; q += kTailThreads) {
      const int a = q / (N * N) + 1, b = (q / N) % N + 1, c = q % N + 1;
      const int g = a * D::SX + b * D::SY + c;
      t[g] = r[g] - ((P)6 * e[g] - (e[g - D::SX] + e[g + D::SX] + e[g - D::SY] + e[g + D::SY] + e[g - 1] + e[g + 1]));
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NC * NC * NC; q += kTailThreads) {
      const int a = q / (NC * NC) + 1, b = (q / NC) % NC + 1, c = q % NC + 1;
      const int f = (2 * a) * D::SX + (2 * b) * D::SY + 2 * c;
      P acc = 0;
#pragma unroll
      for (int da = -1; da <= 1; ++da)
#pragma unroll
        for (int db = -1; db <= 1; ++db)
#pragma unroll
          for (int dc = -1; dc <= 1; ++dc) acc += fw_weight<P>(da, db, dc) * t[f + da * D::SX + db * D::SY + dc];
      const int gc = a * C::SX + b * C::SY + c;
      rc[gc] = acc * (P)0.5;
      ec[gc] = (P)0;
    }
    __syncthreads();
    c_cycle<P, NC>(ec, w1, w6, pre, post);
    for (int q = threadIdx.x; q < N * N * N; q += kTailThreads) {
      const int a = q / (N * N) + 1, b = (q / N) % N + 1, c = q % N + 1;
      const int a0 = a >> 1, a1 = (a + 1) >> 1, b0 = b >> 1, b1 = (b + 1) >> 1, c0 = c >> 1, c1 = (c + 1) >> 1;
      auto E = [&](int x, int y, int z) { return ec[x * C::SX + y * C::SY + z]; };
      const P v = (E(a0, b0, c0) + E(a1, b0, c0) + E(a0, b1, c0) + E(a1, b1, c0) + E(a0, b0, c1) + E(a1, b0, c1) +
                   E(a0, b1, c1) + E(a1, b1, c1)) *
                  (P)0.125;
      e[a * D::SX + b * D::SY + c] += v;
    }
    __syncthreads();
    c_smooth<P, N>(e, r, post, w1, w6);
  }
}

template <int N>
constexpr size_t chain_elems() {
  if constexpr (N == 1)
    return 2 * TailDim<1>::SIZE;
  else
    return 3 * TailDim<N>::SIZE + chain_elems<(N - 1) / 2>();
}

template <class P>
__global__ __launch_bounds__(kTailThreads) void k_coarse_tail15(TailArgs<P> A) {
  if (*A.done) return;
  extern __shared__ __attribute__((aligned(16))) char tail_lds[];
  P* s = reinterpret_cast<P*>(tail_lds);
  using D = TailDim<15>;
  constexpr int total = (int)chain_elems<15>();
  for (int q = threadIdx.x; q < total; q += kTailThreads) s[q] = (P)0;
  __syncthreads();
  for (int q = threadIdx.x; q < 15 * 15 * 15; q += kTailThreads) {
    const int a = q / 225 + 1, b = (q / 15) % 15 + 1, c = q % 15 + 1;
    s[D::SIZE + a * D::SX + b * D::SY + c] = A.gr[a * A.gsx + b * A.gsy + c];
  }
  __syncthreads();
  c_cycle<P, 15>(s, A.w1, A.w6, A.pre, A.post);
  for (int q = threadIdx.x; q < 15 * 15 * 15; q += kTailThreads) {
    const int a = q / 225 + 1, b = (q / 15) % 15 + 1, c = q % 15 + 1;
    A.ge[a * A.gsx + b * A.gsy + c] = s[a * D::SX + b * D::SY + c];
  }
}

// Face conditions on phi (Field.cpp:10-96): D -> 0, N -> copy the neighbour.
__global__ void k_face(double* phi, int cx, int cy, int cz, int axis, int side, int neumann) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  int nu, nv;
  if (axis == 0) {
    nu = cz - 2;
    nv = cy - 2;
  } else if (axis == 1) {
    nu = cz - 2;
    nv = cx - 2;
  } else {
    nu = cy - 2;
    nv = cx - 2;
  }
  if (u >= nu || v >= nv) return;
  const int nmax = axis == 0 ? cx : (axis == 1 ? cy : cz);
  const int bi = side == 0 ? 1 : nmax - 2, ni = side == 0 ? 2 : nmax - 3;
  int64_t t, s;
  const int64_t sI = (int64_t)cy * cz, sJ = cz;
  if (axis == 0) {
    t = bi * sI + (v + 1) * sJ + (u + 1);
    s = ni * sI + (v + 1) * sJ + (u + 1);
  } else if (axis == 1) {
    t = (v + 1) * sI + bi * sJ + (u + 1);
    s = (v + 1) * sI + ni * sJ + (u + 1);
  } else {
    t = (v + 1) * sI + (u + 1) * sJ + bi;
    s = (v + 1) * sI + (u + 1) * sJ + ni;
  }
  phi[t] = neumann ? phi[s] : 0.0;
}

Lv lv_of(const MgLevel& m) { return Lv{m.n[0], m.n[1], m.n[2], m.sx, m.sy}; }

template <class P>
P* E_(MgLevel& m);
template <>
double* E_<double>(MgLevel& m) {
  return m.e64.p + m.off;
}
template <>
float* E_<float>(MgLevel& m) {
  return m.e32.p + m.off;
}
template <class P>
P* R_(MgLevel& m);
template <>
double* R_<double>(MgLevel& m) {
  return m.r64.p + m.off;
}
template <>
float* R_<float>(MgLevel& m) {
  return m.r32.p + m.off;
}
template <class P>
P* T_(MgLevel& m);
template <>
double* T_<double>(MgLevel& m) {
  return m.t64.p + m.off;
}
template <>
float* T_<float>(MgLevel& m) {
  return m.t32.p + m.off;
}

dim3 grid_full(const MgLevel& m) { return dim3(div_up(m.n[2], 64), m.n[1], m.n[0]); }
// kRows rows per thread
dim3 grid_full4(const MgLevel& m) { return dim3(div_up(m.n[2], 64), div_up(m.n[1], kRows), m.n[0]); }
dim3 grid_half4(const MgLevel& m) { return dim3(div_up((m.n[2] + 1) / 2, 64), div_up(m.n[1], kRows), m.n[0]); }

constexpr int kResidBlocks = 2048;  // workgroups of the level-0 residual and update passes

// Levels with at least this many unknowns form their residual in a pass of
// its own before restricting it (below, one launch forming it on the fly costs less).
constexpr int64_t kSplitRestrict = (int64_t)1 << 20;

int coarse_sweeps(const MgLevel& m) { return 4 * std::max({m.n[0], m.n[1], m.n[2]}) + 8; }

// First level handled by the LDS tail kernel and its LDS bytes, for element size `esz`.
void plan_tail(Ctx& c, size_t esz) {
  const int nl = (int)c.mg.size();
  c.mg_tail = nl;  // none
  c.mg_tail_bytes = 0;
  size_t bytes = 0;
  for (int l = nl - 1; l >= 0; --l) {
    const MgLevel& m = c.mg[l];
    // the tail keeps its levels as padded (n+2)^3 arrays
    const size_t lsz = (size_t)(m.n[0] + 2) * (m.n[1] + 2) * (m.n[2] + 2);
    const size_t lb = 3 * lsz * esz;
    if (bytes + lb > c.lds_max || nl - l > kTailMaxLevels) break;
    bytes += lb;
    c.mg_tail = l;
    c.mg_tail_bytes = bytes;
  }
  // the compile-time 15 -> 7 -> 3 -> 1 kernel when the tail is exactly that chain
  c.mg_tail15 = false;
  if (c.mg_tail < nl && nl - c.mg_tail == 4) {
    bool ok = true;
    const int want[4] = {15, 7, 3, 1};
    for (int q = 0; q < 4; ++q)
      for (int a = 0; a < 3; ++a) ok = ok && c.mg[c.mg_tail + q].n[a] == want[q];
    c.mg_tail15 = ok;
  }
}

template <class P>
TailArgs<P> tail_args(Ctx& c, const int* done) {
  TailArgs<P> A{};
  const int l0 = c.mg_tail;
  A.nl = (int)c.mg.size() - l0;
  int off = 0;
  for (int q = 0; q < A.nl; ++q) {
    const MgLevel& m = c.mg[l0 + q];
    TailLv& L = A.lv[q];
    L.n0 = m.n[0];
    L.n1 = m.n[1];
    L.n2 = m.n[2];
    L.sy = m.n[2] + 2;
    L.sx = (m.n[1] + 2) * (m.n[2] + 2);
    L.size = (m.n[0] + 2) * L.sx;
    L.oe = off;
    L.orr = off + L.size;
    L.ot = off + 2 * L.size;
    off += 3 * L.size;
  }
  MgLevel& m0 = c.mg[l0];
  A.ge = E_<P>(m0);
  A.gr = R_<P>(m0);
  A.gsx = m0.sx;
  A.gsy = m0.sy;
  A.w1 = (P)(1.0 - c.d.mg_omega);
  A.w6 = (P)(c.d.mg_omega / 6.0);
  A.pre = c.d.mg_pre;
  A.post = c.d.mg_post;
  A.coarse_sweeps = coarse_sweeps(c.mg.back());
  A.total = off;
  A.done = done;
  return A;
}

Rb rb_of(Ctx& c) {
  const MgLevel& m = c.mg[0];
  return Rb{c.erb.p, c.rrb.p, c.rb_cs, c.rb_sx, c.rb_sy, m.n[0], m.n[1], m.n[2]};
}

// the colour-split finest level, kRows rows per thread: the row slots h that
// hold interior points (0 .. n2 / 2) for the first sweep, which takes both
// colours of a slot; a half-sweep takes one colour's points (grid_half4)
dim3 grid_rb(const Ctx& c) {
  const MgLevel& m = c.mg[0];
  return dim3(div_up((m.n[2] >> 1) + 1, 64), div_up(m.n[1], kRows), m.n[0]);
}

template <class P>
void smooth(Ctx& c, MgLevel& m, int sweeps, double w, bool from_zero, const int* done) {
  const P w1 = (P)(1.0 - w), w6 = (P)(w / 6.0);
  if constexpr (std::is_same<P, float>::value) {
    if (c.mg_rb && &m == &c.mg[0]) {
      const Rb R = rb_of(c);
      for (int s = 0; s < sweeps; ++s) {
        if (from_zero && s == 0) {
          k_smooth_first2_rb<<<grid_rb(c), 64, 0, c.stream>>>(R, w1, w6, done);
          continue;
        }
        for (int col = 0; col < 2; ++col) k_smooth_rb<<<grid_half4(m), 64, 0, c.stream>>>(R, col, w1, w6, done);
      }
      return;
    }
  }
  for (int s = 0; s < sweeps; ++s) {
    if (from_zero && s == 0) {
      k_smooth_first2<P><<<grid_half4(m), 64, 0, c.stream>>>(E_<P>(m), R_<P>(m), lv_of(m), w1, w6, done);
      continue;
    }
    for (int col = 0; col < 2; ++col)
      k_smooth<P><<<grid_half4(m), 64, 0, c.stream>>>(E_<P>(m), R_<P>(m), lv_of(m), col, w1, w6, done);
  }
}

// One V-cycle from level l down. zero: the level's correction starts from
// e = 0 (always so below the finest level); otherwise the finest level's
// cycle continues from the correction already in e (a second cycle on the
// same right-hand side).
template <class P>
void vcycle(Ctx& c, int l, const int* done, bool zero = true) {
  MgLevel& m = c.mg[l];
  const double w = c.d.mg_omega;
  if (l == c.mg_tail) {
    TailArgs<P> A = tail_args<P>(c, done);
    if (c.mg_tail15)
      k_coarse_tail15<P><<<1, kTailThreads, c.mg_tail_bytes, c.stream>>>(A);
    else
      k_coarse_tail<P><<<1, kTailThreads, c.mg_tail_bytes, c.stream>>>(A);
    return;
  }
  if (l == (int)c.mg.size() - 1) {
    // coarsest level outside the tail kernel: Gauss-Seidel sweeps from zero
    if (m.n[0] == 1 && m.n[1] == 1 && m.n[2] == 1) {
      k_coarse_single<P><<<1, 1, 0, c.stream>>>(E_<P>(m), R_<P>(m), lv_of(m), done);
    } else {
      smooth<P>(c, m, coarse_sweeps(m), 1.0, true, done);
    }
    return;
  }
  MgLevel& mc = c.mg[l + 1];
  // every level starts its cycle from e = 0 (the correction equation)
  if (zero && c.d.mg_pre == 0) k_zero_level<P><<<grid_full(m), 64, 0, c.stream>>>(E_<P>(m), lv_of(m), done);
  smooth<P>(c, m, c.d.mg_pre, w, zero, done);
  bool rb = false;
  if constexpr (std::is_same<P, float>::value) rb = c.mg_rb && l == 0;
  if (rb) {
    if constexpr (std::is_same<P, float>::value) {
      k_resid_level_rb<<<grid_rb(c), 64, 0, c.stream>>>(T_<P>(m), rb_of(c), lv_of(m), done);
      k_restrict_t<P><<<grid_full(mc), 64, 0, c.stream>>>(R_<P>(mc), lv_of(mc), T_<P>(m), lv_of(m), done);
    }
  } else if ((int64_t)m.n[0] * m.n[1] * m.n[2] >= kSplitRestrict) {
    k_resid_level<P><<<grid_full4(m), 64, 0, c.stream>>>(T_<P>(m), E_<P>(m), R_<P>(m), lv_of(m), done);
    k_restrict_t<P><<<grid_full(mc), 64, 0, c.stream>>>(R_<P>(mc), lv_of(mc), T_<P>(m), lv_of(m), done);
  } else {
    k_restrict_resid<P><<<grid_full(mc), 64, 0, c.stream>>>(R_<P>(mc), lv_of(mc), E_<P>(m), R_<P>(m), lv_of(m),
                                                            done);
  }
  vcycle<P>(c, l + 1, done);
  if (rb) {
    if constexpr (std::is_same<P, float>::value)
      k_prolong_rb<<<grid_full4(m), 64, 0, c.stream>>>(rb_of(c), E_<P>(mc), lv_of(mc), done);
  } else {
    k_prolong<P><<<grid_full4(m), 64, 0, c.stream>>>(E_<P>(m), lv_of(m), E_<P>(mc), lv_of(mc), done);
  }
  smooth<P>(c, m, c.d.mg_post, w, false, done);
}

template <class P>
int solve_impl(Ctx& c, double tol, int max_cycles, double* relres_out, const double* rho0) {
  MgLevel& m0 = c.mg[0];
  const Lv L0 = lv_of(m0);
  // level-0 arrays share the glue layout; the MG origin is glue node (1,1,1)
  const int64_t off = m0.off;
  double* phi = c.phi.p + off;
  const double* rho = rho0 + off;
  P* r0 = R_<P>(m0);
  const double rc = 1.0 / c.d.eps0;  // pow(dx, 2) / eps0 with dx = 1
  const int nb = kResidBlocks;  // workgroups of the level-0 residual passes (8 per CU)
  double* part = c.mg_part.p;
  double* dst = c.mg_part.p + 2 * nb;
  int* st = c.mg_state.p;
  const int* done = st;
  if (c.mg_tail_esz != sizeof(P)) {
    plan_tail(c, sizeof(P));
    c.mg_tail_esz = sizeof(P);
    if (c.mg_tail < (int)c.mg.size() && c.mg_tail_bytes > 65536) {
      FOPIS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_coarse_tail<P>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.mg_tail_bytes));
      FOPIS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_coarse_tail15<P>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.mg_tail_bytes));
    }
  }

  // level-0 FP64 residual kept by the cycle updates (the level's own array when P is double)
  double* r64 = std::is_same<P, double>::value ? reinterpret_cast<double*>(r0) : c.mg_r64.p + off;
  const int* never = st + 3;  // stays 0: the final residual always runs
  k_solve_begin<<<1, 1, 0, c.stream>>>(st);
  // the level-0 passes, on the colour-split arrays when the level has them
  bool rb = false;
  if constexpr (std::is_same<P, float>::value) rb = c.mg_rb;
  auto resid0 = [&](bool store, const int* flag) {
    if constexpr (std::is_same<P, float>::value) {
      if (rb) {
        k_resid0_rb<<<nb, 256, 0, c.stream>>>(phi, rho, rc, rb_of(c), store, L0, part, flag, store ? r64 : nullptr);
        return;
      }
    }
    k_resid0<P><<<nb, 256, 0, c.stream>>>(phi, rho, rc, store ? r0 : (P*)nullptr, L0, part, flag,
                                          store ? r64 : (double*)nullptr);
  };
  resid0(true, done);
  k_norm<<<1, 256, 0, c.stream>>>(part, nb, dst, st, tol, max_cycles, 0, 0);
  k_keep_start<<<1, 1, 0, c.stream>>>(dst);
  // Queue as many cycles as the previous two solves needed, then more. With
  // pairs, two V-cycles share one FP64 update and norm check: the second
  // runs on the same right-hand side, continuing from the first's correction
  // (the residual it restricts is then r - A e, formed in P, whose rounding
  // stays far below the ~30x a V-cycle removes). Between the two, one read of
  // e and r gives the residual the first cycle left (k_pair_resid); when that
  // already meets the tolerance the second cycle's kernels all return at
  // once, so a solve applies exactly the cycles it needs, as with single
  // cycles. A pair saves the FP64 update (reads and writes phi and r in FP64)
  // and a norm check for the price of that one read.
  const int pred = std::max(1, std::min(c.mg_hist[0], c.mg_hist[1]));
  // (a finest level inside the single-workgroup tail, or the coarsest level
  // itself, always solves from zero: no pairs there)
  const bool pairs = c.mg_pairs && c.mg_tail > 0 && (int)c.mg.size() > 1;
  int npair = pairs ? pred / 2 : 0;
  int batch = pred - 2 * npair;
  int32_t hs[7];
  double hd[4];
  int* skip = st + 4;
  // One cycle (or pair): the V-cycle(s), the FP64 update and the norm
  // check. Their ~30 (~60) launches are captured once into a graph of kernel
  // nodes (every pointer and size in them is fixed for the context; the
  // tolerance and the cycle cap are arguments of the norm kernel, so the
  // graphs are rebuilt when they change) and replayed.
  auto cycle = [&](int nv) {
    vcycle<P>(c, 0, done);
    if (nv == 2) {
      if constexpr (std::is_same<P, float>::value) {
        if (rb)
          k_pair_resid_rb<<<nb, 256, 0, c.stream>>>(rb_of(c), L0, part, done);
        else
          k_pair_resid<P><<<nb, 256, 0, c.stream>>>(E_<P>(m0), r0, L0, part, done);
      } else {
        k_pair_resid<P><<<nb, 256, 0, c.stream>>>(E_<P>(m0), r0, L0, part, done);
      }
      k_pair_check<<<1, 256, 0, c.stream>>>(part, nb, dst, st, tol, skip);
      vcycle<P>(c, 0, skip, false);
    }
    if constexpr (std::is_same<P, float>::value) {
      if (rb)
        k_update_rb<<<nb, 256, 0, c.stream>>>(phi, rb_of(c), r64, L0, part, done);
      else
        k_update<P><<<nb, 256, 0, c.stream>>>(phi, E_<P>(m0), r64, r0, L0, part, done);
    } else {
      k_update<P><<<nb, 256, 0, c.stream>>>(phi, E_<P>(m0), r64, r0, L0, part, done);
    }
    k_norm<<<1, 256, 0, c.stream>>>(part, nb, dst, st, tol, max_cycles, nv, 1, nv == 2 ? skip : nullptr);
  };
  if (c.mg_use_graph && (!c.mg_graph[0] || c.mg_graph_esz != sizeof(P) || c.mg_graph_tol != tol ||
                         c.mg_graph_cap != max_cycles)) {
    for (int k = 0; k < 2; ++k) {
      if (c.mg_graph[k]) {
        FOPIS_HIP(hipGraphExecDestroy(c.mg_graph[k]));
        c.mg_graph[k] = nullptr;
      }
      hipGraph_t g = nullptr;
      FOPIS_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      cycle(k + 1);
      const hipError_t ec = hipStreamEndCapture(c.stream, &g);
      if (ec != hipSuccess || !g)
        throw Error(FOPIS_E_DEVICE, std::string("V-cycle capture: ") + hipGetErrorString(ec));
      const hipError_t ei = hipGraphInstantiate(&c.mg_graph[k], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        c.mg_graph[k] = nullptr;
        throw Error(FOPIS_E_DEVICE, std::string("V-cycle graph: ") + hipGetErrorString(ei));
      }
    }
    c.mg_graph_esz = sizeof(P);
    c.mg_graph_tol = tol;
    c.mg_graph_cap = max_cycles;
  }
  auto launch = [&](int nv) {
    if (c.mg_use_graph)
      FOPIS_HIP(hipGraphLaunch(c.mg_graph[nv - 1], c.stream));
    else
      cycle(nv);
  };
  for (;;) {
    for (int q = 0; q < npair; ++q) launch(2);
    for (int q = 0; q < batch; ++q) launch(1);
    if (pairs) {  // later rounds: one pair (up to two cycles) at a time
      npair = 1;
      batch = 0;
    } else {
      npair = 0;
      batch = 2;
    }
    FOPIS_HIP(hipMemcpyAsync(hs, st, sizeof(hs), hipMemcpyDeviceToHost, c.stream));
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    if (!hs[0]) continue;
    // confirm on the true residual b - A phi before stopping
    resid0(false, never);
    k_norm_final<<<1, 256, 0, c.stream>>>(part, nb, dst);
    FOPIS_HIP(hipMemcpyAsync(hd, dst, sizeof(hd), hipMemcpyDeviceToHost, c.stream));
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    const bool zero_rhs = hd[1] == 0.0 && hd[0] < 1e-300;
    if (hd[2] <= tol || zero_rhs || hs[1] >= max_cycles) break;
    // the recursively updated residual has drifted from the true one: the
    // further cycles start from the true residual
    resid0(true, never);
    k_reopen<<<1, 1, 0, c.stream>>>(st);
  }
  const int cycles = hs[1];
  c.mg_updates_last = hs[5];
  c.mg_pair_passes_last = hs[6];
  c.mg_hist[1] = c.mg_hist[0];
  c.mg_hist[0] = cycles;
  // face conditions in the reference's order: x low/up, y, z
  for (int axis = 0; axis < 3; ++axis)
    for (int side = 0; side < 2; ++side) {
      const char bcc = c.d.boundary[2 * axis + side];
      if (bcc == 'P') continue;
      int nu, nv;
      if (axis == 0) {
        nu = c.cz - 2;
        nv = c.cy - 2;
      } else if (axis == 1) {
        nu = c.cz - 2;
        nv = c.cx - 2;
      } else {
        nu = c.cy - 2;
        nv = c.cx - 2;
      }
      k_face<<<dim3(div_up(nu, 64), nv), 64, 0, c.stream>>>(c.phi.p, c.cx, c.cy, c.cz, axis, side, bcc == 'N');
    }
  FOPIS_HIP(hipGetLastError());
  *relres_out = hd[2];
  c.relres0_last = hd[3];
  return cycles;
}

}  // namespace

void poisson_init(Ctx& c) {
  c.mg.clear();
  int n[3] = {c.nx - 2, c.ny - 2, c.nz - 2};
  if (n[0] < 1 || n[1] < 1 || n[2] < 1) throw Error(FOPIS_E_ARG, "grid too small for the Poisson solve");
  const bool f32 = c.d.mg_precision == FOPIS_FP32;
  c.mg.reserve(16);
  for (int l = 0;; ++l) {
    c.mg.emplace_back();
    MgLevel& m = c.mg.back();
    for (int q = 0; q < 3; ++q) m.n[q] = n[q];
    if (l == 0) {
      m.sx = (int64_t)c.cy * c.cz;
      m.sy = c.cz;
      m.size = c.nnode;
      m.off = (int64_t)c.id3(1, 1, 1);
    } else {
      m.sx = (int64_t)(n[1] + 2) * (n[2] + 2);
      m.sy = n[2] + 2;
      m.size = (size_t)(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
    }
    if (f32) {
      m.e32.alloc(m.size);
      m.r32.alloc(m.size);
      m.t32.alloc(m.size);
      FOPIS_HIP(hipMemsetAsync(m.e32.p, 0, m.size * sizeof(float), c.stream));
      FOPIS_HIP(hipMemsetAsync(m.r32.p, 0, m.size * sizeof(float), c.stream));
      FOPIS_HIP(hipMemsetAsync(m.t32.p, 0, m.size * sizeof(float), c.stream));
    } else {
      m.e64.alloc(m.size);
      m.r64.alloc(m.size);
      m.t64.alloc(m.size);
      FOPIS_HIP(hipMemsetAsync(m.e64.p, 0, m.size * sizeof(double), c.stream));
      FOPIS_HIP(hipMemsetAsync(m.r64.p, 0, m.size * sizeof(double), c.stream));
      FOPIS_HIP(hipMemsetAsync(m.t64.p, 0, m.size * sizeof(double), c.stream));
    }
    bool can = true;
    for (int q = 0; q < 3; ++q) can = can && (n[q] >= 3) && (n[q] % 2 == 1);
    if (!can) break;
    for (int q = 0; q < 3; ++q) n[q] = (n[q] - 1) / 2;
  }
  if (f32) {
    c.mg_r64.alloc(c.nnode);
    FOPIS_HIP(hipMemsetAsync(c.mg_r64.p, 0, c.nnode * sizeof(double), c.stream));
  }
  // the finest level stored by colour: FP32 levels, a level large enough to
  // form its residual in a pass of its own (and so never inside the LDS tail),
  // pre-smoothing (its first sweep starts from zero); FOPIS_MG_RB=0 keeps one array
  const char* rbs = std::getenv("FOPIS_MG_RB");
  const MgLevel& l0 = c.mg[0];
  c.mg_rb = f32 && !(rbs && rbs[0] == '0') && c.mg.size() > 1 && c.d.mg_pre > 0 &&
            (int64_t)l0.n[0] * l0.n[1] * l0.n[2] >= kSplitRestrict;
  if (c.mg_rb) {
    c.rb_sy = (l0.n[2] + 3) / 2;
    c.rb_sx = (int64_t)(l0.n[1] + 2) * c.rb_sy;
    c.rb_cs = (int64_t)(l0.n[0] + 2) * c.rb_sx;
    c.erb.alloc(2 * c.rb_cs);
    c.rrb.alloc(2 * c.rb_cs);
    FOPIS_HIP(hipMemsetAsync(c.erb.p, 0, 2 * c.rb_cs * sizeof(float), c.stream));
    FOPIS_HIP(hipMemsetAsync(c.rrb.p, 0, 2 * c.rb_cs * sizeof(float), c.stream));
  }
  c.mg_state.alloc(8);
  FOPIS_HIP(hipMemsetAsync(c.mg_state.p, 0, 8 * sizeof(int32_t), c.stream));
  c.mg_part.alloc(2 * kResidBlocks + 4);
  const char* g = std::getenv("FOPIS_MG_GRAPH");
  c.mg_use_graph = !(g && g[0] == '0');
  const char* pr2 = std::getenv("FOPIS_MG_PAIRS");
  c.mg_pairs = !(pr2 && pr2[0] == '0');
  const char* cy = std::getenv("FOPIS_CARRY");
  c.carry = !(cy && cy[0] == '0');
  int dev = 0;
  FOPIS_HIP(hipGetDevice(&dev));
  hipDeviceProp_t pr;
  FOPIS_HIP(hipGetDeviceProperties(&pr, dev));
  c.lds_max = pr.sharedMemPerBlock;
  c.mg_tail_esz = 0;
  c.mg_hist[0] = c.mg_hist[1] = 1;
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

void poisson_free(Ctx& c) {
  for (auto& g : c.mg_graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

int poisson_solve(Ctx& c, double tol, int max_cycles, double* relres, const double* rhs_override) {
  Phase ph(c, "poisson");
  const double* rho0 = rhs_override ? rhs_override : c.rho.p;
  int cyc;
  if (c.d.mg_precision == FOPIS_FP32)
    cyc = solve_impl<float>(c, tol, max_cycles, relres, rho0);
  else
    cyc = solve_impl<double>(c, tol, max_cycles, relres, rho0);
  c.mg_cycles_last = cyc;
  c.mg_cycles_total += cyc;
  c.mg_solves_total += 1;
  c.mg_updates_total += c.mg_updates_last;
  c.mg_pair_passes_total += c.mg_pair_passes_last;
  c.relres_last = *relres;
  return cyc;
}

}  // namespace fopis
