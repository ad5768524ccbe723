// Conductors by the capacitance-matrix method, FP64 on the device.
//
// Reference behaviour followed (hsimyu/fopis):
//   construct / cell map       src/Spacecraft/Spacecraft.cpp:16-152
//   capacity matrix setup      src/Grid/RootGrid/RootGridObject.cpp:77-148 (B by unit-charge solves, C = B^-1)
//   sumWholeCharge             src/Spacecraft/Spacecraft.cpp:615-626
//   getTotalCharge             src/Spacecraft/Spacecraft.cpp:248-259
//   perfect-conductor charge   src/Spacecraft/Spacecraft.cpp:708-758
//   Cmatrix::total             include/Spacecraft/Cmatrix.hpp:36-44
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "fopis_internal.hpp"

#pragma clang fp contract(off)

namespace fopis {

namespace {

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_set_node_cmat(int32_t* node_cmat, const int64_t* cmat_node, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) node_cmat[cmat_node[i]] = i;
}

__global__ void k_unit_rhs(double* rho0, int64_t nn, int64_t node) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) rho0[i] = (i == node) ? 1.0 : 0.0;
}

__global__ void k_zero_d(double* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// B(col,row) = phi at the row node for the unit charge at the col node
__global__ void k_gather_col(double* B, int n, int col, const double* phi, const int64_t* cmat_node) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row < n) B[(size_t)col * n + row] = phi[cmat_node[row]];
}

__global__ void k_identity(double* A, int n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (int64_t)n * n) A[i] = (i / n == i % n) ? 1.0 : 0.0;
}

// Gauss-Jordan with partial pivoting, step c: pivot search (one block).
__global__ __launch_bounds__(1024) void k_gj_pivot(const double* A, int n, int c, int* piv, double* pval) {
  __shared__ double bv[1024];
  __shared__ int bi[1024];
  double best = -1.0;
  int arg = c;
  for (int r = c + threadIdx.x; r < n; r += 1024) {
    const double v = fabs(A[(size_t)r * n + c]);
    if (v > best) {
      best = v;
      arg = r;
    }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = arg;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double v2 = bv[threadIdx.x + o];
      const int i2 = bi[threadIdx.x + o];
      // largest magnitude, first row on ties (a sequential scan's choice)
      if (v2 > bv[threadIdx.x] || (v2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
        bv[threadIdx.x] = v2;
        bi[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *piv = bi[0];
    *pval = A[(size_t)bi[0] * n + c];
  }
}

// swap rows c and piv, then divide row c by the pivot value
__global__ void k_gj_swap_scale(double* A, double* I, int n, int c, const int* piv, const double* pval) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int p = *piv;
  const double d = *pval;
  double a = A[(size_t)p * n + q], b = I[(size_t)p * n + q];
  if (p != c) {
    A[(size_t)p * n + q] = A[(size_t)c * n + q];
    I[(size_t)p * n + q] = I[(size_t)c * n + q];
  }
  A[(size_t)c * n + q] = a / d;
  I[(size_t)c * n + q] = b / d;
}

__global__ void k_gj_getcol(const double* A, int n, int c, double* f) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) f[r] = A[(size_t)r * n + c];
}

// every other row r: row_r -= f[r] * row_c (rows with f == 0 unchanged)
__global__ void k_gj_elim(double* A, double* I, int n, int c, const double* f) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (q >= n || r == c) return;
  const double fr = f[r];
  if (fr == 0.0) return;
  A[(size_t)r * n + q] -= fr * A[(size_t)c * n + q];
  I[(size_t)r * n + q] -= fr * I[(size_t)c * n + q];
}

// colsum[j] = sum_i C(i,j) (i ascending) and the Cmatrix::total order
__global__ void k_colsum(const double* C, int n, double* colsum) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += C[(size_t)i * n + j];
  colsum[j] = s;
}

// Sum of all elements of C, as the sum of its column sums (one block, fixed tree).
__global__ __launch_bounds__(256) void k_total(const double* colsum, int n, double* out) {
  __shared__ double s[256];
  double r = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) r += colsum[j];
  s[threadIdx.x] = r;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// sumWholeCharge: charge_map += tmp; tmp = 0
__global__ void k_sum_whole(double* cmap, double* tmp, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cmap[i] += tmp[i];
  tmp[i] = 0.0;
}

// Scalars of the redistribution in one block: Q (getTotalCharge), the sum
// over C(i,j) phi_j (by column sums) and the conductor potential V.
__global__ __launch_bounds__(256) void k_obj_scalars(const double* cmap, int S, int n, const double* colsum,
                                                     const double* phi, const int64_t* cmat_node, double total,
                                                     int fixed, double fixed_pot, double* scal, double* phis) {
  __shared__ double sq[256], sc[256];
  double q = 0.0, ct = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) {
    for (int s = 0; s < S; ++s) q += cmap[(size_t)s * n + j];
    const double pj = phi[cmat_node[j]];
    phis[j] = pj;
    ct += colsum[j] * pj;
  }
  sq[threadIdx.x] = q;
  sc[threadIdx.x] = ct;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sq[threadIdx.x] += sq[threadIdx.x + o];
      sc[threadIdx.x] += sc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double Q = sq[0], ctp = sc[0];
    scal[0] = fixed ? fixed_pot : (ctp + Q) / total;
    scal[1] = Q;
    scal[2] = ctp;
  }
}

// delta_rho_i = sum_j C(i,j) (V - phi_j), one wavefront per row; added to rho[0]
__global__ __launch_bounds__(256) void k_obj_drho(const double* C, int n, const double* phis, const double* scal,
                                                  const int64_t* cmat_node, double* rho0) {
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (w >= n) return;
  const double V = scal[0];
  const double* row = C + (size_t)w * n;
  double acc = 0.0;
  for (int j = lane; j < n; j += 64) acc += row[j] * (V - phis[j]);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if (lane == 0) rho0[cmat_node[w]] += acc;
}

}  // namespace

static int fdiv_floor(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

void objects_set(Ctx& c, int o, const uint8_t* mask, int ncmat, const int32_t* ijk, int fixed, double fixed_pot) {
  if (o < 0 || o >= FOPIS_MAX_OBJECTS) throw Error(FOPIS_E_ARG, "object index out of range");
  if (ncmat < 0) throw Error(FOPIS_E_ARG, "negative ncmat");
  ObjectDev& ob = c.obj[o];
  hipStream_t st = c.stream;
  const size_t nm = c.ncell_d;
  // merge this object's cells into the shared mask (value o+1)
  std::vector<uint8_t> hm(nm, 0);
  if (c.have_mask) FOPIS_HIP(hipMemcpy(hm.data(), c.mask.p, nm, hipMemcpyDeviceToHost));
  for (size_t q = 0; q < nm; ++q) {
    if (hm[q] == (uint8_t)(o + 1)) hm[q] = 0;
    if (mask[q]) {
      if (hm[q] != 0) throw Error(FOPIS_E_ARG, "objects overlap");
      hm[q] = (uint8_t)(o + 1);
    }
  }
  c.mask.alloc(nm);
  FOPIS_HIP(hipMemcpy(c.mask.p, hm.data(), nm, hipMemcpyHostToDevice));
  c.have_mask = true;
  // per tile: does a conductor cell lie in the tile or its 26 neighbours?
  // (a particle of the tile that ends in a cell of that box needs no mask read
  // when the answer is no)
  {
    const int m1 = c.nx + 1, m2 = c.ny + 1, m3 = c.nz + 1;  // mask extent; cell ci has mask index ci + 1
    std::vector<uint8_t> near((size_t)c.ntiles, 0);
    for (int mi = 0; mi < m1; ++mi)
      for (int mj = 0; mj < m2; ++mj)
        for (int mk = 0; mk < m3; ++mk) {
          if (!hm[((size_t)mi * m2 + mj) * m3 + mk]) continue;
          const int ci = mi - 1, cj = mj - 1, ck = mk - 1;
          // tiles whose neighbourhood box [8(t-1), 8(t+2)) holds this cell
          const int t0[3] = {fdiv_floor(ci, kTile), fdiv_floor(cj, kTile), fdiv_floor(ck, kTile)};
          for (int a0 = t0[0] - 1; a0 <= t0[0] + 1; ++a0)
            for (int a1 = t0[1] - 1; a1 <= t0[1] + 1; ++a1)
              for (int a2 = t0[2] - 1; a2 <= t0[2] + 1; ++a2)
                if (a0 >= 0 && a1 >= 0 && a2 >= 0 && a0 < c.nt[0] && a1 < c.nt[1] && a2 < c.nt[2])
                  near[((size_t)a0 * c.nt[1] + a1) * c.nt[2] + a2] = 1;
        }
    c.tile_near.alloc(std::max(c.ntiles, 1));
    FOPIS_HIP(hipMemcpy(c.tile_near.p, near.data(), near.size(), hipMemcpyHostToDevice));
  }
  ob.ncmat = ncmat;
  ob.fixed = fixed != 0;
  ob.fixed_potential = fixed_pot;
  ob.defined = ncmat > 0;
  ob.cmat_ijk.assign(ijk, ijk + 3 * (size_t)ncmat);
  std::vector<int64_t> nodes(ncmat);
  for (int q = 0; q < ncmat; ++q) {
    const int i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
    if (i < 0 || j < 0 || k < 0 || i >= c.cx || j >= c.cy || k >= c.cz)
      throw Error(FOPIS_E_ARG, "cmat node outside the grid");
    nodes[q] = (int64_t)c.id3(i, j, k);
  }
  ob.node_cmat.alloc(c.nnode);
  k_fill_i32<<<div_up((int64_t)c.nnode, 256), 256, 0, st>>>(ob.node_cmat.p, (int64_t)c.nnode, -1);
  ob.cmat_node.alloc(std::max(ncmat, 1));
  if (ncmat > 0) {
    FOPIS_HIP(hipMemcpyAsync(ob.cmat_node.p, nodes.data(), ncmat * sizeof(int64_t), hipMemcpyHostToDevice, st));
    k_set_node_cmat<<<div_up(ncmat, 256), 256, 0, st>>>(ob.node_cmat.p, ob.cmat_node.p, ncmat);
  }
  ob.C.alloc(std::max<size_t>((size_t)ncmat * ncmat, 1));
  ob.colsum.alloc(std::max(ncmat, 1));
  ob.tmp_charge.alloc(std::max<size_t>((size_t)c.S * ncmat, 1));
  ob.charge_map.alloc(std::max<size_t>((size_t)c.S * ncmat, 1));
  ob.absorbed.alloc(2 * (size_t)c.S);  // [S] captured, then [S] captured with no entry face found
  ob.scal.alloc(4 + std::max(ncmat, 1));
  FOPIS_HIP(hipMemsetAsync(ob.C.p, 0, ob.C.n * sizeof(double), st));
  FOPIS_HIP(hipMemsetAsync(ob.tmp_charge.p, 0, ob.tmp_charge.n * sizeof(double), st));
  FOPIS_HIP(hipMemsetAsync(ob.charge_map.p, 0, ob.charge_map.n * sizeof(double), st));
  FOPIS_HIP(hipMemsetAsync(ob.absorbed.p, 0, 2 * c.S * sizeof(unsigned long long), st));
  FOPIS_HIP(hipMemsetAsync(ob.scal.p, 0, ob.scal.n * sizeof(double), st));
  ob.current.assign(c.S, 0.0);
  ob.host_current_add.assign(c.S, 0.0);
  ob.host_current_pending.assign(c.S, 0.0);
  ob.total_cmat = 0.0;
  ob.cells.assign(mask, mask + nm);
  ob.lit_key = -1;
  ob.n_lit = 0;
  c.nobj = std::max(c.nobj, o + 1);
  FOPIS_HIP(hipStreamSynchronize(st));
}

// The cmat nodes that emit when light travels along +axis (sign > 0) or
// -axis: Spacecraft::is{X,Y,Z}surfacePoint (Spacecraft.cpp:461-504). A node
// whose cell and the cell behind it along the axis both lie in the cell map
// emits if exactly the one facing the light is inside the object (surface
// minus for sign > 0, plus otherwise); elsewhere it emits if the four nodes
// of the face square it spans are all cmat nodes. The list is made once per
// (axis, sign) on the host and kept on the device.
int objects_lit_nodes(Ctx& c, int o, int axis, int sign, const int32_t** dev_ijk) {
  ObjectDev& ob = c.obj[o];
  const int key = 2 * axis + (sign > 0 ? 0 : 1);
  if (ob.lit_key != key) {
    const int n1 = c.nx + 1, n2 = c.ny + 1, n3 = c.nz + 1;  // cell map extent
    auto in_map = [&](int i, int j, int k) { return i >= 0 && j >= 0 && k >= 0 && i < n1 && j < n2 && k < n3; };
    auto inside = [&](int i, int j, int k) { return ob.cells[((size_t)i * n2 + j) * n3 + k] != 0; };
    std::vector<int64_t> nodes(ob.ncmat);
    for (int q = 0; q < ob.ncmat; ++q)
      nodes[q] = (int64_t)c.id3(ob.cmat_ijk[3 * q], ob.cmat_ijk[3 * q + 1], ob.cmat_ijk[3 * q + 2]);
    std::vector<int64_t> sorted = nodes;
    std::sort(sorted.begin(), sorted.end());
    auto is_cmat = [&](int i, int j, int k) {
      if (i < 0 || j < 0 || k < 0 || i >= c.cx || j >= c.cy || k >= c.cz) return false;
      return std::binary_search(sorted.begin(), sorted.end(), (int64_t)c.id3(i, j, k));
    };
    const int t1 = axis == 0 ? 1 : 0, t2 = axis == 2 ? 1 : 2;  // the face's two tangential axes
    std::vector<int32_t> lit;
    for (int q = 0; q < ob.ncmat; ++q) {
      int p[3] = {ob.cmat_ijk[3 * q], ob.cmat_ijk[3 * q + 1], ob.cmat_ijk[3 * q + 2]};
      int b[3] = {p[0], p[1], p[2]};
      b[axis] -= 1;
      bool on;
      if (in_map(p[0], p[1], p[2]) && in_map(b[0], b[1], b[2])) {
        const bool here = inside(p[0], p[1], p[2]), behind = inside(b[0], b[1], b[2]);
        on = sign > 0 ? (here && !behind) : (!here && behind);
      } else {
        int u[3] = {p[0], p[1], p[2]}, w[3] = {p[0], p[1], p[2]}, uw[3] = {p[0], p[1], p[2]};
        u[t1] += 1;
        w[t2] += 1;
        uw[t1] += 1;
        uw[t2] += 1;
        on = is_cmat(p[0], p[1], p[2]) && is_cmat(u[0], u[1], u[2]) && is_cmat(w[0], w[1], w[2]) &&
             is_cmat(uw[0], uw[1], uw[2]);
      }
      if (on) lit.insert(lit.end(), p, p + 3);
    }
    ob.n_lit = (int)(lit.size() / 3);
    ob.lit_ijk.alloc(std::max<size_t>(lit.size(), 3));
    if (!lit.empty())
      FOPIS_HIP(hipMemcpyAsync(ob.lit_ijk.p, lit.data(), lit.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                               c.stream));
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    ob.lit_key = key;
  }
  *dev_ijk = ob.lit_ijk.p;
  return ob.n_lit;
}

static void finish_cmatrix(Ctx& c, ObjectDev& ob) {
  const int n = ob.ncmat;
  k_colsum<<<div_up(n, 256), 256, 0, c.stream>>>(ob.C.p, n, ob.colsum.p);
  k_total<<<1, 256, 0, c.stream>>>(ob.colsum.p, n, ob.scal.p + 3);
  FOPIS_HIP(hipMemcpyAsync(&ob.total_cmat, ob.scal.p + 3, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

// ---- the unit-charge responses B from the exact discrete Green's function
//
// With Dirichlet faces on every axis, the 7-point operator A (6 phi minus the
// six neighbours, dx = 1) on the m0 x m1 x m2 interior nodes is diagonalised
// by the sine transform: A = S diag(lambda) S with S = S0 (x) S1 (x) S2,
// S_d[k][u] = sqrt(2/(m_d+1)) sin(pi k u/(m_d+1)) (k, u = 1..m_d, orthogonal
// and symmetric) and lambda(k) = sum_d 4 sin^2(pi k_d / (2(m_d+1))). The
// response at node i to a unit charge at node j (RootGridObject.cpp:102-125
// solves A phi = rho/eps0 with rho = 1 at j) is therefore
//   B(j,i) = (1/eps0) sum_k prod_d S_d[k_d][u_d(i)] S_d[k_d][u_d(j)] / lambda(k),
// the limit the reference's unit solves approach. The surface nodes lie in a
// box of w_d nodes per axis, and the sum separates over the axes:
//   T[k_b][a_c,b_c][k_a] = sum_{k_c} S_c[k_c][a_c] S_c[k_c][b_c] / lambda   (GEMM, batched over k_b)
//   U[a_b,b_b][a_c,b_c][k_a] = sum_{k_b} S_b[k_b][a_b] S_b[k_b][b_b] T       (one GEMM)
//   B(j,i) = (1/eps0) sum_{k_a} S_a[k_a][a_a(i)] S_a[k_a][a_a(j)] U          (one wavefront per pair)
// with axis c the one of smallest box width and a the one of largest. FP64
// throughout; the GEMMs are plain LDS-tiled FMA kernels.

// C[b] (M x N) = A (M x K) x B[b] (K x N), row-major, A shared by the batch.
constexpr int kGT = 64, kGK = 16;
__global__ __launch_bounds__(256) void k_dgemm(const double* __restrict__ A, const double* __restrict__ Bm,
                                               double* __restrict__ Cm, int M, int N, int K, int64_t sB,
                                               int64_t sC) {
  __shared__ double As[kGK][kGT + 1];
  __shared__ double Bs[kGK][kGT + 1];
  const double* Bb = Bm + (int64_t)blockIdx.z * sB;
  double* Cb = Cm + (int64_t)blockIdx.z * sC;
  const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;  // 4x4 outputs: rows tr+16u, cols tc+16v
  double acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += kGK) {
    for (int q = threadIdx.x; q < kGK * kGT; q += 256) {
      const int kk = q / kGT, mm = q % kGT;
      // A tile transposed into As[k][m] (reads along k are strided: A is small)
      As[kk][mm] = (m0 + mm < M && k0 + kk < K) ? A[(int64_t)(m0 + mm) * K + k0 + kk] : 0.0;
      Bs[kk][mm] = (k0 + kk < K && n0 + mm < N) ? Bb[(int64_t)(k0 + kk) * N + n0 + mm] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGK; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = As[kk][tr + 16 * u];
#pragma unroll
      for (int v = 0; v < 4; ++v) b[v] = Bs[kk][tc + 16 * v];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = __builtin_fma(a[u], b[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = m0 + tr + 16 * u, col = n0 + tc + 16 * v;
      if (r < M && col < N) Cb[(int64_t)r * N + col] = acc[u][v];
    }
}

// invl[kb][kc][ka] = 1 / (la[ka] + lb[kb] + lc[kc])
__global__ void k_inv_lambda(const double* la, const double* lb, const double* lc, int ma, int mb, int mc,
                             double* invl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)ma * mb * mc) return;
  const int ka = (int)(i % ma), kc = (int)((i / ma) % mc), kb = (int)(i / ((int64_t)ma * mc));
  invl[i] = 1.0 / (la[ka] + lb[kb] + lc[kc]);
}

// P[(a,b)][k] = S[k][a] S[k][b] for a, b < w
__global__ void k_pair_products(const double* S, int m, int w, double* P) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)w * w * m) return;
  const int k = (int)(i % m), ab = (int)(i / m), a = ab / w, b = ab % w;
  P[i] = S[(int64_t)k * w + a] * S[(int64_t)k * w + b];
}

// B(col,row): one wavefront per (col, row) pair, lanes over k_a; a bounded
// grid strides over the pairs (n^2 wavefronts of 64 lanes would pass the
// 2^32 work-items a launch can hold once n exceeds 8192)
__global__ __launch_bounds__(256) void k_green_pairs(const double* __restrict__ U, const double* __restrict__ Sa,
                                                     const int32_t* __restrict__ box, int n, int ma, int wa,
                                                     int wb, int wc, double scale, double* __restrict__ B) {
  const int lane = threadIdx.x & 63;
  const int64_t npair = (int64_t)n * n;
  for (int64_t pair = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); pair < npair; pair += (int64_t)gridDim.x * 4) {
    const int col = (int)(pair / n), row = (int)(pair % n);
    const int aa = box[3 * col], ab = box[3 * col + 1], ac = box[3 * col + 2];
    const int ba = box[3 * row], bb = box[3 * row + 1], bc = box[3 * row + 2];
    const double* u = U + (((int64_t)ab * wb + bb) * (wc * wc) + (int64_t)ac * wc + bc) * ma;
    double s = 0.0;
    for (int k = lane; k < ma; k += 64) s += Sa[(int64_t)k * wa + aa] * Sa[(int64_t)k * wa + ba] * u[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) B[(size_t)col * n + row] = s * scale;
  }
}

// Fills B (n x n, (col,row) at col*n+row) from the sine-transform Green's
// function; false when a face is not Dirichlet or a node is not an interior
// unknown (the caller then solves for the columns).
static bool spectral_unit_responses(Ctx& c, ObjectDev& ob, double* B) {
  for (int q = 0; q < 6; ++q)
    if (c.d.boundary[q] != 'D') return false;
  const int n = ob.ncmat;
  const int m[3] = {c.nx - 2, c.ny - 2, c.nz - 2};
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
  for (int q = 0; q < n; ++q)
    for (int d = 0; d < 3; ++d) {
      const int u = ob.cmat_ijk[3 * q + d] - 1;  // interior unknown index along d (1..m)
      if (u < 1 || u > m[d]) return false;
      lo[d] = std::min(lo[d], u);
      hi[d] = std::max(hi[d], u);
    }
  int w[3], ord[3] = {0, 1, 2};
  for (int d = 0; d < 3; ++d) w[d] = hi[d] - lo[d] + 1;
  // ord = {a, b, c}: a largest width, c smallest
  std::sort(ord, ord + 3, [&](int x, int y) { return w[x] > w[y] || (w[x] == w[y] && x < y); });
  const int A_ = ord[0], B_ = ord[1], C_ = ord[2];
  {
    // beyond 16 GB of intermediate responses (surface nodes spread over a
    // wide box) the caller's per-node solves are the cheaper route
    const double nU = (double)w[C_] * w[C_] * m[A_];
    if (nU > INT32_MAX || (double)w[B_] * w[B_] * nU * sizeof(double) > 16.0 * (1 << 30)) return false;
  }
  hipStream_t st = c.stream;
  // sine tables S_d[k][a] (node lo_d + a) and eigenvalues, in FP64 on the host
  DBuf<double> S[3], L[3];
  for (int d = 0; d < 3; ++d) {
    const double md = (double)(m[d] + 1);
    std::vector<double> hs((size_t)m[d] * w[d]), hl(m[d]);
    const double nrm = std::sqrt(2.0 / md);
    for (int k = 1; k <= m[d]; ++k) {
      const double h = std::sin(M_PI * k / (2.0 * md));
      hl[k - 1] = 4.0 * h * h;
      for (int a = 0; a < w[d]; ++a) hs[(size_t)(k - 1) * w[d] + a] = nrm * std::sin(M_PI * k * (double)(lo[d] + a) / md);
    }
    S[d].alloc(hs.size());
    L[d].alloc(hl.size());
    // synchronous: the host tables go out of scope with this iteration
    FOPIS_HIP(hipStreamSynchronize(st));
    FOPIS_HIP(hipMemcpy(S[d].p, hs.data(), hs.size() * sizeof(double), hipMemcpyHostToDevice));
    FOPIS_HIP(hipMemcpy(L[d].p, hl.data(), hl.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  const int ma = m[A_], mb = m[B_], mc = m[C_], wa = w[A_], wb = w[B_], wc = w[C_];
  const int64_t nU = (int64_t)wc * wc * ma;
  DBuf<double> invl, Pc, Pb, T, U;
  invl.alloc((size_t)mb * mc * ma);
  k_inv_lambda<<<div_up((int64_t)ma * mb * mc, 256), 256, 0, st>>>(L[A_].p, L[B_].p, L[C_].p, ma, mb, mc, invl.p);
  Pc.alloc((size_t)wc * wc * mc);
  k_pair_products<<<div_up((int64_t)wc * wc * mc, 256), 256, 0, st>>>(S[C_].p, mc, wc, Pc.p);
  Pb.alloc((size_t)wb * wb * mb);
  k_pair_products<<<div_up((int64_t)wb * wb * mb, 256), 256, 0, st>>>(S[B_].p, mb, wb, Pb.p);
  // T[kb] (wc^2 x ma) = Pc (wc^2 x mc) x invl[kb] (mc x ma)
  T.alloc((size_t)mb * wc * wc * ma);
  k_dgemm<<<dim3(div_up(ma, kGT), div_up(wc * wc, kGT), mb), 256, 0, st>>>(
      Pc.p, invl.p, T.p, wc * wc, ma, mc, (int64_t)mc * ma, (int64_t)wc * wc * ma);
  invl.release();
  // U (wb^2 x wc^2 ma) = Pb (wb^2 x mb) x T (mb x wc^2 ma)
  U.alloc((size_t)wb * wb * nU);
  k_dgemm<<<dim3(div_up(nU, kGT), div_up(wb * wb, kGT), 1), 256, 0, st>>>(Pb.p, T.p, U.p, wb * wb, (int)nU, mb, 0,
                                                                          0);
  T.release();
  // each node's box coordinates in (a, b, c) order
  std::vector<int32_t> hbox(3 * (size_t)n);
  for (int q = 0; q < n; ++q) {
    hbox[3 * q] = ob.cmat_ijk[3 * q + A_] - 1 - lo[A_];
    hbox[3 * q + 1] = ob.cmat_ijk[3 * q + B_] - 1 - lo[B_];
    hbox[3 * q + 2] = ob.cmat_ijk[3 * q + C_] - 1 - lo[C_];
  }
  DBuf<int32_t> box;
  box.alloc(hbox.size());
  FOPIS_HIP(hipMemcpyAsync(box.p, hbox.data(), hbox.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  const int64_t pairs = (int64_t)n * n;
  const int64_t gblocks = std::min<int64_t>(div_up(pairs, 4), (int64_t)1 << 20);
  k_green_pairs<<<(unsigned)gblocks, 256, 0, st>>>(U.p, S[A_].p, box.p, n, ma, wa, wb, wc, 1.0 / c.d.eps0, B);
  FOPIS_HIP(hipGetLastError());
  FOPIS_HIP(hipStreamSynchronize(st));
  return true;
}

void objects_build_cmatrix(Ctx& c, int o) {
  ObjectDev& ob = c.obj[o];
  const int n = ob.ncmat;
  if (n <= 0) return;
  hipStream_t st = c.stream;
  const int64_t nn = (int64_t)c.nnode;
  DBuf<double> B;
  B.alloc((size_t)n * n);
  const char* sp = std::getenv("FOPIS_CMAT_SPECTRAL");
  const bool spectral = !(sp && sp[0] == '0');  // FOPIS_CMAT_SPECTRAL=0: one solve per node
  if (!spectral || !spectral_unit_responses(c, ob, B.p)) {
    // one unit-charge multigrid solve per surface node (RootGridObject.cpp:102-125)
    std::vector<int64_t> nodes(n);
    FOPIS_HIP(hipMemcpy(nodes.data(), ob.cmat_node.p, n * sizeof(int64_t), hipMemcpyDeviceToHost));
    k_zero_d<<<div_up(nn, 256), 256, 0, st>>>(c.phi.p, nn);
    for (int col = 0; col < n; ++col) {
      k_unit_rhs<<<div_up(nn, 256), 256, 0, st>>>(c.rho.p, nn, nodes[col]);
      double rr = 0.0;
      poisson_solve(c, 1e-11, 400, &rr, c.rho.p);
      k_gather_col<<<div_up(n, 256), 256, 0, st>>>(B.p, n, col, c.phi.p, ob.cmat_node.p);
    }
    k_zero_d<<<div_up(nn, 256), 256, 0, st>>>(c.rho.p, nn);
  } else {
    // the reference leaves phi and rho cleared after the unit solves (RootGridObject.cpp:83-85)
    k_zero_d<<<div_up(nn, 256), 256, 0, st>>>(c.phi.p, nn);
  }
  // C = B^-1 by Gauss-Jordan with partial pivoting
  k_identity<<<div_up((int64_t)n * n, 256), 256, 0, st>>>(ob.C.p, n);
  DBuf<double> f;
  f.alloc(n + 2);
  DBuf<int> piv;
  piv.alloc(1);
  for (int col = 0; col < n; ++col) {
    k_gj_pivot<<<1, 1024, 0, st>>>(B.p, n, col, piv.p, f.p + n);
    k_gj_swap_scale<<<div_up(n, 256), 256, 0, st>>>(B.p, ob.C.p, n, col, piv.p, f.p + n);
    k_gj_getcol<<<div_up(n, 256), 256, 0, st>>>(B.p, n, col, f.p);
    k_gj_elim<<<dim3(div_up(n, 256), n), 256, 0, st>>>(B.p, ob.C.p, n, col, f.p);
  }
  FOPIS_HIP(hipGetLastError());
  finish_cmatrix(c, ob);
}

void objects_set_cmatrix(Ctx& c, int o, const double* C) {
  ObjectDev& ob = c.obj[o];
  const size_t n2 = (size_t)ob.ncmat * ob.ncmat;
  if (n2 == 0) return;
  FOPIS_HIP(hipMemcpy(ob.C.p, C, n2 * sizeof(double), hipMemcpyHostToDevice));
  finish_cmatrix(c, ob);
}

void objects_get_cmatrix(Ctx& c, int o, double* C) {
  ObjectDev& ob = c.obj[o];
  const size_t n2 = (size_t)ob.ncmat * ob.ncmat;
  if (n2 == 0) return;
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  FOPIS_HIP(hipMemcpy(C, ob.C.p, n2 * sizeof(double), hipMemcpyDeviceToHost));
}

// RootGrid::resetObjects -> Spacecraft::resetCurrent at the top of mainLoopES
// (RootGrid.cpp:82-86, RootGridObject.cpp:162-170, Spacecraft.cpp:303-307):
// the current counts one step's captured and emitted superparticles. Emission
// staged before the step belongs to it.
void objects_reset_step(Ctx& c) {
  for (int o = 0; o < c.nobj; ++o) {
    ObjectDev& ob = c.obj[o];
    if (!ob.defined) continue;
    FOPIS_HIP(hipMemsetAsync(ob.absorbed.p, 0, 2 * c.S * sizeof(unsigned long long), c.stream));
    ob.host_current_add = ob.host_current_pending;
    ob.host_current_pending.assign(c.S, 0.0);
  }
}

void objects_sum_whole_charge(Ctx& c) {
  for (int o = 0; o < c.nobj; ++o) {
    ObjectDev& ob = c.obj[o];
    if (!ob.defined) continue;
    const int64_t n = (int64_t)c.S * ob.ncmat;
    k_sum_whole<<<div_up(n, 256), 256, 0, c.stream>>>(ob.charge_map.p, ob.tmp_charge.p, n);
  }
  FOPIS_HIP(hipGetLastError());
}

void objects_redistribute(Ctx& c) {
  for (int o = 0; o < c.nobj; ++o) {
    ObjectDev& ob = c.obj[o];
    if (!ob.defined) continue;
    const int n = ob.ncmat;
    if (ob.total_cmat == 0.0 && !ob.fixed) throw Error(FOPIS_E_STATE, "capacity matrix not set");
    k_obj_scalars<<<1, 256, 0, c.stream>>>(ob.charge_map.p, c.S, n, ob.colsum.p, c.phi.p, ob.cmat_node.p,
                                           ob.total_cmat, ob.fixed ? 1 : 0, ob.fixed_potential, ob.scal.p,
                                           ob.scal.p + 4);
    k_obj_drho<<<div_up((int64_t)n * 64, 256), 256, 0, c.stream>>>(ob.C.p, n, ob.scal.p + 4, ob.scal.p,
                                                                   ob.cmat_node.p, c.rho.p);
  }
  FOPIS_HIP(hipGetLastError());
}

void objects_add_charge(Ctx& c, int o, const double* dq, const double* dcur) {
  ObjectDev& ob = c.obj[o];
  const size_t n = (size_t)c.S * ob.ncmat;
  if (dq && n > 0) {
    std::vector<double> cur(n);
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    FOPIS_HIP(hipMemcpy(cur.data(), ob.tmp_charge.p, n * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < n; ++q) cur[q] += dq[q];
    FOPIS_HIP(hipMemcpy(ob.tmp_charge.p, cur.data(), n * sizeof(double), hipMemcpyHostToDevice));
  }
  if (dcur)
    for (int s = 0; s < c.S; ++s) ob.host_current_pending[s] += dcur[s];
}

void objects_state(Ctx& c, int o, double* pot, double* q, double* cur, double* cmap) {
  ObjectDev& ob = c.obj[o];
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  double sc[4] = {0, 0, 0, 0};
  FOPIS_HIP(hipMemcpy(sc, ob.scal.p, 4 * sizeof(double), hipMemcpyDeviceToHost));
  if (pot) *pot = sc[0];
  if (q) *q = sc[1];
  if (cur) {
    std::vector<unsigned long long> ab(c.S);
    FOPIS_HIP(hipMemcpy(ab.data(), ob.absorbed.p, c.S * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int s = 0; s < c.S; ++s)
      cur[s] = (double)ab[s] * (c.d.species[s].charge * c.d.species[s].size) + ob.host_current_add[s] +
               ob.host_current_pending[s];
  }
  if (cmap && ob.ncmat > 0)
    FOPIS_HIP(hipMemcpy(cmap, ob.charge_map.p, (size_t)c.S * ob.ncmat * sizeof(double), hipMemcpyDeviceToHost));
}

void objects_unattributed(Ctx& c, int o, int64_t* out) {
  ObjectDev& ob = c.obj[o];
  std::vector<unsigned long long> ab(2 * (size_t)c.S, 0ull);
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  if (ob.absorbed.p)
    FOPIS_HIP(hipMemcpy(ab.data(), ob.absorbed.p, 2 * c.S * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int s = 0; s < c.S; ++s) out[s] = (int64_t)ab[c.S + s];
}

void objects_set_charge_map(Ctx& c, int o, const double* cmap) {
  ObjectDev& ob = c.obj[o];
  if (ob.ncmat <= 0) return;
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  FOPIS_HIP(hipMemcpy(ob.charge_map.p, cmap, (size_t)c.S * ob.ncmat * sizeof(double), hipMemcpyHostToDevice));
}

// ---- warm starts carried over between steps
//
// The first solve of a step has the deposited charge alone on its right-hand
// side, the second adds the conductor's redistributed charge; the previous
// step's final potential carries that step's conductor response psi, which
// the first solve must remove and the second must add back. Starting the
// first solve from phi - psi and the second from (its first potential) + psi
// leaves each with the change of one step to correct. Only the starting
// points change: every solve still iterates to the same residual.
__global__ void k_carry_first(double* phi, const double* psi, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) phi[i] -= psi[i];
}

// psi <- phi (the first solve's potential); phi += old psi when `add`
__global__ void k_carry_second(double* phi, double* psi, int64_t n, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = phi[i];
  if (add) phi[i] = f + psi[i];
  psi[i] = f;
}

// psi <- phi - psi (the final potential minus the first solve's)
__global__ void k_carry_end(const double* phi, double* psi, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) psi[i] = phi[i] - psi[i];
}

void objects_carry_first(Ctx& c) {
  if (!c.carry || c.psi_step != c.timestep - 1 || !c.psi.p) return;
  const int64_t n = (int64_t)c.nnode;
  k_carry_first<<<div_up(n, 256), 256, 0, c.stream>>>(c.phi.p, c.psi.p, n);
  FOPIS_HIP(hipGetLastError());
}

void objects_carry_second(Ctx& c) {
  if (!c.carry) return;
  const int64_t n = (int64_t)c.nnode;
  const bool add = c.psi_step == c.timestep - 1 && c.psi.p;
  c.psi.alloc(c.nnode);
  k_carry_second<<<div_up(n, 256), 256, 0, c.stream>>>(c.phi.p, c.psi.p, n, add ? 1 : 0);
  FOPIS_HIP(hipGetLastError());
}

void objects_carry_end(Ctx& c) {
  if (!c.carry || !c.psi.p) return;
  const int64_t n = (int64_t)c.nnode;
  k_carry_end<<<div_up(n, 256), 256, 0, c.stream>>>(c.phi.p, c.psi.p, n);
  FOPIS_HIP(hipGetLastError());
  c.psi_step = c.timestep;
}

}  // namespace fopis
