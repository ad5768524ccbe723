// Grid-side kernels: reduction of the per-tile charge/count windows into
// rho and density, the E field, field access and field energy.
//
// Reference behaviour followed (hsimyu/fopis):
//   rho[0] = sum of species   src/Grid/RootGrid/RootGridField.cpp:1205-1214
//   density                   src/Grid/RootGrid/RootGridField.cpp:1129-1150
//   edge E and nodal E_ref    src/Grid/RootGrid/RootGridField.cpp:456-582
//   field energy              src/Grid/Field.cpp:194-213
#include "fopis_internal.hpp"

#pragma clang fp contract(off)

namespace fopis {

namespace {

__device__ __forceinline__ int fdivd(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
__device__ __forceinline__ int cdivd(int a, int b) { return -fdivd(-a, b); }

// One thread per glue node: rho[s+1] = the far deposits (only for species in
// far_mask; read and cleared) plus the windows of every tile covering the
// node, in tile order (deterministic); rho[0] = sum over s.
template <class R>
__global__ void k_reduce_rho(const R* __restrict__ part, double* __restrict__ rho, double* __restrict__ far,
                             uint32_t far_mask, int S, int ntiles, int cx, int cy, int cz, int T0, int T1, int T2) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)cx * cy * cz;
  if (g >= nn) return;
  const int k = (int)(g % cz), j = (int)((g / cz) % cy), i = (int)(g / ((int64_t)cy * cz));
  const int p0 = i - 1, p1 = j - 1, p2 = k - 1;  // particle node index
  const int a0 = max(0, cdivd(p0 - (kTile + 1), kTile)), b0 = min(T0 - 1, fdivd(p0 + 1, kTile));
  const int a1 = max(0, cdivd(p1 - (kTile + 1), kTile)), b1 = min(T1 - 1, fdivd(p1 + 1, kTile));
  const int a2 = max(0, cdivd(p2 - (kTile + 1), kTile)), b2 = min(T2 - 1, fdivd(p2 + 1, kTile));
  double tot = 0.0;
  for (int s = 0; s < S; ++s) {
    const R* P = part + (size_t)s * ntiles * (kRW * kRW * kRW);
    double acc = 0.0;
    if ((far_mask >> s) & 1u) {
      acc = far[(size_t)s * nn + g];
      far[(size_t)s * nn + g] = 0.0;
    }
    for (int tx = a0; tx <= b0; ++tx)
      for (int ty = a1; ty <= b1; ++ty)
        for (int tz = a2; tz <= b2; ++tz) {
          const int t = (tx * T1 + ty) * T2 + tz;
          const int w0 = p0 - tx * kTile + 1, w1 = p1 - ty * kTile + 1, w2 = p2 - tz * kTile + 1;
          acc += (double)P[(size_t)t * (kRW * kRW * kRW) + (w0 * kRW + w1) * kRW + w2];
        }
    rho[(size_t)(s + 1) * nn + g] = acc;
    tot += acc;
  }
  rho[g] = tot;
}

// One thread per density cell ((cx-1)^3, glue-relative): density[s] = the far
// counts (species in far_mask; read and cleared) plus the windows' counts.
__global__ void k_reduce_dens(const int32_t* __restrict__ part, double* __restrict__ dens, double* __restrict__ far,
                              uint32_t far_mask, const double* size, int S, int ntiles, int cx, int cy, int cz, int T0,
                              int T1, int T2) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int dx_ = cx - 1, dy_ = cy - 1, dz_ = cz - 1;
  const int64_t nn = (int64_t)dx_ * dy_ * dz_;
  if (g >= nn) return;
  const int k = (int)(g % dz_), j = (int)((g / dz_) % dy_), i = (int)(g / ((int64_t)dy_ * dz_));
  const int c0 = i - 1, c1 = j - 1, c2 = k - 1;
  const int a0 = max(0, cdivd(c0 - kTile, kTile)), b0 = min(T0 - 1, fdivd(c0 + 1, kTile));
  const int a1 = max(0, cdivd(c1 - kTile, kTile)), b1 = min(T1 - 1, fdivd(c1 + 1, kTile));
  const int a2 = max(0, cdivd(c2 - kTile, kTile)), b2 = min(T2 - 1, fdivd(c2 + 1, kTile));
  for (int s = 0; s < S; ++s) {
    const int32_t* P = part + (size_t)s * ntiles * (kDW * kDW * kDW);
    int64_t n = 0;
    for (int tx = a0; tx <= b0; ++tx)
      for (int ty = a1; ty <= b1; ++ty)
        for (int tz = a2; tz <= b2; ++tz) {
          const int t = (tx * T1 + ty) * T2 + tz;
          const int w0 = c0 - tx * kTile + 1, w1 = c1 - ty * kTile + 1, w2 = c2 - tz * kTile + 1;
          n += P[(size_t)t * (kDW * kDW * kDW) + (w0 * kDW + w1) * kDW + w2];
        }
    // density += size once per particle (RootGridField.cpp:1142): n additions
    // of an integer-valued size are exact, so n * size equals them.
    double base = 0.0;
    if ((far_mask >> s) & 1u) {
      base = far[(size_t)s * nn + g];
      far[(size_t)s * nn + g] = 0.0;
    }
    dens[(size_t)s * nn + g] = base + (double)n * size[s];
  }
}

struct Faces {
  int b[6];  // 1 if the face is physical (non-periodic)
};

// Nodal reference E (RootGridField.cpp:489-582) from phi, per glue node.
__global__ void k_eref(const double* __restrict__ phi, double* __restrict__ er, int cx, int cy, int cz, Faces f) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)cx * cy * cz;
  if (g >= nn) return;
  const int k = (int)(g % cz), j = (int)((g / cz) % cy), i = (int)(g / ((int64_t)cy * cz));
  const double pdx = 1.0 / 1.0;
  const int64_t sI = (int64_t)cy * cz, sJ = cz;
  double ex = 0.0, ey = 0.0, ez = 0.0;
  const bool in_i = i >= 1 && i <= cx - 2, in_j = j >= 1 && j <= cy - 2, in_k = k >= 1 && k <= cz - 2;
  if (in_i && in_j && in_k) {
    const double* P = phi + g;
    // x
    if (i == 1 && f.b[0])
      ex = 0.5 * (P[2 * sI] - 4.0 * P[sI] + 3.0 * P[0]);
    else if (i == cx - 2 && f.b[1])
      ex = 0.5 * (-P[-2 * sI] + 4.0 * P[-sI] - 3.0 * P[0]);
    else
      ex = 0.5 * ((P[-sI] - P[0]) * pdx + (P[0] - P[sI]) * pdx);
    // y
    if (j == 1 && f.b[2])
      ey = 0.5 * (P[2 * sJ] - 4.0 * P[sJ] + 3.0 * P[0]);
    else if (j == cy - 2 && f.b[3])
      ey = 0.5 * (-P[-2 * sJ] + 4.0 * P[-sJ] - 3.0 * P[0]);
    else
      ey = 0.5 * ((P[-sJ] - P[0]) * pdx + (P[0] - P[sJ]) * pdx);
    // z
    if (k == 1 && f.b[4])
      ez = 0.5 * (P[2] - 4.0 * P[1] + 3.0 * P[0]);
    else if (k == cz - 2 && f.b[5])
      ez = 0.5 * (-P[-2] + 4.0 * P[-1] - 3.0 * P[0]);
    else
      ez = 0.5 * ((P[-1] - P[0]) * pdx + (P[0] - P[1]) * pdx);
  }
  er[g] = ex;
  er[nn + g] = ey;
  er[2 * nn + g] = ez;
}

// Edge E (RootGridField.cpp:456-486) into an array of the given dims.
__global__ void k_edge(const double* __restrict__ phi, double* __restrict__ out, int axis, int cx, int cy, int cz) {
  const int ex = cx - (axis == 0), ey = cy - (axis == 1), ez = cz - (axis == 2);
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)ex * ey * ez;
  if (g >= nn) return;
  const int k = (int)(g % ez), j = (int)((g / ez) % ey), i = (int)(g / ((int64_t)ey * ez));
  const int64_t sI = (int64_t)cy * cz, sJ = cz;
  const int64_t p = i * sI + j * sJ + k;
  const int64_t d = axis == 0 ? sI : (axis == 1 ? sJ : 1);
  out[g] = (phi[p] - phi[p + d]) * (1.0 / 1.0);
}

// Sum of squared edge fields over Field::getEfieldEnergy's index ranges.
__global__ __launch_bounds__(256) void k_efield_energy(const double* __restrict__ phi, int cx, int cy, int cz,
                                                       double* part) {
  __shared__ double s[256];
  const int64_t nn = (int64_t)cx * cy * cz, sI = (int64_t)cy * cz, sJ = cz;
  double acc = 0.0;
  for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < nn; g += (int64_t)gridDim.x * 256) {
    const int k = (int)(g % cz), j = (int)((g / cz) % cy), i = (int)(g / sI);
    if (i < 1 || i > cx - 2 || j < 1 || j > cy - 2 || k < 1 || k > cz - 2) continue;
    const double p = phi[g];
    if (i < cx - 2) {
      const double e = (p - phi[g + sI]);
      acc += e * e;
    }
    if (j < cy - 2) {
      const double e = (p - phi[g + sJ]);
      acc += e * e;
    }
    if (k < cz - 2) {
      const double e = (p - phi[g + 1]);
      acc += e * e;
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_serial_sum(const double* part, int n, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *out = s;
  }
}

Faces faces(const Ctx& c) {
  Faces f;
  for (int q = 0; q < 6; ++q) f.b[q] = c.d.boundary[q] != 'P';
  return f;
}

DBuf<double>& sizes_buf(Ctx& c) { return c.species_sizes; }

}  // namespace

void fields_init(Ctx& c) {
  c.phi.alloc(c.nnode);
  c.rho.alloc((size_t)(c.S + 1) * c.nnode);
  c.density.alloc((size_t)c.S * c.ncell_d);
  c.eref.alloc(3 * c.nnode);
  c.red.alloc(4096);
  FOPIS_HIP(hipMemsetAsync(c.phi.p, 0, c.phi.n * sizeof(double), c.stream));
  FOPIS_HIP(hipMemsetAsync(c.rho.p, 0, c.rho.n * sizeof(double), c.stream));
  FOPIS_HIP(hipMemsetAsync(c.density.p, 0, c.density.n * sizeof(double), c.stream));
  FOPIS_HIP(hipMemsetAsync(c.eref.p, 0, c.eref.n * sizeof(double), c.stream));
  auto& sz = sizes_buf(c);
  sz.alloc(FOPIS_MAX_SPECIES);
  double h[FOPIS_MAX_SPECIES] = {0};
  for (int s = 0; s < c.S; ++s) h[s] = c.d.species[s].size;
  FOPIS_HIP(hipMemcpyAsync(sz.p, h, sizeof(h), hipMemcpyHostToDevice, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

void fields_reduce_windows(Ctx& c, bool do_rho, bool do_dens, uint32_t far_mask) {
  hipStream_t st = c.stream;
  if (do_rho) {
    const int64_t nn = (int64_t)c.nnode;
    // charge windows are FP64 for both particle precisions
    k_reduce_rho<double><<<div_up(nn, 256), 256, 0, st>>>(c.rho_part.p, c.rho.p, c.rho_far.p, far_mask, c.S,
                                                         c.ntiles, c.cx, c.cy, c.cz, c.nt[0], c.nt[1], c.nt[2]);
  }
  if (do_dens) {
    k_reduce_dens<<<div_up((int64_t)c.ncell_d, 256), 256, 0, st>>>(c.dens_part.p, c.density.p, c.dens_far.p,
                                                                   far_mask, sizes_buf(c).p, c.S, c.ntiles, c.cx,
                                                                   c.cy, c.cz, c.nt[0], c.nt[1], c.nt[2]);
  }
  FOPIS_HIP(hipGetLastError());
}

void fields_efield(Ctx& c) {
  Phase ph(c, "efield");
  k_eref<<<div_up((int64_t)c.nnode, 256), 256, 0, c.stream>>>(c.phi.p, c.eref.p, c.cx, c.cy, c.cz, faces(c));
  FOPIS_HIP(hipGetLastError());
}

static size_t field_len(const Ctx& c, int which) {
  switch (which) {
    case FOPIS_PHI:
    case FOPIS_RHO:
    case FOPIS_EXREF:
    case FOPIS_EYREF:
    case FOPIS_EZREF:
      return c.nnode;
    case FOPIS_DENSITY:
      return c.ncell_d;
    case FOPIS_EX:
      return (size_t)(c.cx - 1) * c.cy * c.cz;
    case FOPIS_EY:
      return (size_t)c.cx * (c.cy - 1) * c.cz;
    case FOPIS_EZ:
      return (size_t)c.cx * c.cy * (c.cz - 1);
  }
  throw Error(FOPIS_E_ARG, "unknown field selector");
}

static double* field_ptr(Ctx& c, int which, int s) {
  switch (which) {
    case FOPIS_PHI:
      return c.phi.p;
    case FOPIS_RHO:
      if (s < 0 || s > c.S) throw Error(FOPIS_E_ARG, "rho index out of range");
      return c.rho.p + (size_t)s * c.nnode;
    case FOPIS_DENSITY:
      if (s < 0 || s >= c.S) throw Error(FOPIS_E_ARG, "density species out of range");
      return c.density.p + (size_t)s * c.ncell_d;
    case FOPIS_EXREF:
    case FOPIS_EYREF:
    case FOPIS_EZREF:
      return c.eref.p + (size_t)(which - FOPIS_EXREF) * c.nnode;
  }
  return nullptr;
}

void fields_get(Ctx& c, int which, int s, double* out) {
  const size_t n = field_len(c, which);
  if (which >= FOPIS_EX && which <= FOPIS_EZ) {
    DBuf<double> tmp;
    tmp.alloc(n);
    k_edge<<<div_up((int64_t)n, 256), 256, 0, c.stream>>>(c.phi.p, tmp.p, which - FOPIS_EX, c.cx, c.cy, c.cz);
    FOPIS_HIP(hipMemcpyAsync(out, tmp.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FOPIS_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  FOPIS_HIP(hipMemcpyAsync(out, field_ptr(c, which, s), n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

void fields_set(Ctx& c, int which, int s, const double* in) {
  if (which >= FOPIS_EX && which <= FOPIS_EZ) throw Error(FOPIS_E_ARG, "edge fields are derived from phi");
  const size_t n = field_len(c, which);
  FOPIS_HIP(hipMemcpyAsync(field_ptr(c, which, s), in, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
}

void fields_energy_queue(Ctx& c, double* part, double* sum) {
  k_efield_energy<<<kEnergyBlocks, 256, 0, c.stream>>>(c.phi.p, c.cx, c.cy, c.cz, part);
  k_serial_sum<<<1, 1, 0, c.stream>>>(part, kEnergyBlocks, sum);
  FOPIS_HIP(hipGetLastError());
}

double fields_energy(Ctx& c) {
  c.red.alloc(std::max<size_t>(c.red.n, kEnergyBlocks + 1));
  fields_energy_queue(c, c.red.p, c.red.p + kEnergyBlocks);
  double e = 0.0;
  FOPIS_HIP(hipMemcpyAsync(&e, c.red.p + kEnergyBlocks, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FOPIS_HIP(hipStreamSynchronize(c.stream));
  return 0.5 * c.d.eps0 * e;
}

}  // namespace fopis
