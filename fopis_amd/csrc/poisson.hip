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
#include "fopis_internal.hpp"

#pragma clang fp contract(off)

namespace fopis {

namespace {

struct Lv {
  int n0, n1, n2;
  int64_t sx, sy;
};

// b = rc * rho at the level-0 interior; r = b - A phi written as P;
// per-block partial sums of r^2 and b^2 in FP64.
template <class P>
__global__ __launch_bounds__(256) void k_resid0(const double* __restrict__ phi, const double* __restrict__ rho,
                                                double rc, P* __restrict__ r, Lv L, double* part) {
  __shared__ double s1[256], s2[256];
  const int64_t nint = (int64_t)L.n0 * L.n1 * L.n2;
  double rr = 0.0, bb = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < nint; q += (int64_t)gridDim.x * 256) {
    const int c = (int)(q % L.n2) + 1, b = (int)((q / L.n2) % L.n1) + 1, a = (int)(q / ((int64_t)L.n1 * L.n2)) + 1;
    const int64_t g = a * L.sx + b * L.sy + c;
    const double bv = rc * rho[g];
    const double ap = 6.0 * phi[g] - (phi[g - L.sx] + phi[g + L.sx] + phi[g - L.sy] + phi[g + L.sy] + phi[g - 1] +
                                      phi[g + 1]);
    const double rv = bv - ap;
    r[g] = (P)rv;
    rr += rv * rv;
    bb += bv * bv;
  }
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

// Final sum of the block partials: one block, fixed tree (deterministic).
__global__ __launch_bounds__(256) void k_sum2(const double* part, int nb, double* out) {
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
    out[0] = s1[0];
    out[1] = s2[0];
  }
}

// Red-black SOR sweep over points with (a+b+c) % 2 == color.
template <class P>
__global__ void k_smooth(P* __restrict__ e, const P* __restrict__ r, Lv L, int color, P w) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y + 1, a = blockIdx.z + 1;
  const int c0 = ((a + b + 1) & 1) == color ? 1 : 2;
  const int c = c0 + 2 * m;
  if (c > L.n2) return;
  const int64_t g = a * L.sx + b * L.sy + c;
  const P nb = e[g - L.sx] + e[g + L.sx] + e[g - L.sy] + e[g + L.sy] + e[g - 1] + e[g + 1];
  e[g] = ((P)1 - w) * e[g] + w * (nb + r[g]) / (P)6;
}

template <class P>
__global__ void k_resid(P* __restrict__ t, const P* __restrict__ e, const P* __restrict__ r, Lv L) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int b = blockIdx.y + 1, a = blockIdx.z + 1;
  if (c > L.n2) return;
  const int64_t g = a * L.sx + b * L.sy + c;
  t[g] = r[g] - ((P)6 * e[g] - (e[g - L.sx] + e[g + L.sx] + e[g - L.sy] + e[g + L.sy] + e[g - 1] + e[g + 1]));
}

// Full weighting to the coarse grid, scaled by 4 for the unscaled 2h operator.
template <class P>
__global__ void k_restrict(P* __restrict__ rc, Lv C, const P* __restrict__ t, Lv F) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int b = blockIdx.y + 1, a = blockIdx.z + 1;
  if (c > C.n2) return;
  const int64_t f = (2 * a) * F.sx + (2 * b) * F.sy + 2 * c;
  P acc = 0;
#pragma unroll
  for (int da = -1; da <= 1; ++da)
#pragma unroll
    for (int db = -1; db <= 1; ++db)
#pragma unroll
      for (int dc = -1; dc <= 1; ++dc) {
        const int nz = (da != 0) + (db != 0) + (dc != 0);
        const P wgt = nz == 0 ? (P)1 : (nz == 1 ? (P)0.5 : (nz == 2 ? (P)0.25 : (P)0.125));
        acc += wgt * t[f + da * F.sx + db * F.sy + dc];
      }
  rc[a * C.sx + b * C.sy + c] = acc * (P)0.5;  // (4/8) * sum
}

// e_fine += trilinear interpolation of e_coarse (coarse node A sits at fine 2A).
template <class P>
__global__ void k_prolong(P* __restrict__ e, Lv F, const P* __restrict__ ec, Lv C) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int b = blockIdx.y + 1, a = blockIdx.z + 1;
  if (c > F.n2) return;
  const int a0 = a >> 1, a1 = (a + 1) >> 1, b0 = b >> 1, b1 = (b + 1) >> 1, c0 = c >> 1, c1 = (c + 1) >> 1;
  auto E = [&](int x, int y, int z) { return ec[x * C.sx + y * C.sy + z]; };
  const P v = (E(a0, b0, c0) + E(a1, b0, c0) + E(a0, b1, c0) + E(a1, b1, c0) + E(a0, b0, c1) + E(a1, b0, c1) +
               E(a0, b1, c1) + E(a1, b1, c1)) *
              (P)0.125;
  e[a * F.sx + b * F.sy + c] += v;
}

template <class P>
__global__ void k_zero(P* __restrict__ e, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) e[i] = (P)0;
}

template <class P>
__global__ void k_add(double* __restrict__ phi, const P* __restrict__ e, Lv L) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int b = blockIdx.y + 1, a = blockIdx.z + 1;
  if (c > L.n2) return;
  const int64_t g = a * L.sx + b * L.sy + c;
  phi[g] += (double)e[g];
}

template <class P>
__global__ void k_coarse_single(P* e, const P* r, Lv L) {
  // one unknown: A = 6
  const int64_t g = L.sx + L.sy + 1;
  e[g] = r[g] / (P)6;
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
dim3 grid_half(const MgLevel& m) { return dim3(div_up((m.n[2] + 1) / 2, 64), m.n[1], m.n[0]); }

template <class P>
void smooth(Ctx& c, MgLevel& m, int sweeps, double w) {
  for (int s = 0; s < sweeps; ++s)
    for (int col = 0; col < 2; ++col)
      k_smooth<P><<<grid_half(m), 64, 0, c.stream>>>(E_<P>(m), R_<P>(m), lv_of(m), col, (P)w);
}

template <class P>
void vcycle(Ctx& c, int l) {
  MgLevel& m = c.mg[l];
  const double w = c.d.mg_omega;
  if (l == (int)c.mg.size() - 1) {
    if (m.n[0] == 1 && m.n[1] == 1 && m.n[2] == 1) {
      k_coarse_single<P><<<1, 1, 0, c.stream>>>(E_<P>(m), R_<P>(m), lv_of(m));
    } else {
      const int sweeps = 4 * std::max({m.n[0], m.n[1], m.n[2]}) + 8;
      smooth<P>(c, m, sweeps, 1.0);
    }
    return;
  }
  MgLevel& mc = c.mg[l + 1];
  smooth<P>(c, m, c.d.mg_pre, w);
  k_resid<P><<<grid_full(m), 64, 0, c.stream>>>(T_<P>(m), E_<P>(m), R_<P>(m), lv_of(m));
  k_restrict<P><<<grid_full(mc), 64, 0, c.stream>>>(R_<P>(mc), lv_of(mc), T_<P>(m), lv_of(m));
  k_zero<P><<<div_up((int64_t)mc.size, 256), 256, 0, c.stream>>>(E_<P>(mc) - mc.off, (int64_t)mc.size);
  vcycle<P>(c, l + 1);
  k_prolong<P><<<grid_full(m), 64, 0, c.stream>>>(E_<P>(m), lv_of(m), E_<P>(mc), lv_of(mc));
  smooth<P>(c, m, c.d.mg_post, w);
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
  const int nb = 512;
  c.red.alloc(std::max<size_t>(c.red.n, 2 * nb + 2));
  double nrm[2];
  auto residual = [&]() {
    k_resid0<P><<<nb, 256, 0, c.stream>>>(phi, rho, rc, r0, L0, c.red.p);
    k_sum2<<<1, 256, 0, c.stream>>>(c.red.p, nb, c.red.p + 2 * nb);
    FOPIS_HIP(hipMemcpyAsync(nrm, c.red.p + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FOPIS_HIP(hipStreamSynchronize(c.stream));
  };
  auto rel = [&]() { return nrm[1] > 0.0 ? std::sqrt(nrm[0] / nrm[1]) : std::sqrt(nrm[0]); };
  residual();
  int cycles = 0;
  while (rel() > tol && cycles < max_cycles) {
    if (nrm[1] == 0.0 && nrm[0] < 1e-300) break;
    k_zero<P><<<div_up((int64_t)m0.size, 256), 256, 0, c.stream>>>(E_<P>(m0) - off, (int64_t)m0.size);
    vcycle<P>(c, 0);
    k_add<P><<<grid_full(m0), 64, 0, c.stream>>>(phi, E_<P>(m0), L0);
    residual();
    ++cycles;
  }
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
  *relres_out = rel();
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
  FOPIS_HIP(hipStreamSynchronize(c.stream));
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
  c.relres_last = *relres;
  return cyc;
}

}  // namespace fopis
