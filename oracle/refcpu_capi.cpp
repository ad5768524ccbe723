// ORACLE — test infrastructure only (see refcpu.hpp). extern "C" wrapper so
// the tests and the CPU-baseline leg of the benchmark can drive the
// restatement through ctypes.
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>

#include "refcpu.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

using refcpu::Grid;

static thread_local std::string g_err;

template <class F>
static int guard(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

static std::vector<double>* fieldPtr(Grid* g, int which, int s) {
  switch (which) {
    case 0: return &g->phi;
    case 1: return (s >= 0 && s <= g->S) ? &g->rho[s] : nullptr;
    case 2: return (s >= 0 && s < g->S) ? &g->density[s] : nullptr;
    case 3: return &g->ex;
    case 4: return &g->ey;
    case 5: return &g->ez;
    case 6: return &g->exref;
    case 7: return &g->eyref;
    case 8: return &g->ezref;
    case 9: return &g->poisson_residual;
    case 10: return &g->poisson_error;
    default: return nullptr;
  }
}

extern "C" {

const char* rc_error() { return g_err.c_str(); }

void rc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void* rc_create(const refcpu::Params* p) {
  try {
    return new Grid(*p);
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void rc_destroy(void* h) { delete static_cast<Grid*>(h); }

#define G static_cast<Grid*>(h)
int rc_load_initial_particles(void* h) { return guard([&] { G->loadInitialParticles(); return 0; }); }
int rc_initialize(void* h) { return guard([&] { G->initialize(); return 0; }); }
int rc_main_loop(void* h) {
  return guard([&] {
    G->mainLoopES();
    ++G->timestep;
    return 0;
  });
}
int rc_update_velocity(void* h) { return guard([&] { G->updateParticleVelocityES(); return 0; }); }
int rc_update_position(void* h) { return guard([&] { G->updateParticlePositionES(); return 0; }); }
int rc_inject(void* h) { return guard([&] { G->injectParticlesFromBoundary(); return 0; }); }
int rc_emit(void* h) { return guard([&] { G->emitParticlesFromObjects(); return 0; }); }
int rc_update_density(void* h) { return guard([&] { G->updateDensity(); return 0; }); }
int rc_update_rho(void* h) { return guard([&] { G->updateRho(); return 0; }); }
int rc_deposit_only(void* h) { return guard([&] { G->depositOnly(); return 0; }); }
int rc_solve_poisson(void* h) { return guard([&] { G->solvePoisson(); return G->last_sor_iterations; }); }
int rc_solve_poisson_relres(void* h, double tol, int maxloop) {
  return guard([&] { return G->solvePoissonToRelres(tol, maxloop); });
}
int rc_do_psor(void* h, int n) {
  return guard([&] {
    for (int q = 0; q < n; ++q) G->doPSOR();
    return 0;
  });
}
int rc_set_boundary_phi(void* h) { return guard([&] { G->setBoundaryConditionPhi(); return 0; }); }
double rc_relres(void* h) { return G->relres(); }
double rc_check_phi_residual(void* h) { return G->checkPhiResidual(); }
int rc_update_efield(void* h) { return guard([&] { G->updateEfield(); return 0; }); }
int rc_remove_invalid(void* h) { return guard([&] { G->removeInvalidParticles(); return 0; }); }
int rc_build_cmatrix(void* h) { return guard([&] { G->buildCmatrix(); return 0; }); }
int rc_initialize_charge_offset(void* h) { return guard([&] { G->initializeChargeMapOffset(); return 0; }); }
void rc_set_timestep(void* h, int t) { G->timestep = t; }
int rc_get_timestep(void* h) { return G->timestep; }
double rc_eps0(void* h) { return G->u.eps0; }
double rc_particle_energy(void* h) { return G->particleEnergy(); }
double rc_efield_energy(void* h) { return G->efieldEnergy(); }

// out: [charge, mass, size, charge_of_superparticle, thermal_velocity, deviation, drift x,y,z]
int rc_species_info(void* h, int s, double* out) {
  if (s < 0 || s >= G->S) return -1;
  const auto& t = G->sp[s];
  out[0] = t.charge;
  out[1] = t.mass;
  out[2] = t.size;
  out[3] = t.chargeOfSuperParticle();
  out[4] = t.thermalVelocity(G->u);
  out[5] = t.deviation(G->u);
  for (int a = 0; a < 3; ++a) out[6 + a] = t.drift[a];
  return 0;
}

void rc_static_b(void* h, double* out) {
  for (int a = 0; a < 3; ++a) out[a] = G->b[a];
}

int rc_calc_flux(void* h, int s, double* out) {
  if (s < 0 || s >= G->S) return -1;
  auto f = refcpu::calcFlux(G->sp[s], G->u);
  for (int a = 0; a < 6; ++a) out[a] = f[a];
  return 0;
}

int64_t rc_num_particles(void* h, int s) { return (s >= 0 && s < G->S) ? (int64_t)G->particles[s].size() : -1; }

int rc_get_particles(void* h, int s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                     int32_t* valid) {
  if (s < 0 || s >= G->S) return -1;
  const auto& pa = G->particles[s];
  for (size_t q = 0; q < pa.size(); ++q) {
    x[q] = pa[q].x;
    y[q] = pa[q].y;
    z[q] = pa[q].z;
    vx[q] = pa[q].vx;
    vy[q] = pa[q].vy;
    vz[q] = pa[q].vz;
    if (valid) valid[q] = pa[q].isValid;
  }
  return 0;
}

int rc_set_particles(void* h, int s, int64_t n, const double* x, const double* y, const double* z,
                     const double* vx, const double* vy, const double* vz, const int32_t* valid) {
  if (s < 0 || s >= G->S) return -1;
  auto& pa = G->particles[s];
  pa.resize((size_t)n);
  for (int64_t q = 0; q < n; ++q) pa[q] = refcpu::Particle{x[q], y[q], z[q], vx[q], vy[q], vz[q], s, valid ? valid[q] : 1};
  return 0;
}

int64_t rc_field_size(void* h, int which, int s) {
  auto* f = fieldPtr(G, which, s);
  return f ? (int64_t)f->size() : -1;
}

int rc_get_field(void* h, int which, int s, double* out) {
  auto* f = fieldPtr(G, which, s);
  if (!f) return -1;
  std::memcpy(out, f->data(), f->size() * sizeof(double));
  return 0;
}

int rc_set_field(void* h, int which, int s, const double* in) {
  auto* f = fieldPtr(G, which, s);
  if (!f) return -1;
  std::memcpy(f->data(), in, f->size() * sizeof(double));
  return 0;
}

int rc_num_objects(void* h) { return (int)G->objects.size(); }

// out: [ncmat, potential, total_charge, total_cmat_value, defined]
int rc_object_info(void* h, int o, double* out) {
  if (o < 0 || o >= (int)G->objects.size()) return -1;
  const auto& sc = G->objects[o];
  out[0] = sc.num_cmat;
  out[1] = sc.potential;
  out[2] = sc.total_charge;
  out[3] = sc.total_cmat_value;
  out[4] = sc.defined ? 1.0 : 0.0;
  return 0;
}

// glue-relative node ijk of every cmat index, 3*ncmat ints
int rc_object_cmat_nodes(void* h, int o, int32_t* ijk) {
  if (o < 0 || o >= (int)G->objects.size()) return -1;
  for (const auto& nd : G->objects[o].cmat_rel) {
    ijk[3 * nd.first + 0] = nd.second.i;
    ijk[3 * nd.first + 1] = nd.second.j;
    ijk[3 * nd.first + 2] = nd.second.k;
  }
  return 0;
}

// glue-relative cell map, (nx+1)*(ny+1)*(nz+1) bytes
int rc_object_cell_map(void* h, int o, uint8_t* out) {
  if (o < 0 || o >= (int)G->objects.size()) return -1;
  const auto& m = G->objects[o].cell_map;
  std::memcpy(out, m.data(), m.size());
  return 0;
}

// element (col,row) at col*n+row; mode 0 = get, 1 = set (then totals are refreshed)
int rc_object_cmatrix(void* h, int o, double* buf, int mode) {
  if (o < 0 || o >= (int)G->objects.size()) return -1;
  auto& sc = G->objects[o];
  if (mode == 0) {
    std::memcpy(buf, sc.cmat.data(), sc.cmat.size() * sizeof(double));
  } else {
    std::memcpy(sc.cmat.data(), buf, sc.cmat.size() * sizeof(double));
    sc.updateTotalCmatValue();
  }
  return 0;
}

// which 0 = charge_map, 1 = temporary_charge_map, 2 = current; mode 0 get, 1 set
int rc_object_charge(void* h, int o, int which, double* buf, int mode) {
  if (o < 0 || o >= (int)G->objects.size()) return -1;
  auto& sc = G->objects[o];
  std::vector<double>* v = which == 0 ? &sc.charge_map : (which == 1 ? &sc.temporary_charge_map : &sc.current);
  if (mode == 0) std::memcpy(buf, v->data(), v->size() * sizeof(double));
  else std::memcpy(v->data(), buf, v->size() * sizeof(double));
  return 0;
}

int rc_object_redistribute(void* h, int o) {
  if (o < 0 || o >= (int)G->objects.size()) return -1;
  return guard([&] {
    G->objects[o].redistributeCharge(G->rho[0], G->phi, G->cy, G->cz, G->S);
    return 0;
  });
}

// Known answer for the engine: the 10000th output of a default-seeded
// std::mt19937 is 4123659995 ([rand.predef] in the C++ standard).
uint32_t rc_mt19937_10000th() {
  std::mt19937 m;
  m.discard(9999);
  return (uint32_t)m();
}

int64_t rc_params_size() { return (int64_t)sizeof(refcpu::Params); }
}
