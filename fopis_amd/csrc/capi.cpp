// extern "C" boundary of libfopis_hip.so (include/fopis.h): argument checks,
// status codes, the context lifecycle and the order of the step's stages.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "fopis_internal.hpp"

namespace fopis {

static thread_local std::string g_last_error;

void timer_begin(Ctx& c, const char* name, hipEvent_t* ev) {
  if (!c.timing) return;
  FOPIS_HIP(hipEventCreate(ev));
  FOPIS_HIP(hipEventRecord(*ev, c.stream));
  (void)name;
}

void timer_end(Ctx& c, const char* name, hipEvent_t ev) {
  if (!c.timing || !ev) return;
  hipEvent_t e2;
  if (hipEventCreate(&e2) != hipSuccess) return;
  if (hipEventRecord(e2, c.stream) != hipSuccess) return;
  PhaseTimer& t = c.timers[name];
  t.pending.emplace_back(ev, e2);
  if (t.pending.size() > 512) timers_collect(c);
}

void timers_collect(Ctx& c) {
  for (auto& kv : c.timers) {
    PhaseTimer& t = kv.second;
    for (auto& pr : t.pending) {
      if (hipEventSynchronize(pr.second) == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
          t.total_ms += ms;
          t.n += 1;
        }
      }
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    t.pending.clear();
  }
}

static void check_desc(const fopis_desc& d) {
  if (d.nx < 4 || d.ny < 4 || d.nz < 4) throw Error(FOPIS_E_ARG, "nx, ny, nz must be >= 4");
  if (d.nspecies < 1 || d.nspecies > FOPIS_MAX_SPECIES) throw Error(FOPIS_E_ARG, "nspecies out of range");
  if (std::strlen(d.boundary) != 6) throw Error(FOPIS_E_ARG, "boundary must have 6 letters");
  for (int q = 0; q < 6; ++q) {
    const char b = d.boundary[q];
    if (b == 'P') throw Error(FOPIS_E_ARG, "periodic faces are not supported by this build");
    if (b != 'D' && b != 'N')
      throw Error(FOPIS_E_ARG, std::string("Unknown boundary condition type was used: boundary = ") + b);
  }
  for (int s = 0; s < d.nspecies; ++s)
    if (!(d.species[s].mass > 0.0)) throw Error(FOPIS_E_ARG, "species mass must be positive");
  if (!(d.eps0 > 0.0)) throw Error(FOPIS_E_ARG, "eps0 must be positive");
  if (d.precision != FOPIS_FP64 && d.precision != FOPIS_FP32) throw Error(FOPIS_E_ARG, "bad precision");
  if (d.mg_precision != FOPIS_FP64 && d.mg_precision != FOPIS_FP32) throw Error(FOPIS_E_ARG, "bad mg precision");
}

// With the learned guess on, the first Poisson solve of a step starts from
// the network's potential instead of the previous step's.
// The first solve of a step and where it starts: the network's potential
// (guess mode 1), or the warm start (in the fused step with conductors, the
// previous potential less its conductor response), which guess mode 2
// corrects by the network's potential for its residual and mode 3 records.
static void first_solve(Ctx& c, bool carry, double* rr) {
  Phase ph(c, "poisson_first");
  const int mode = guess_mode(c);
  if (mode == 1) {
    guess_apply(c);
  } else {
    if (carry && c.nobj > 0) objects_carry_first(c);
    if (mode >= 2) guess_before_solve(c);
  }
  poisson_solve(c, c.d.poisson_tol, c.d.mg_max_cycles, rr, nullptr);
  if (mode == 3) guess_after_solve(c);
}

// In the fused step (carry), the two solves of a step with conductors start
// from the warm starts objects_carry_* forms.
static void update_rho(Ctx& c, bool push, bool carry = false) {
  particles_deposit_bin(c, push);
  objects_sum_whole_charge(c);
  if (c.nobj > 0) {
    double rr = 0.0;
    first_solve(c, carry, &rr);
    c.mg_first_cycles_total += c.mg_cycles_last;
    c.relres0_first += c.relres0_last;
    ++c.n_first;
    // the object charge is redistributed from this potential: it must have converged
    if (!(rr <= c.d.poisson_tol))
      throw Error(FOPIS_E_NOCONV, "first Poisson solve stopped at relres " + std::to_string(rr));
    objects_redistribute(c);
    if (carry) objects_carry_second(c);
  }
}

}  // namespace fopis

using namespace fopis;

struct fopis_ctx {
  Ctx c;
};

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return FOPIS_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return FOPIS_E_STATE;
  }
}

#define NEED_CTX                                 \
  if (!ctx) {                                    \
    g_last_error = "null context";               \
    return FOPIS_E_ARG;                          \
  }

extern "C" {

const char* fopis_last_error(void) { return g_last_error.c_str(); }

int fopis_device_info(char* buf, int buflen) {
  return guarded([&] {
    int dev = 0;
    FOPIS_HIP(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    FOPIS_HIP(hipGetDeviceProperties(&pr, dev));
    const std::string name = pr.name[0] ? std::string(pr.name) + " " : std::string();
    std::string s = std::string("libfopis_hip 0.1; device ") + name + "(" + pr.gcnArchName + "), " +
                    std::to_string(pr.multiProcessorCount) + " CUs, " +
                    std::to_string((long long)(pr.totalGlobalMem >> 20)) + " MiB";
    if (buf && buflen > 0) {
      std::strncpy(buf, s.c_str(), buflen - 1);
      buf[buflen - 1] = 0;
    }
  });
}

int fopis_create(const fopis_desc* desc, fopis_ctx** out) {
  if (!desc || !out) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  *out = nullptr;
  fopis_ctx* h = new fopis_ctx();
  const int rc = guarded([&] {
    Ctx& c = h->c;
    c.d = *desc;
    check_desc(c.d);
    if (c.d.mg_omega == 0.0) c.d.mg_omega = 1.25;
    if (c.d.mg_pre == 0) c.d.mg_pre = 2;
    if (c.d.mg_post == 0) c.d.mg_post = 2;
    if (c.d.mg_max_cycles == 0) c.d.mg_max_cycles = 100;
    if (c.d.poisson_tol == 0.0) c.d.poisson_tol = 1e-8;
    if (c.d.capacity_slack == 0.0) c.d.capacity_slack = 1.25;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(FOPIS_E_DEVICE, "no HIP device");
    c.S = c.d.nspecies;
    c.nx = c.d.nx;
    c.ny = c.d.ny;
    c.nz = c.d.nz;
    c.cx = c.nx + 2;
    c.cy = c.ny + 2;
    c.cz = c.nz + 2;
    c.Nc[0] = c.nx - 1;
    c.Nc[1] = c.ny - 1;
    c.Nc[2] = c.nz - 1;
    for (int q = 0; q < 3; ++q) c.nt[q] = div_up(c.Nc[q], kTile);
    c.ntiles = c.nt[0] * c.nt[1] * c.nt[2];
    c.nnode = (size_t)c.cx * c.cy * c.cz;
    c.ncell_d = (size_t)(c.cx - 1) * (c.cy - 1) * (c.cz - 1);
    c.fp32 = c.d.precision == FOPIS_FP32;
    FOPIS_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    fields_init(c);
    particles_init(c);
    poisson_init(c);
  });
  if (rc != FOPIS_OK) {
    if (h->c.stream) (void)hipStreamDestroy(h->c.stream);
    delete h;
    return rc;
  }
  *out = h;
  return FOPIS_OK;
}

int fopis_destroy(fopis_ctx* ctx) {
  if (!ctx) return FOPIS_OK;
  (void)hipStreamSynchronize(ctx->c.stream);
  timers_collect(ctx->c);
  hipStream_t st = ctx->c.stream;
  guess_free(ctx->c);
  poisson_free(ctx->c);
  delete ctx;
  if (st) (void)hipStreamDestroy(st);
  return FOPIS_OK;
}

int fopis_set_particles(fopis_ctx* ctx, int s, int64_t n, const double* x, const double* y, const double* z,
                        const double* vx, const double* vy, const double* vz) {
  NEED_CTX
  return guarded([&] {
    if (s < 0 || s >= ctx->c.S) throw Error(FOPIS_E_ARG, "species out of range");
    if (n < 0 || (n > 0 && (!x || !y || !z || !vx || !vy || !vz))) throw Error(FOPIS_E_ARG, "bad particle arrays");
    const double* h[6] = {x, y, z, vx, vy, vz};
    particles_set(ctx->c, s, n, h);
  });
}

int fopis_load_maxwellian(fopis_ctx* ctx, int s, int64_t n, uint64_t seed, const double* lo, const double* hi,
                          const double* drift, double sigma) {
  NEED_CTX
  return guarded([&] {
    if (s < 0 || s >= ctx->c.S || n < 0 || !lo || !hi || !drift || !(sigma >= 0.0))
      throw Error(FOPIS_E_ARG, "bad argument");
    particles_load_maxwellian(ctx->c, s, n, seed, lo, hi, drift, sigma);
  });
}

int fopis_num_particles(fopis_ctx* ctx, int s, int64_t* n) {
  NEED_CTX
  return guarded([&] {
    if (s < 0 || s >= ctx->c.S || !n) throw Error(FOPIS_E_ARG, "bad argument");
    *n = particles_count(ctx->c, s);
  });
}

int fopis_get_particles(fopis_ctx* ctx, int s, int64_t cap, double* x, double* y, double* z, double* vx, double* vy,
                        double* vz, int64_t* n) {
  NEED_CTX
  return guarded([&] {
    if (s < 0 || s >= ctx->c.S || !n) throw Error(FOPIS_E_ARG, "bad argument");
    double* h[6] = {x, y, z, vx, vy, vz};
    particles_get(ctx->c, s, cap, h, n);
  });
}

int fopis_add_particles(fopis_ctx* ctx, int s, int64_t n, const double* x, const double* y, const double* z,
                        const double* vx, const double* vy, const double* vz) {
  NEED_CTX
  return guarded([&] {
    if (s < 0 || s >= ctx->c.S) throw Error(FOPIS_E_ARG, "species out of range");
    if (n < 0 || (n > 0 && (!x || !y || !z || !vx || !vy || !vz))) throw Error(FOPIS_E_ARG, "bad particle arrays");
    const double* h[6] = {x, y, z, vx, vy, vz};
    particles_stage(ctx->c, s, n, h);
  });
}

int fopis_sort_by_cell(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] { particles_sort_by_cell(ctx->c); });
}

int fopis_get_cell_keys(fopis_ctx* ctx, int s, int64_t cap, int32_t* keys, int64_t* n) {
  NEED_CTX
  return guarded([&] {
    if (s < 0 || s >= ctx->c.S || !n) throw Error(FOPIS_E_ARG, "bad argument");
    particles_cell_keys(ctx->c, s, cap, keys, n);
  });
}

int fopis_update_particle_velocity_es(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] { particles_velocity(ctx->c); });
}

int fopis_update_particle_position_es(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] { particles_position(ctx->c); });
}

int fopis_update_density(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] { particles_density_only(ctx->c); });
}

int fopis_update_rho(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] { update_rho(ctx->c, false); });
}

int fopis_solve_poisson(fopis_ctx* ctx, int32_t* cycles, double* relres) {
  NEED_CTX
  int rc = guarded([&] {
    Ctx& c = ctx->c;
    double rr = 0.0;
    const int cyc = poisson_solve(c, c.d.poisson_tol, c.d.mg_max_cycles, &rr, nullptr);
    if (cycles) *cycles = cyc;
    if (relres) *relres = rr;
    if (!(rr <= c.d.poisson_tol)) throw Error(FOPIS_E_NOCONV, "Poisson solve stopped at relres " + std::to_string(rr));
  });
  return rc;
}

int fopis_update_efield(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] { fields_efield(ctx->c); });
}

int fopis_initialize(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] {
    Ctx& c = ctx->c;
    update_rho(c, false);
    double rr = 0.0;
    poisson_solve(c, c.d.poisson_tol, c.d.mg_max_cycles, &rr, nullptr);
    fields_efield(c);
  });
}

int fopis_main_loop_es(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] {
    Ctx& c = ctx->c;
    objects_reset_step(c);
    update_rho(c, true, true);
    double rr = 0.0;
    if (c.nobj == 0) {
      first_solve(c, false, &rr);
      c.mg_first_cycles_total += c.mg_cycles_last;
      c.relres0_first += c.relres0_last;
      ++c.n_first;
    } else {
      poisson_solve(c, c.d.poisson_tol, c.d.mg_max_cycles, &rr, nullptr);
      c.relres0_other += c.relres0_last;
      ++c.n_other;
      objects_carry_end(c);
    }
    if (!(rr <= c.d.poisson_tol)) throw Error(FOPIS_E_NOCONV, "Poisson solve stopped at relres " + std::to_string(rr));
    fields_efield(c);
    ++c.timestep;
  });
}

int fopis_get_field(fopis_ctx* ctx, int which, int s, double* out) {
  NEED_CTX
  return guarded([&] {
    if (!out) throw Error(FOPIS_E_ARG, "null output");
    fields_get(ctx->c, which, s, out);
  });
}

int fopis_set_field(fopis_ctx* ctx, int which, int s, const double* in) {
  NEED_CTX
  return guarded([&] {
    if (!in) throw Error(FOPIS_E_ARG, "null input");
    fields_set(ctx->c, which, s, in);
  });
}

int fopis_energy(fopis_ctx* ctx, double* out) {
  NEED_CTX
  return guarded([&] {
    if (!out) throw Error(FOPIS_E_ARG, "null output");
    out[0] = particles_energy(ctx->c);
    out[1] = fields_energy(ctx->c);
  });
}

int fopis_solver_start_residuals(fopis_ctx* ctx, double* out) {
  NEED_CTX
  return guarded([&] {
    if (!out) throw Error(FOPIS_E_ARG, "null output");
    const Ctx& c = ctx->c;
    out[0] = c.relres0_last;
    out[1] = c.n_first ? c.relres0_first / (double)c.n_first : 0.0;
    out[2] = c.n_other ? c.relres0_other / (double)c.n_other : 0.0;
  });
}

int fopis_energy_noalloc(fopis_ctx* ctx, double* out) {
  NEED_CTX
  return guarded([&] {
    if (!out) throw Error(FOPIS_E_ARG, "null output");
    energy_noalloc(ctx->c, out);
  });
}

int fopis_set_object(fopis_ctx* ctx, int obj, const uint8_t* cell_mask, int32_t ncmat, const int32_t* cmat_ijk,
                     int32_t fixed, double fixed_potential) {
  NEED_CTX
  return guarded([&] {
    if (!cell_mask || (ncmat > 0 && !cmat_ijk)) throw Error(FOPIS_E_ARG, "null object arrays");
    objects_set(ctx->c, obj, cell_mask, ncmat, cmat_ijk, fixed, fixed_potential);
  });
}

int fopis_build_cmatrix(fopis_ctx* ctx, int obj) {
  NEED_CTX
  return guarded([&] {
    if (obj < 0 || obj >= ctx->c.nobj) throw Error(FOPIS_E_ARG, "object index out of range");
    objects_build_cmatrix(ctx->c, obj);
  });
}

int fopis_get_cmatrix(fopis_ctx* ctx, int obj, double* out) {
  NEED_CTX
  return guarded([&] {
    if (obj < 0 || obj >= ctx->c.nobj || !out) throw Error(FOPIS_E_ARG, "bad argument");
    objects_get_cmatrix(ctx->c, obj, out);
  });
}

int fopis_set_cmatrix(fopis_ctx* ctx, int obj, const double* in) {
  NEED_CTX
  return guarded([&] {
    if (obj < 0 || obj >= ctx->c.nobj || !in) throw Error(FOPIS_E_ARG, "bad argument");
    objects_set_cmatrix(ctx->c, obj, in);
  });
}

int fopis_add_object_charge(fopis_ctx* ctx, int obj, const double* dq, const double* dcurrent) {
  NEED_CTX
  return guarded([&] {
    if (obj < 0 || obj >= ctx->c.nobj) throw Error(FOPIS_E_ARG, "object index out of range");
    objects_add_charge(ctx->c, obj, dq, dcurrent);
  });
}

int fopis_get_object_state(fopis_ctx* ctx, int obj, double* potential, double* total_charge, double* current,
                           double* charge_map) {
  NEED_CTX
  return guarded([&] {
    if (obj < 0 || obj >= ctx->c.nobj) throw Error(FOPIS_E_ARG, "object index out of range");
    objects_state(ctx->c, obj, potential, total_charge, current, charge_map);
  });
}

int fopis_object_unattributed(fopis_ctx* ctx, int obj, int64_t* out) {
  NEED_CTX
  return guarded([&] {
    if (obj < 0 || obj >= ctx->c.nobj || !out) throw Error(FOPIS_E_ARG, "bad argument");
    objects_unattributed(ctx->c, obj, out);
  });
}

int fopis_set_object_charge_map(fopis_ctx* ctx, int obj, const double* charge_map) {
  NEED_CTX
  return guarded([&] {
    if (obj < 0 || obj >= ctx->c.nobj || !charge_map) throw Error(FOPIS_E_ARG, "bad argument");
    objects_set_charge_map(ctx->c, obj, charge_map);
  });
}

int fopis_phase_time(fopis_ctx* ctx, const char* phase, double* avg_ms, int64_t* n) {
  NEED_CTX
  return guarded([&] {
    if (!phase || !avg_ms || !n) throw Error(FOPIS_E_ARG, "null argument");
    timers_collect(ctx->c);
    auto it = ctx->c.timers.find(phase);
    if (it == ctx->c.timers.end() || it->second.n == 0) {
      *avg_ms = 0.0;
      *n = 0;
      return;
    }
    *avg_ms = it->second.total_ms / (double)it->second.n;
    *n = it->second.n;
  });
}

int fopis_reset_timers(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] {
    timers_collect(ctx->c);
    ctx->c.timers.clear();
    ctx->c.mg_cycles_total = 0;
    ctx->c.mg_solves_total = 0;
    ctx->c.mg_updates_total = 0;
    ctx->c.mg_pair_passes_total = 0;
    ctx->c.mg_first_cycles_total = 0;
    ctx->c.relres0_first = ctx->c.relres0_other = 0.0;
    ctx->c.n_first = ctx->c.n_other = 0;
  });
}

int fopis_solver_stats(fopis_ctx* ctx, int32_t* last_cycles, double* last_relres, int64_t* total_cycles,
                       int64_t* total_solves) {
  NEED_CTX
  return guarded([&] {
    const Ctx& c = ctx->c;
    if (last_cycles) *last_cycles = c.mg_cycles_last;
    if (last_relres) *last_relres = c.relres_last;
    if (total_cycles) *total_cycles = c.mg_cycles_total;
    if (total_solves) *total_solves = c.mg_solves_total;
  });
}

int fopis_moments(fopis_ctx* ctx, int s, double* out) {
  NEED_CTX
  if (!out || s < 0 || s >= ctx->c.S) {
    g_last_error = "bad species or null output";
    return FOPIS_E_ARG;
  }
  return guarded([&] { particles_moments(ctx->c, s, out); });
}

int fopis_solver_passes(fopis_ctx* ctx, int64_t* out) {
  NEED_CTX
  if (!out) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] {
    out[0] = ctx->c.mg_updates_total;
    out[1] = ctx->c.mg_pair_passes_total;
  });
}

int fopis_first_solve_cycles(fopis_ctx* ctx, int64_t* total_cycles) {
  NEED_CTX
  if (!total_cycles) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] { *total_cycles = ctx->c.mg_first_cycles_total; });
}

int fopis_inject_maxwellian(fopis_ctx* ctx, int s, const int64_t* count6, uint64_t seed, const double* drift,
                            double vth, double sigma) {
  NEED_CTX
  if (!count6 || !drift || s < 0 || s >= ctx->c.S) {
    g_last_error = "bad species or null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] { particles_inject(ctx->c, s, count6, seed, drift, vth, sigma); });
}

int fopis_emit_photoelectrons(fopis_ctx* ctx, int obj, int s, const double* shine, double amount, double vth,
                              uint64_t seed) {
  NEED_CTX
  return guarded([&] {
    Ctx& c = ctx->c;
    if (obj < 0 || obj >= c.nobj || !c.obj[obj].defined) throw Error(FOPIS_E_ARG, "object not defined");
    if (s < 0 || s >= c.S || !shine) throw Error(FOPIS_E_ARG, "bad species or null shine vector");
    int axis = -1;
    for (int a = 0; a < 3; ++a)
      if (std::fabs(shine[a]) == 1.0) {
        axis = a;
        break;
      }
    if (axis < 0)
      throw Error(FOPIS_E_ARG, "shine vector must be [+-1,0,0], [0,+-1,0] or [0,0,+-1]");
    if (!(amount >= 0.0) || amount > 1e6) throw Error(FOPIS_E_ARG, "emission amount per node out of range");
    if (!(vth > 0.0)) throw Error(FOPIS_E_ARG, "vth must be positive");
    // the reference's loop `for (int n = 0; n < amount; ++n)` runs ceil(amount) times
    const int per_node = (int)std::ceil(amount);
    particles_emit(c, obj, s, axis, shine[axis] > 0.0 ? 1 : -1, per_node, vth, seed);
  });
}

int fopis_guess_init(fopis_ctx* ctx, uint64_t seed) {
  NEED_CTX
  return guarded([&] { guess_init(ctx->c, seed); });
}

int fopis_guess_num_params(fopis_ctx* ctx, int64_t* n) {
  NEED_CTX
  if (!n) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] { *n = guess_num_params(ctx->c); });
}

int fopis_guess_get_params(fopis_ctx* ctx, float* params, double* out_scale) {
  NEED_CTX
  if (!params) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] { guess_get_params(ctx->c, params, out_scale); });
}

int fopis_guess_set_params(fopis_ctx* ctx, const float* params, double out_scale) {
  NEED_CTX
  if (!params) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] { guess_set_params(ctx->c, params, out_scale); });
}

int fopis_guess_set_mode(fopis_ctx* ctx, int mode) {
  NEED_CTX
  return guarded([&] { guess_set_mode(ctx->c, mode); });
}

int fopis_guess_recorded(fopis_ctx* ctx, int32_t* n) {
  NEED_CTX
  return guarded([&] {
    if (!n) throw Error(FOPIS_E_ARG, "null output");
    *n = guess_recorded(ctx->c);
  });
}

int fopis_guess_set_loss(fopis_ctx* ctx, int kind) {
  NEED_CTX
  return guarded([&] { guess_set_loss(ctx->c, kind); });
}

int fopis_guess_apply(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] { guess_apply(ctx->c); });
}

int fopis_guess_dataset_alloc(fopis_ctx* ctx, int32_t capacity) {
  NEED_CTX
  return guarded([&] { guess_dataset_alloc(ctx->c, capacity); });
}

int fopis_guess_dataset_store(fopis_ctx* ctx, int32_t slot, double* phi_hat_rms) {
  NEED_CTX
  return guarded([&] {
    const double r = guess_dataset_store(ctx->c, slot);
    if (phi_hat_rms) *phi_hat_rms = r;
  });
}

int fopis_guess_dataset_get(fopis_ctx* ctx, int32_t slot, uint16_t* rho_hat, uint8_t* mask, uint16_t* phi_hat) {
  NEED_CTX
  return guarded([&] { guess_dataset_get(ctx->c, slot, rho_hat, mask, phi_hat); });
}

int fopis_guess_dataset_set(fopis_ctx* ctx, int32_t slot, const uint16_t* rho_hat, const uint8_t* mask,
                            const uint16_t* phi_hat) {
  NEED_CTX
  if (!rho_hat || !mask || !phi_hat) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] { guess_dataset_set(ctx->c, slot, rho_hat, mask, phi_hat); });
}

int fopis_guess_grad(fopis_ctx* ctx, int32_t slot, double* loss, float* grad) {
  NEED_CTX
  return guarded([&] {
    const double l = guess_grad(ctx->c, slot, grad);
    if (loss) *loss = l;
  });
}

int fopis_guess_train(fopis_ctx* ctx, int32_t npairs, int32_t epochs, double lr, double lr_end, uint64_t seed,
                      double* epoch_loss) {
  NEED_CTX
  return guarded([&] { guess_train(ctx->c, npairs, epochs, lr, lr_end, seed, epoch_loss); });
}

int fopis_guess_eval(fopis_ctx* ctx, int32_t begin, int32_t end, double* loss) {
  NEED_CTX
  if (!loss) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] { *loss = guess_eval(ctx->c, begin, end); });
}

int fopis_guess_synth_boxes(fopis_ctx* ctx, int32_t nbox, const int32_t* lohi, const double* charge, uint64_t seed) {
  NEED_CTX
  if (nbox > 0 && (!lohi || !charge)) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] {
    if (nbox == 0)
      guess_clear_boxes(ctx->c);
    else
      guess_synth_boxes(ctx->c, nbox, lohi, charge, seed);
  });
}

int fopis_nn_conv3d(int32_t n0, int32_t n1, int32_t n2, int32_t cin, int32_t cout, int32_t up, int32_t relu,
                    const uint16_t* x, const float* w, const float* b, const uint16_t* add, uint16_t* y, int32_t reps,
                    float* ms) {
  if (!x || !w || !b || !y) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] { nn_conv3d(n0, n1, n2, cin, cout, up, relu, x, w, b, add, y, reps, ms); });
}

int fopis_nn_conv3d_backward(int32_t n0, int32_t n1, int32_t n2, int32_t cin, int32_t cout, int32_t up,
                             const uint16_t* x, const float* w, const uint16_t* gy, float* dw, float* db,
                             uint16_t* dx) {
  if (!x || !w || !gy || !dw || !db) {
    g_last_error = "null argument";
    return FOPIS_E_ARG;
  }
  return guarded([&] { nn_conv3d_backward(n0, n1, n2, cin, cout, up, x, w, gy, dw, db, dx); });
}

int fopis_particle_ledger(fopis_ctx* ctx, int s, int64_t* out) {
  NEED_CTX
  return guarded([&] {
    if (s < 0 || s >= ctx->c.S) throw Error(FOPIS_E_ARG, "species out of range");
    if (!out) throw Error(FOPIS_E_ARG, "null output");
    particles_ledger(ctx->c, s, out);
  });
}

int fopis_synchronize(fopis_ctx* ctx) {
  NEED_CTX
  return guarded([&] { FOPIS_HIP(hipStreamSynchronize(ctx->c.stream)); });
}

}  // extern "C"
