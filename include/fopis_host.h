/*
 * fopis_host.h — C entry points of the host facade (fopis::GpuRootGrid,
 * fopis_amd/csrc/host.cpp) that sits on top of the hot-path ABI in fopis.h.
 *
 * The facade keeps the reference's input and set-up surface: it reads the
 * reference JSON schema (src/Misc/Initialize.cpp:158-421), normalises units
 * like Normalizer (include/normalizer.hpp), draws particles from the
 * reference's per-species mt19937 streams (src/Particle/ParticleType.cpp:11-27,
 * AmbientParticleType.cpp:60-157), reads spacecraft .obj files
 * (src/Spacecraft/ObjectUtils.cpp:9-333), generates boundary injection
 * (src/Grid/RootGrid/RootGridParticle.cpp:3-130) and photoelectron emission
 * (src/Spacecraft/Spacecraft.cpp:829-970) on the host, and drives the device
 * step through fopis.h. These entry points exist for the Python binding
 * (tests, benchmark); C++ callers use fopis::GpuRootGrid directly.
 */
#ifndef FOPIS_HOST_H
#define FOPIS_HOST_H

#include <stdint.h>

#include "fopis.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fopis_host fopis_host;

typedef struct fopis_host_options {
  int32_t load_particles;   /* 1: initial load (RootGrid.cpp:57-79); 0: none              */
  int32_t device_loader;    /* 1: counter-based device generator instead of mt19937        */
  uint64_t device_seed;     /* seed of the device generator                                */
  int32_t ppc_override;     /* >0 replaces particle_per_cell of every ambient species      */
  int32_t nx_override;      /* >0 replaces nx = ny = nz                                    */
  int32_t build_cmatrix;    /* 1: unit-charge solves on the device (RootGridObject.cpp:77) */
  int32_t initialize;       /* 1: Grid::initialize after the load                          */
  int32_t host_threads;     /* threads for the host mt19937 streams (0 -> 6)               */
  char cmatrix_cache[512];  /* directory of cached C matrices ("" = none)                  */
  int32_t no_device;        /* 1: parse, normalise and read geometry only (no GPU needed)  */
  int32_t device_injection; /* 1: new particles drawn on the device: boundary injection
                               (counts and residues still as RootGridParticle.cpp:3-45)
                               and photoemission (fopis_emit_photoelectrons)              */
  char resume_dir[512];     /* job_type "load": resume files directory ("" = <input dir>/resume) */
} fopis_host_options;

/* Reads the JSON input and builds the grid, objects and particles. */
int fopis_host_open(const char* json_path, const fopis_host_options* opt, fopis_host** out);
int fopis_host_close(fopis_host* h);
/* RootGrid::saveResumeData with Environment::saveInfo (RootGridIO.cpp:42-53,
 * Environment.cpp:421-438): writes dir/environment.h5 {dx, dt, timestep} and
 * dir/grid_id_0.h5 {<species>/{size, generated_counts, x, y, z, vx, vy, vz},
 * Potential/phi, Efield/{ex, ey, ez}, Bfield/{bx, by, bz}, <object>/charge_map}
 * in the reference's names, types and shapes. Needs libfopis_io.so (HDF5). */
int fopis_host_save_resume(fopis_host* h, const char* dir);
/* Environment::loadInfo + RootGrid::loadResumeData + Grid::initializeForLoad
 * (Environment.cpp:379-419, RootGridIO.cpp:25-40,55-131,183-225,264-271,
 * Grid.cpp:38-50): dx and dt must match the input; particles, the random
 * streams' generated counts, phi and the objects' charge maps are restored,
 * then rho and the reference E are recomputed. */
int fopis_host_load_resume(fopis_host* h, const char* dir);
/* The device context underneath (owned by the facade). */
fopis_ctx* fopis_host_ctx(fopis_host* h);
/* Normalised description the facade built (species order = reference ids). */
int fopis_host_desc(fopis_host* h, fopis_desc* out);
/* One RootGrid::mainLoopES: emission and injection on the host, the step on the device. */
int fopis_host_main_loop_es(fopis_host* h);
int fopis_host_timestep(fopis_host* h);
/* Redraws every ambient species' initial load on the device with a new seed
 * (the RootGrid constructor's load, RootGrid.cpp:57-79; device loader only). */
int fopis_host_reload(fopis_host* h, uint64_t seed);
/* Particles the next injection call would add, without staging them
 * (for parity of the host generator): returns the count, fills up to cap. */
int64_t fopis_host_generate_injection(fopis_host* h, int s, int64_t cap, double* x, double* y, double* z, double* vx,
                                      double* vy, double* vz);
/* Photoelectrons of species s the next emission would add, drawn on the host
 * from the reference's streams without staging them (Spacecraft.cpp:829-900:
 * positions after this step's move, those leaving the domain dropped), and
 * object 0's surface charge change for s (ncmat doubles) and current change.
 * For parity of the host emitter: returns the count, fills up to cap. */
int64_t fopis_host_generate_emission(fopis_host* h, int s, int64_t cap, double* x, double* y, double* z, double* vx,
                                     double* vy, double* vz, double* dq, double* dcurrent);
/* The initial load of species s as the host draws it from the reference's
 * streams (RootGrid.cpp:57-79), without uploading: returns the count. */
int64_t fopis_host_generate_initial(fopis_host* h, int s, int64_t cap, double* x, double* y, double* z, double* vx,
                                    double* vy, double* vz);
/* Particle energy normalisation and flux of species s (AmbientParticleType::calcFlux). */
int fopis_host_flux(fopis_host* h, int s, double* flux6);
const char* fopis_host_last_error(void);
/* Object geometry as the facade read it: the (nx+1)^3 cell mask and the
 * glue-relative ijk of the ncmat surface nodes in cmat order. */
int fopis_host_object(fopis_host* h, int obj, uint8_t* mask, int32_t* ncmat, int32_t* cmat_ijk, int32_t cap);
/* Host wall time (ms) of the last step spent generating emission and injection. */
double fopis_host_generate_ms(fopis_host* h);

#ifdef __cplusplus
}
#endif

#endif /* FOPIS_HOST_H */
