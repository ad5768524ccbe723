/*
 * fopis.h — C ABI of the MI355X electrostatic PIC step (libfopis_hip.so).
 *
 * The reference (hsimyu/fopis) has no foreign-function interface: its hot path
 * sits behind the C++ class surface of Grid/RootGrid/Field/Spacecraft. Each
 * entry point below replaces one of those methods and says which (file:line in
 * the reference). A C++ host facade (fopis_amd/csrc/host.cpp, class
 * fopis::GpuRootGrid, C entry points in fopis_host.h) keeps the reference's
 * method names on top of this ABI.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative FOPIS_E_* code on
 *    failure; fopis_last_error() then holds a message (thread-local). Nothing
 *    throws across the ABI.
 *  - Plain pointers and sizes only; all host buffers stay owned by the caller
 *    and are copied. The library owns every device buffer.
 *  - Arrays of grid quantities use the reference's layout: C order, glue layer
 *    included, node arrays (nx+2)*(ny+2)*(nz+2) doubles (include/field.hpp:5-36,
 *    src/Grid/Grid.cpp:148-196).
 *  - All quantities are normalised as in the reference (dx = dt = 1, masses in
 *    electron masses, charges in elementary charges; include/normalizer.hpp).
 *  - One host thread drives one context; work is ordered on the context's
 *    HIP stream; calls that return data synchronise that stream.
 */
#ifndef FOPIS_H
#define FOPIS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FOPIS_MAX_SPECIES 8
#define FOPIS_MAX_OBJECTS 4

enum fopis_status {
  FOPIS_OK = 0,
  FOPIS_E_ARG = -1,      /* bad argument (std::invalid_argument in the reference) */
  FOPIS_E_STATE = -2,    /* call out of order (std::logic_error)                  */
  FOPIS_E_DEVICE = -3,   /* HIP runtime error / no device / library not built     */
  FOPIS_E_NOMEM = -4,    /* device allocation failed                              */
  FOPIS_E_NOCONV = -5    /* Poisson solve hit its cycle limit before the tolerance */
};

enum fopis_precision { FOPIS_FP64 = 0, FOPIS_FP32 = 1 };

/* Field selectors for fopis_get_field / fopis_set_field (include/field.hpp:57-99). */
enum fopis_field {
  FOPIS_PHI = 0,       /* (nx+2)^3                                   */
  FOPIS_RHO = 1,       /* (nx+2)^3; index 0 = total, s+1 = species s */
  FOPIS_DENSITY = 2,   /* (nx+1)^3 per species (cell-centred)        */
  FOPIS_EX = 3,        /* (nx+1)(ny+2)(nz+2)                         */
  FOPIS_EY = 4,        /* (nx+2)(ny+1)(nz+2)                         */
  FOPIS_EZ = 5,        /* (nx+2)(ny+2)(nz+1)                         */
  FOPIS_EXREF = 6,     /* (nx+2)^3                                   */
  FOPIS_EYREF = 7,
  FOPIS_EZREF = 8
};

typedef struct fopis_ctx fopis_ctx;

typedef struct fopis_species_desc {
  double charge; /* ParticleType::charge, elementary charges (particle_type.hpp:89)   */
  double mass;   /* ParticleType::mass, electron masses (particle_type.hpp:88)        */
  double size;   /* ParticleType::size, real particles per superparticle (:96)        */
} fopis_species_desc;

typedef struct fopis_desc {
  int32_t nx, ny, nz;        /* RootGrid nx/ny/nz: node counts incl. the two face nodes */
  char boundary[8];          /* "DDDDDD" (Environment::boundary); D and N supported     */
  int32_t nspecies;
  fopis_species_desc species[FOPIS_MAX_SPECIES];
  double static_b[3];        /* normalised static B (Initialize.cpp:306-327)            */
  double eps0;               /* Normalizer::eps0                                        */
  int32_t precision;         /* particles, push and deposit: FOPIS_FP64 | FOPIS_FP32    */
  int32_t mg_precision;      /* V-cycle smoothing: FOPIS_FP64 | FOPIS_FP32              */
  double mg_omega;           /* red-black SOR relaxation in the V-cycle (0 -> 1.25)     */
  int32_t mg_pre, mg_post;   /* smoothing sweeps (0 -> 2, 2)                            */
  int32_t mg_max_cycles;     /* defect-correction cycle limit (0 -> 100)                */
  double poisson_tol;        /* ||b - A phi||_2 / ||b||_2 target (0 -> 1e-8)            */
  double capacity_slack;     /* per-tile capacity factor (0 -> 1.25)                    */
} fopis_desc;

/* ---- lifecycle ------------------------------------------------------------ */

/* RootGrid::RootGrid field allocation (RootGrid.cpp:7-44, Grid.cpp:148-196). */
int fopis_create(const fopis_desc* desc, fopis_ctx** out);
int fopis_destroy(fopis_ctx* ctx);
const char* fopis_last_error(void);
/* Library version string and the device it runs on. */
int fopis_device_info(char* buf, int buflen);

/* ---- particles -------------------------------------------------------------- */

/* Replaces the initial load's push_back into Grid::particles[s]
 * (RootGrid.cpp:57-79): n particles of species s, FP64 SoA, replace the
 * species' current set, including particles staged for the next deposit.
 * Particles are tile-binned on the device. */
int fopis_set_particles(fopis_ctx* ctx, int s, int64_t n, const double* x, const double* y, const double* z,
                        const double* vx, const double* vy, const double* vz);
/* Draws n particles of species s on the device and replaces the species'
 * set: positions uniform in [lo, hi) per axis, velocities drift + sigma *
 * N(0,1), from a counter-based generator (Philox4x32-10, Box-Muller) keyed by
 * seed; particles born inside a conductor are dropped, as
 * Spacecraft::removeInnerParticle does. A statistically equivalent stand-in
 * for the reference's mt19937 initial load (RootGrid.cpp:57-79,
 * AmbientParticleType.cpp:76-93) at sizes where drawing on the host is slow. */
int fopis_load_maxwellian(fopis_ctx* ctx, int s, int64_t n, uint64_t seed, const double* lo, const double* hi,
                          const double* drift, double sigma);
/* RootGrid::injectParticlesFromBoundary (RootGridParticle.cpp:3-130) with the
 * particles drawn on the device: count6[face] particles (faces x-, x+, y-, y+,
 * z-, z+; the caller keeps the reference's floor(dt*A*flux + residue) counts),
 * normal velocity N(drift, vth) conditioned on pointing inward (the law of
 * AmbientParticleType::generateNewInjectionVelocity, AmbientParticleType.cpp:
 * 101-157, drawn exactly by its inverse CDF instead of by redrawing),
 * tangential N(drift, sigma), position uniform over the face and within v of
 * it. Counter-based (Philox4x32-10) stream keyed by seed; the particles join
 * at the next deposit like fopis_add_particles. Statistically equivalent to
 * the reference's mt19937 streams, not bit-identical. */
int fopis_inject_maxwellian(fopis_ctx* ctx, int s, const int64_t* count6, uint64_t seed, const double* drift,
                            double vth, double sigma);
/* Photoemission drawn on the device: replaces Spacecraft::emitParticles'
 * photoelectron branch (Spacecraft.cpp:829-900, called from
 * RootGrid::emitParticlesFromObjects, RootGridObject.cpp:175-181). Every cmat
 * node of object obj on the face lit by shine (an axis unit vector; the
 * node rule of Spacecraft::is{X,Y,Z}surfacePoint, Spacecraft.cpp:461-504)
 * emits ceil(amount) superparticles of species s (amount = dx^2 j / (|q| size),
 * the reference's per-node count): speed vth, polar angle asin(2u - 1) to the
 * outward normal (PhotoElectronParticleType.cpp:9-106), born inside the
 * object within one step of the face. Each one's charge leaves the surface
 * at its face crossing point, bilinearly over that face square's cmat nodes
 * (Spacecraft.cpp:945-970), and counts in the object's current for the next
 * step. The particles join the next deposit moved by v (the reference emits
 * between the velocity and the position update); those that end outside
 * the domain are dropped when binned. Philox stream keyed by seed:
 * statistically equivalent to the reference's mt19937 streams. */
int fopis_emit_photoelectrons(fopis_ctx* ctx, int obj, int s, const double* shine, double amount, double vth,
                              uint64_t seed);
/* Number of live particles of species s (Grid::getValidParticleNumber,
 * GridParticleTreatment.cpp:120-128). */
int fopis_num_particles(fopis_ctx* ctx, int s, int64_t* n);
/* Copies the live particles of species s out in device order (tile-major),
 * followed by the particles staged for the next deposit (added from the
 * host, then drawn on the device, those inside the domain). Particles a
 * standalone position update absorbed are not copied. cap = buffer length;
 * the count is that of fopis_num_particles. */
int fopis_get_particles(fopis_ctx* ctx, int s, int64_t cap, double* x, double* y, double* z, double* vx,
                        double* vy, double* vz, int64_t* n);
/* Stages new particles of species s (injected at the boundary or emitted by
 * an object) to join at the next deposit: replaces particles[pid].push_back in
 * RootGrid::injectParticlesFromBoundary (RootGridParticle.cpp:46-127) and
 * Spacecraft::emitParticles (Spacecraft.cpp:887-897). */
int fopis_add_particles(fopis_ctx* ctx, int s, int64_t n, const double* x, const double* y, const double* z,
                        const double* vx, const double* vy, const double* vz);
/* Stable sort of every species by cell inside its particle tiles (tiles of
 * 8x8x8 cells, the unit the device keeps particles binned by): the order is
 * what std::stable_sort gives on the device order with the key
 * tile * 512 + ((i % 8) * 8 + j % 8) * 8 + k % 8, where (i, j, k) = floor(x, y, z)
 * and tile = ((i / 8) * T1 + j / 8) * T2 + k / 8 (T = tiles per axis). Cells
 * of one tile are contiguous and in the reference's (i, j, k) order
 * (position.hpp:59-69); on a grid of one tile this is the reference cell-key
 * order. A particle at the upper face (x = nx - 1) counts in the last cell. */
int fopis_sort_by_cell(fopis_ctx* ctx);
/* Reference cell keys of species s in device order. */
int fopis_get_cell_keys(fopis_ctx* ctx, int s, int64_t cap, int32_t* keys, int64_t* n);

/* ---- the step, method by method -------------------------------------------- */

/* Grid::updateParticleVelocityES (GridParticleTreatmentES.cpp:13-88). */
int fopis_update_particle_velocity_es(fopis_ctx* ctx);
/* RootGrid::updateParticlePositionES (GridParticleTreatmentES.cpp:90-181,
 * grid.hpp:309-395): x += v; a particle leaving a physical face is invalid. */
int fopis_update_particle_position_es(fopis_ctx* ctx);
/* RootGrid::updateDensity (RootGridField.cpp:1129-1150). */
int fopis_update_density(fopis_ctx* ctx);
/* RootGrid::updateRho (RootGridField.cpp:1153-1252): conductor absorption,
 * CIC deposit, rho[0] = sum, and with objects the first solve plus the
 * capacitance redistribution (Spacecraft.cpp:708-758). Also drops invalid
 * particles and re-bins the rest (GridParticleTreatment.cpp:93-102). */
int fopis_update_rho(fopis_ctx* ctx);
/* RootGrid::solvePoisson (RootGridField.cpp:41-103): FP64 defect correction
 * around V-cycles to poisson_tol, warm-started from the current phi. Either
 * out pointer may be NULL. */
int fopis_solve_poisson(fopis_ctx* ctx, int32_t* cycles, double* relres);
/* RootGrid::updateEfield + updateReferenceEfield (RootGridField.cpp:456-582). */
int fopis_update_efield(fopis_ctx* ctx);
/* RootGrid::mainLoopES (RootGrid.cpp:82-143) on the device, fused: one
 * kernel per species pushes, moves, absorbs, counts, deposits and re-bins;
 * then rho reduction, conductor correction, Poisson and E. Particles staged
 * by fopis_add_particles join this step's deposit. */
int fopis_main_loop_es(fopis_ctx* ctx);
/* Grid::initialize (Grid.cpp:23-36): updateRho, solvePoisson, updateEfield. */
int fopis_initialize(fopis_ctx* ctx);

/* ---- fields ---------------------------------------------------------------- */

int fopis_get_field(fopis_ctx* ctx, int which, int s, double* out);
int fopis_set_field(fopis_ctx* ctx, int which, int s, const double* in);
/* Grid::getParticleEnergy (GridParticleTreatment.cpp:104-118) and
 * Field::getEfieldEnergy (Field.cpp:194-213), FP64 reductions on the device.
 * out[0] = particle energy, out[1] = field energy. */
int fopis_energy(fopis_ctx* ctx, double* out);
/* fopis_energy without any device allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * partial sums live in a fixed array of the library, one per device, of
 * (FOPIS_MAX_SPECIES + 1) * 1024 + FOPIS_MAX_SPECIES + 1 doubles (74 KB),
 * held until the process ends; a context with more species than that is
 * refused. The result equals fopis_energy's bit for bit (same kernels, same
 * summation order). Every call on a device shares that device's array:
 * calls that overlap on one device race. */
int fopis_energy_noalloc(fopis_ctx* ctx, double* out);

/* ---- conductors (Spacecraft) ------------------------------------------------- */

/* Spacecraft::construct for a perfect conductor (Spacecraft.cpp:16-152):
 * cell_mask is the reference object_cell_map, (nx+1)(ny+1)(nz+1) bytes over
 * glue-relative cells; cmat_ijk holds the glue-relative node of each of the
 * ncmat capacitance-matrix nodes. fixed != 0 holds the potential at
 * fixed_potential (normalised). */
int fopis_set_object(fopis_ctx* ctx, int obj, const uint8_t* cell_mask, int32_t ncmat, const int32_t* cmat_ijk,
                     int32_t fixed, double fixed_potential);
/* RootGrid::initializeObjectsCmatrix (RootGridObject.cpp:77-148): one unit
 * charge solve per surface node on the device, then C = B^-1 in FP64. */
int fopis_build_cmatrix(fopis_ctx* ctx, int obj);
/* Read or load C (element (col,row) at col*ncmat+row, Cmatrix.hpp:17-20);
 * loading replaces IO::loadCmatrixData (IO.cpp:59-120). */
int fopis_get_cmatrix(fopis_ctx* ctx, int obj, double* out);
int fopis_set_cmatrix(fopis_ctx* ctx, int obj, const double* in);
/* Adds host-side surface charge (photoemission, Spacecraft.cpp:945-970) to
 * the object's temporary charge map, S*ncmat doubles, species-major. */
int fopis_add_object_charge(fopis_ctx* ctx, int obj, const double* dq, const double* dcurrent);
/* State after the last redistribution: potential, total charge, per-species
 * current (S: the superparticle charge captured and emitted in the current
 * step, reset at the start of each fopis_main_loop_es like
 * Spacecraft::resetCurrent; emission staged before a step counts in it),
 * charge_map (S*ncmat, may be NULL). */
int fopis_get_object_state(fopis_ctx* ctx, int obj, double* potential, double* total_charge, double* current,
                           double* charge_map);
/* Superparticles of each species (out[S]) that the last deposit captured in
 * object obj without finding the face they entered by: their charge counts in
 * the current but on no surface node, as in Spacecraft::
 * distributeInnerParticleCharge, which prints "Surface cannot detect" for
 * them (Spacecraft.cpp:573-589). */
int fopis_object_unattributed(fopis_ctx* ctx, int obj, int64_t* out);
/* Sets charge_map (S*ncmat), as initializeChargeMapOffset does (Spacecraft.cpp:160-183). */
int fopis_set_object_charge_map(fopis_ctx* ctx, int obj, const double* charge_map);

/* ---- measurement ------------------------------------------------------------- */

/* Average device time (ms) of the named phase over the calls since the last
 * reset, measured with HIP events on the context's stream. Phases:
 * "particles" (inbox upload, fused tile kernel, arrivals, capture and the
 * window reduction of one deposit), inside it "tile" (the tile kernels of all
 * species), "bin" (arrivals) and "reduce" (windows into rho and density);
 * "poisson" (each solve), "poisson_first" (the learned guess, when on, plus
 * the first solve of a step with objects), "guess" (network inference),
 * "efield", "velocity" and "position" (the standalone updates).
 * n = number of samples. */
int fopis_phase_time(fopis_ctx* ctx, const char* phase, double* avg_ms, int64_t* n);
int fopis_reset_timers(fopis_ctx* ctx);
/* Poisson statistics: cycles and relres of the last solve, totals since the
 * last fopis_reset_timers (the reference prints the SOR count per solve,
 * RootGridField.cpp:86-88). Any pointer may be NULL. */
int fopis_solver_stats(fopis_ctx* ctx, int32_t* last_cycles, double* last_relres, int64_t* total_cycles,
                       int64_t* total_solves);
/* Relative residuals ||b - A phi||_2 / ||b||_2 the solves started from:
 * out[0] the last solve's, out[1] the mean over the first solve of each
 * main_loop_es step since fopis_reset_timers, out[2] the mean over the
 * steps' second solves (after the conductor charge; 0 without conductors).
 * The warm-start quality behind the cycle counts; the reference's SOR starts
 * from the previous potential and reports no such figure
 * (RootGridField.cpp:41-103). */
int fopis_solver_start_residuals(fopis_ctx* ctx, double* out);
/* Phase-space moments of species s in FP64 over its live particles:
 * out[0] count, out[1..3] mean x, y, z, out[4..6] mean vx, vy, vz,
 * out[7..9] variance of vx, vy, vz (the particle side of the FP32-vs-FP64
 * comparison; the reference reports energies only, Grid::getParticleEnergy,
 * GridParticleTreatment.cpp:104-118). */
int fopis_moments(fopis_ctx* ctx, int s, double* out);
/* Multigrid cycles summed over the first Poisson solve of each main_loop_es
 * step since fopis_reset_timers (the solve the learned guess starts; phase
 * "poisson_first" of fopis_phase_time times guess + that solve). */
int fopis_first_solve_cycles(fopis_ctx* ctx, int64_t* total_cycles);
/* Full-grid passes of the Poisson solves since fopis_reset_timers beyond the
 * V-cycles themselves: out[0] FP64 defect updates (phi += e, r -= A e; one
 * per cycle, or one per pair of cycles that share a right-hand side), out[1]
 * residual passes between the two cycles of a pair (the bytes a solve
 * moves, for its roofline). The reference's SOR has no such split
 * (RootGridField.cpp:65-103). */
int fopis_solver_passes(fopis_ctx* ctx, int64_t* out);
/* Particle ledger of species s since the context was created: out[0]
 * staged particles (added, injected or emitted) that joined the particle
 * set inside the domain, out[1] particles absorbed at a domain face by a
 * position update (grid.hpp:309-395), out[2] particles captured by a
 * conductor (Spacecraft::distributeInnerParticleCharge, Spacecraft.cpp:
 * 512-592). Between two calls the live count (fopis_num_particles) changes
 * by exactly out[0] - out[1] - out[2]; set, load and sort calls replace the
 * species and are outside that balance. The reference keeps no such count;
 * its particle arrays are the ledger (Grid::getValidParticleNumber). */
int fopis_particle_ledger(fopis_ctx* ctx, int s, int64_t* out);
/* Blocks until the context's stream is idle. */
int fopis_synchronize(fopis_ctx* ctx);

/* ---- learned initial guess for the Poisson solve ---------------------------
 *
 * No counterpart in the reference, whose solve is SOR warm-started from the
 * previous potential (RootGridField.cpp:41-103). A 3-D convolutional
 * encoder-decoder (4 levels, 16-64 channels, 3^3 kernels, BF16 weights and
 * activations, FP32 accumulation) maps (rho[0], conductor node mask) to phi at
 * the interior nodes; the multigrid solve then iterates from it to the same
 * residual as always. The network grid has (n-2) rounded up to 8 voxels per
 * axis; voxel (i,j,k) is glue node (i+2, j+2, k+2).
 * Parameters are one float vector: per layer W[27][cin][cout] then b[cout],
 * layers e0a 8->16, e0b 16->16, e1a 16->32, e1b 32->32, e2a 32->48, e2b 48->48,
 * e3a 48->64, e3b 64->64, d2a 64->48, d2b 48->48, d1a 48->32, d1b 32->32,
 * d0a 32->16, d0b 16->16, out 16->16 (channel 0 used); tap (di,dj,dk) is
 * (di+1)*9 + (dj+1)*3 + (dk+1). */

/* Creates the network on the context's grid with a seeded He-normal start. */
int fopis_guess_init(fopis_ctx* ctx, uint64_t seed);
int fopis_guess_num_params(fopis_ctx* ctx, int64_t* n);
/* Master (FP32) parameters and the output scale k: phi = out * k * rms(rho) / eps0. */
int fopis_guess_get_params(fopis_ctx* ctx, float* params, double* out_scale);
int fopis_guess_set_params(fopis_ctx* ctx, const float* params, double out_scale);
/* 0: every solve starts from the previous phi (reference behaviour);
 * 1: the first solve of each step starts from the network's guess for rho;
 * 2: the first solve starts from the previous phi plus the network's guess
 *    for the residual charge eps0 (b - A phi) of that warm start;
 * 3: as 0, and the first solve of each step records the pair (that
 *    residual, phi_solution - phi_start) into the next dataset slot
 *    (fopis_guess_recorded counts them; the counter restarts at 0). */
int fopis_guess_set_mode(fopis_ctx* ctx, int mode);
/* Pairs recorded since mode 3 was set. */
int fopis_guess_recorded(fopis_ctx* ctx, int32_t* n);
/* Training objective: 0 = 0.5*mean((out - phi_hat/k)^2) (default);
 * 1 = 0.5*mean((A(k*out) - rho_hat)^2), A the 7-point operator: the relative
 * residual the multigrid starts from, which is what sets its cycle count. */
int fopis_guess_set_loss(fopis_ctx* ctx, int kind);
/* phi at the interior nodes <- network(rho[0], mask). */
int fopis_guess_apply(fopis_ctx* ctx);
/* Device dataset of training pairs: rho_hat = rho[0]/rms(rho[0]) (BF16), node
 * mask (u8), phi_hat = phi*eps0/rms(rho[0]) (BF16), each network-grid sized. */
int fopis_guess_dataset_alloc(fopis_ctx* ctx, int32_t capacity);
/* Stores the context's current (rho[0], mask, phi) as pair `slot`. */
int fopis_guess_dataset_store(fopis_ctx* ctx, int32_t slot, double* phi_hat_rms);
int fopis_guess_dataset_get(fopis_ctx* ctx, int32_t slot, uint16_t* rho_hat, uint8_t* mask, uint16_t* phi_hat);
int fopis_guess_dataset_set(fopis_ctx* ctx, int32_t slot, const uint16_t* rho_hat, const uint8_t* mask,
                            const uint16_t* phi_hat);
/* Loss (fopis_guess_set_loss) over interior voxels and its parameter
 * gradient on pair `slot` (grad may be NULL); no update. */
int fopis_guess_grad(fopis_ctx* ctx, int32_t slot, double* loss, float* grad);
/* Adam over pairs [0, npairs), `epochs` shuffled passes, learning rate on a
 * cosine from lr to lr_end over the call; epoch_loss[e] = mean loss of pass e
 * (may be NULL). Adam's moments persist across calls. */
int fopis_guess_train(fopis_ctx* ctx, int32_t npairs, int32_t epochs, double lr, double lr_end, uint64_t seed,
                      double* epoch_loss);
int fopis_guess_eval(fopis_ctx* ctx, int32_t begin, int32_t end, double* loss);
/* Training-data conductors: nbox axis-aligned boxes of glue cells
 * lohi[6*b..6*b+5] = lo0, lo1, lo2, hi0, hi1, hi2 (half-open). rho[0] is cleared
 * inside each box and its surface nodes get charge[b] times a smooth seeded
 * modulation; the boxes become the mask the guess sees. nbox = 0 returns to
 * the context's own conductor mask. */
int fopis_guess_synth_boxes(fopis_ctx* ctx, int32_t nbox, const int32_t* lohi, const double* charge, uint64_t seed);

/* One convolution of the network on host buffers (BF16 as uint16 bits):
 * x [n0'][n1'][n2'][cin] (n/2 when up != 0: nearest upsampling on read),
 * w [27][cin][cout] FP32 (rounded to BF16 on the device), b [cout],
 * add [n0][n1][n2][cout] or NULL, y [n0][n1][n2][cout] = act(conv + b + add).
 * reps > 0 also times reps launches into *ms (per launch). */
int fopis_nn_conv3d(int32_t n0, int32_t n1, int32_t n2, int32_t cin, int32_t cout, int32_t up, int32_t relu,
                    const uint16_t* x, const float* w, const float* b, const uint16_t* add, uint16_t* y, int32_t reps,
                    float* ms);
/* Gradients of that convolution for output gradient gy [n0][n1][n2][cout]:
 * dw [27][cin][cout], db [cout] (FP32) and, if dx != NULL (cin multiple of 16),
 * dx [n0][n1][n2][cin] at the output grid (sum it over 2x2x2 for up != 0). */
int fopis_nn_conv3d_backward(int32_t n0, int32_t n1, int32_t n2, int32_t cin, int32_t cout, int32_t up,
                             const uint16_t* x, const float* w, const uint16_t* gy, float* dw, float* db,
                             uint16_t* dx);

#ifdef __cplusplus
}
#endif

#endif /* FOPIS_H */
