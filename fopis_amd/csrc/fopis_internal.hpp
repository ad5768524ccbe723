// Internal types of libfopis_hip.so: device context, particle bins, launch
// helpers. Not part of the ABI (include/fopis.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fopis.h"

namespace fopis {

// Edge of a particle tile in cells. Particles are binned per 8^3-cell tile;
// one workgroup owns a tile per launch and keeps the tile's E nodes and its
// charge window in LDS.
constexpr int kTile = 8;
constexpr int kTileShift = 3;  // log2(kTile)
static_assert((1 << kTileShift) == kTile, "kTile is a power of two");
constexpr int kEW = kTile + 1;  // E-field window: nodes of the tile's cells
constexpr int kRW = kTile + 3;  // charge window: tile cells +-1 after a move
constexpr int kDW = kTile + 2;  // density window: cells of the tile +-1
constexpr int kBlock = 256;
constexpr int kSink = 1 << 16;  // elements of the tile kernel's store sink (a power of two)
// Per-species status words of the particle kernels: [0] strays, [1] lost
// (stray buffer full), [2] deposits outside the tile window, [3] captured by
// a conductor, [4] absorbed at a domain face, [5] staged particles that
// joined the tile regions, [6] captured past the capture buffer.
constexpr int kStat = 8;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FOPIS_HIP(call)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      throw ::fopis::Error(FOPIS_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

// Device buffer owned by the context.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
      p = nullptr;
      throw Error(FOPIS_E_NOMEM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    n = count;
  }
  void swap(DBuf& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
};

// Six SoA component arrays of one particle set, held in one allocation:
// component q starts q * size() elements after component 0, so a kernel can
// address every component of any set from one base and one stride.
template <class R>
struct Soa {
  struct Comp {
    R* p = nullptr;
    size_t n = 0;
  };
  DBuf<R> blk;
  Comp c[6];
  void alloc(size_t n) {
    if (n == c[0].n && blk.p) return;
    blk.alloc(6 * n);
    for (int q = 0; q < 6; ++q) {
      c[q].p = n ? blk.p + (size_t)q * n : nullptr;
      c[q].n = n;
    }
  }
  void release() {
    blk.release();
    for (auto& a : c) a = Comp{};
  }
  size_t size() const { return c[0].n; }
  // grow-only: keeps the buffers when they already hold m elements
  void reserve(size_t m) {
    if (size() < m) alloc(m);
  }
  void swap(Soa& o) {
    blk.swap(o.blk);
    std::swap(c, o.c);
  }
};

// Raw pointers of a SoA set, passed to kernels by value.
template <class R>
struct SoaPtr {
  R* c[6];
};

// Growable pinned host staging of particles added on the host (six FP64
// components), so that their upload is an asynchronous DMA.
struct PinSoa {
  double* c[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  size_t n = 0, cap = 0;
  PinSoa() = default;
  PinSoa(const PinSoa&) = delete;
  PinSoa& operator=(const PinSoa&) = delete;
  PinSoa(PinSoa&& o) noexcept : n(o.n), cap(o.cap) {
    for (int q = 0; q < 6; ++q) {
      c[q] = o.c[q];
      o.c[q] = nullptr;
    }
    o.n = o.cap = 0;
  }
  ~PinSoa() { release(); }
  void release() {
    for (auto& p : c)
      if (p) {
        (void)hipHostFree(p);
        p = nullptr;
      }
    n = cap = 0;
  }
  // Appends m particles; `st` is synchronised before an old buffer is freed.
  void append(const double* const* h, size_t m, hipStream_t st) {
    if (n + m > cap) {
      size_t nc = cap ? 2 * cap : ((size_t)1 << 16);
      while (nc < n + m) nc *= 2;
      if (st) FOPIS_HIP(hipStreamSynchronize(st));
      for (auto& p : c) {
        double* np = nullptr;
        if (hipHostMalloc(reinterpret_cast<void**>(&np), nc * sizeof(double), hipHostMallocDefault) != hipSuccess)
          throw Error(FOPIS_E_NOMEM, "pinned host allocation failed");
        if (n) std::memcpy(np, p, n * sizeof(double));
        if (p) (void)hipHostFree(p);
        p = np;
      }
      cap = nc;
    }
    for (int q = 0; q < 6; ++q) std::memcpy(c[q] + n, h[q], m * sizeof(double));
    n += m;
  }
};

// Device bookkeeping of one species.
template <class R>
struct SpeciesBins {
  Soa<R> reg;            // tile regions
  Soa<R> reg2;           // second buffer of the regions (the binning pass writes it, then they swap)
  DBuf<int64_t> start;   // [ntiles] region start
  DBuf<int32_t> capa;    // [ntiles] region capacity
  DBuf<int32_t> cnt;     // [ntiles] live particles in the region
  Soa<R> out;            // per-tile outboxes of particles leaving the tile
  DBuf<uint8_t> odir;    // direction (0..26) of each outbox entry
  DBuf<int64_t> ostart;  // [ntiles]
  DBuf<int32_t> ocapa;   // [ntiles]
  DBuf<int32_t> ocnt;    // [ntiles]
  DBuf<int32_t> arr;     // [ntiles] arrivals reserved in each region during the arrivals pass (0 between steps)
  Soa<double> inbox;     // staged new particles (FP64 from the host), sorted by tile (grow-only)
  Soa<double> inbox_raw; // their upload in arrival order (grow-only)
  DBuf<int32_t> ibeg;    // [ntiles+1]
  int64_t inbox_n = 0;
  PinSoa host_inbox;     // staged on the host (pinned) until the next deposit
  bool upload_pending = false;  // an upload from host_inbox may still be queued on the stream
  Soa<double> dev_inbox; // particles drawn on the device (injection) until the next deposit
  int64_t dev_n = 0;
  Soa<R> stray;          // particles that did not fit a bin this step
  int64_t stray_cap = 0;
  int64_t total_cap = 0;
  bool has_invalid = false;  // a standalone position update left absorbed particles (x = NaN) in their slots
  // particle ledger since the context was made: staged particles that joined,
  // particles absorbed at a domain face, particles captured by a conductor
  int64_t joined = 0, absorbed = 0, captured = 0;
};

struct ObjectDev {
  bool defined = false;
  int ncmat = 0;
  bool fixed = false;
  double fixed_potential = 0;
  DBuf<int32_t> node_cmat;   // [(nx+2)^3] glue node -> cmat index or -1
  DBuf<int64_t> cmat_node;   // [ncmat] glue linear node index
  DBuf<double> C;            // [ncmat^2], element (col,row) at col*n+row
  DBuf<double> colsum;       // [ncmat] sum_i C(i,j)
  double total_cmat = 0;
  DBuf<double> tmp_charge;   // [S][ncmat] temporary_charge_map
  DBuf<double> charge_map;   // [S][ncmat]
  DBuf<unsigned long long> absorbed;  // [S] absorbed superparticles this step
  DBuf<double> scal;         // [4]: potential, total_charge, ctp, scratch
  std::vector<double> current;         // [S] host copy (Spacecraft::current)
  std::vector<double> host_current_add;      // emission current counted in the current step
  std::vector<double> host_current_pending;  // emission staged for the next step
  double potential = 0, total_charge = 0;
  std::vector<int32_t> cmat_ijk;       // host copy
  std::vector<uint8_t> cells;          // host copy of this object's cell map, (nx+1)^3
  int lit_key = -1;                    // (axis, sign) the lit-node list below was made for
  int n_lit = 0;
  DBuf<int32_t> lit_ijk;               // [n_lit][3] glue ijk of the cmat nodes on the lit face
};

// Per-phase event timer.
struct PhaseTimer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double total_ms = 0;
  int64_t n = 0;
};

struct MgLevel {
  int n[3];            // interior unknowns per axis
  int64_t sx, sy;      // strides (elements) of the level array
  int64_t off = 0;     // element offset of the level's (0,0,0) node in its arrays
  size_t size;         // elements per level array
  DBuf<double> e64, r64, t64;  // FP64 level arrays (used when the V-cycle is FP64)
  DBuf<float> e32, r32, t32;   // FP32 level arrays
};

struct GuessNet;  // learned initial guess (guess.hip)

struct Ctx {
  double* noalloc = nullptr;  // this device's copy of the allocation-free entry points' scratch array
  fopis_desc d{};
  // the particle kernels' sort and binning scratch (owned here, so it goes with the context)
  std::shared_ptr<void> particle_scratch;
  DBuf<double> species_sizes;  // per-species superparticle sizes for the window reduction
  int S = 0;
  int nx = 0, ny = 0, nz = 0;  // node counts (RootGrid nx)
  int cx = 0, cy = 0, cz = 0;  // with glue
  int Nc[3] = {0, 0, 0};       // cells per axis = n - 1
  int nt[3] = {0, 0, 0};       // tiles per axis
  int ntiles = 0;
  size_t nnode = 0;            // cx*cy*cz
  size_t ncell_d = 0;          // (cx-1)^3 density cells
  bool fp32 = false;
  hipStream_t stream = nullptr;
  std::string devname;
  int xcds = 8;

  // fields (FP64, reference glue layout)
  DBuf<double> phi, rho, density, eref;   // rho: (S+1)*nnode, density: S*ncell_d, eref: 3*nnode
  DBuf<double> rho_part;                  // [ntiles][kRW^3] per species launch
  // per species, in the layouts of rho[s + 1] and density[s]: the deposits and
  // counts the tile kernel adds with global atomics (a particle outside its
  // tile's windows); all zero between steps (the window reduction that reads
  // them clears them)
  DBuf<double> rho_far, dens_far;
  DBuf<int32_t> dens_part;                // [ntiles][kDW^3]
  DBuf<double> red;                       // reduction scratch
  DBuf<int32_t> status;                   // [8] stray counts etc.
  DBuf<uint8_t> mask;                     // (nx+1)^3 conductor cell ids (0 none, o+1)
  DBuf<uint8_t> tile_near;                // [ntiles] 1 if a conductor cell lies within one tile of the tile
  DBuf<double> cap_rec;                   // captured particles, [S][cap_cap][8]
  DBuf<double> sink;                      // write target of tile-kernel lanes with no particle to store
  int64_t cap_cap = 0;
  bool have_mask = false;

  std::vector<SpeciesBins<double>> sp64;
  std::vector<SpeciesBins<float>> sp32;

  ObjectDev obj[FOPIS_MAX_OBJECTS];
  int nobj = 0;

  std::vector<MgLevel> mg;
  int mg_cycles_last = 0;
  int64_t mg_cycles_total = 0, mg_solves_total = 0;
  // FP64 defect updates and pair residual passes (two cycles share an update)
  int mg_updates_last = 0, mg_pair_passes_last = 0;
  int64_t mg_updates_total = 0, mg_pair_passes_total = 0;
  int64_t mg_first_cycles_total = 0;  // cycles of the first solve of each step
  double relres_last = 0;
  double relres0_last = 0;     // relative residual the last solve started from
  double relres0_first = 0, relres0_other = 0;  // sums over the first / the other solves of steps
  int64_t n_first = 0, n_other = 0;
  DBuf<double> psi;            // conductor response of the last step: phi after minus phi before the charge
  int psi_step = -1;           // step that formed psi (-1: none)
  bool carry = true;           // warm starts from psi (FOPIS_CARRY=0: the previous phi for both solves)
  DBuf<int32_t> mg_state;     // [4] solve flag and cycle count on the device
  DBuf<double> mg_part;       // per-block residual partials and the norms of the solve (fixed address)
  // captured once: [0] one V-cycle + FP64 update + norm check, [1] two
  // V-cycles (the second continuing from the first's correction) + update + check
  hipGraphExec_t mg_graph[2] = {nullptr, nullptr};
  size_t mg_graph_esz = 0;    // element size the graphs were captured for
  double mg_graph_tol = 0.0;  // tolerance and cycle cap captured in their norm kernels
  int mg_graph_cap = 0;
  bool mg_pairs = true;  // a solve's predicted cycles run two per FP64 update
  // FP32 finest level stored by colour (red-black halves, see poisson.hip):
  // e and r of colour k at erb/rrb + k * rb_cs, point (a, b, c) at
  // a * rb_sx + b * rb_sy + c / 2
  bool mg_rb = false;
  DBuf<float> erb, rrb;
  int64_t rb_cs = 0, rb_sx = 0, rb_sy = 0;
  bool mg_use_graph = true;   // replay the captured cycle (FOPIS_MG_GRAPH=0 launches it kernel by kernel)
  DBuf<double> mg_r64;        // FP64 level-0 residual when the V-cycle is FP32
  int mg_hist[2] = {1, 1};    // cycles of the last two solves (batch size of the next)
  int mg_tail = 0;            // first level solved by the single-workgroup LDS kernel
  size_t mg_tail_bytes = 0;   // its LDS bytes
  size_t mg_tail_esz = 0;     // element size the tail was planned for
  bool mg_tail15 = false;     // the tail is the cubic 15 -> 7 -> 3 -> 1 chain
  size_t lds_max = 65536;     // LDS bytes per workgroup on the device
  GuessNet* guess = nullptr;  // owned; released by guess_free

  std::map<std::string, PhaseTimer> timers;
  bool timing = true;
  int timestep = 1;
  bool tile_double_buffer = true;  // FOPIS_TILE_INPLACE=1 compacts the regions in place instead

  size_t id3(int i, int j, int k) const { return ((size_t)i * cy + j) * cz + k; }
};

// ---- entry points implemented in the .hip units ----
void particles_init(Ctx& c);
void particles_set(Ctx& c, int s, int64_t n, const double* const* host6);
void particles_load_maxwellian(Ctx& c, int s, int64_t n, uint64_t seed, const double* lo, const double* hi,
                               const double* drift, double sigma);
int64_t particles_count(Ctx& c, int s);
void particles_get(Ctx& c, int s, int64_t cap, double* const* host6, int64_t* n);
void particles_stage(Ctx& c, int s, int64_t n, const double* const* host6);
void particles_inject(Ctx& c, int s, const int64_t* count, uint64_t seed, const double* drift, double vth,
                      double sigma);
void particles_velocity(Ctx& c);
void particles_position(Ctx& c);
void particles_density_only(Ctx& c);
void particles_deposit_bin(Ctx& c, bool push);
void particles_sort_by_cell(Ctx& c);
void particles_cell_keys(Ctx& c, int s, int64_t cap, int32_t* keys, int64_t* n);
double particles_energy(Ctx& c);
// Energy reductions: kEnergyBlocks partials per quantity, summed in a fixed order.
constexpr int kEnergyBlocks = 1024;
// Queues sum v^2 of every species into sums[s] (part: S * kEnergyBlocks doubles).
void particles_energy_queue(Ctx& c, double* part, double* sums);
// 1/2 sum_s m_s size_s v2[s].
double particles_energy_from(const Ctx& c, const double* v2);
// Device address of this device's scratch array (set up once per device, with it current).
void noalloc_setup(Ctx& c);
// fopis_energy computed with the partials in that array (allocates nothing).
void energy_noalloc(Ctx& c, double* out);
void particles_moments(Ctx& c, int s, double* out);
void particles_ledger(Ctx& c, int s, int64_t* out);

void fields_init(Ctx& c);
void fields_efield(Ctx& c);
// far_mask: bit s set when species s may have added to rho_far / dens_far
void fields_reduce_windows(Ctx& c, bool do_rho, bool do_dens, uint32_t far_mask = ~0u);
void fields_get(Ctx& c, int which, int s, double* out);
void fields_set(Ctx& c, int which, int s, const double* in);
double fields_energy(Ctx& c);
// Queues the sum of squared edge fields into *sum (part: kEnergyBlocks doubles).
void fields_energy_queue(Ctx& c, double* part, double* sum);

void poisson_init(Ctx& c);
void poisson_free(Ctx& c);
int poisson_solve(Ctx& c, double tol, int max_cycles, double* relres, const double* rhs_override);

void objects_set(Ctx& c, int o, const uint8_t* mask, int ncmat, const int32_t* ijk, int fixed, double fixed_pot);
void objects_build_cmatrix(Ctx& c, int o);
void objects_set_cmatrix(Ctx& c, int o, const double* C);
void objects_get_cmatrix(Ctx& c, int o, double* C);
void objects_sum_whole_charge(Ctx& c);
void objects_redistribute(Ctx& c);
// Superparticles of each species captured in the last deposit whose entry face was not found.
void objects_unattributed(Ctx& c, int o, int64_t* out);
// Warm starts of the two solves of a step with conductors from the last
// step's conductor response psi (phi after the charge redistribution minus
// phi before it): first solve from phi - psi, second from phi + psi.
void objects_carry_first(Ctx& c);
void objects_carry_second(Ctx& c);
void objects_carry_end(Ctx& c);
void objects_add_charge(Ctx& c, int o, const double* dq, const double* dcur);
void objects_state(Ctx& c, int o, double* pot, double* q, double* cur, double* cmap);
void objects_set_charge_map(Ctx& c, int o, const double* cmap);
int objects_lit_nodes(Ctx& c, int o, int axis, int sign, const int32_t** dev_ijk);
void objects_reset_step(Ctx& c);
void particles_emit(Ctx& c, int o, int s, int axis, int sign, int per_node, double vth, uint64_t seed);

void guess_init(Ctx& c, uint64_t seed);
void guess_free(Ctx& c);
int64_t guess_num_params(Ctx& c);
void guess_get_params(Ctx& c, float* out, double* k_out);
void guess_set_params(Ctx& c, const float* in, double k_out);
void guess_set_mode(Ctx& c, int mode);
void guess_set_loss(Ctx& c, int kind);
int guess_mode(const Ctx& c);
void guess_apply(Ctx& c);
// Modes 2 (network correction of the warm start) and 3 (record warm-start pairs).
void guess_before_solve(Ctx& c);
void guess_after_solve(Ctx& c);
int guess_recorded(const Ctx& c);
void guess_dataset_alloc(Ctx& c, int capacity);
double guess_dataset_store(Ctx& c, int slot);
void guess_dataset_get(Ctx& c, int slot, uint16_t* rho_hat, uint8_t* mask, uint16_t* phi_hat);
void guess_dataset_set(Ctx& c, int slot, const uint16_t* rho_hat, const uint8_t* mask, const uint16_t* phi_hat);
double guess_grad(Ctx& c, int slot, float* grad_out);
void guess_train(Ctx& c, int nsamples, int epochs, double lr, double lr_end, uint64_t seed, double* loss_out);
double guess_eval(Ctx& c, int begin, int end);
void guess_synth_boxes(Ctx& c, int nbox, const int32_t* lohi, const double* q, uint64_t seed);
void guess_clear_boxes(Ctx& c);
void nn_conv3d(int n0, int n1, int n2, int cin, int cout, int up, int relu, const uint16_t* x, const float* w,
               const float* b, const uint16_t* add, uint16_t* y, int reps, float* ms);
void nn_conv3d_backward(int n0, int n1, int n2, int cin, int cout, int up, const uint16_t* x, const float* w,
                        const uint16_t* gy, float* dw_out, float* db_out, uint16_t* dx_out);

void timer_begin(Ctx& c, const char* name, hipEvent_t* ev);
void timer_end(Ctx& c, const char* name, hipEvent_t ev);
void timers_collect(Ctx& c);

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Scoped phase timing on the context stream.
struct Phase {
  Ctx& c;
  const char* name;
  hipEvent_t ev = nullptr;
  Phase(Ctx& c_, const char* n) : c(c_), name(n) { timer_begin(c, name, &ev); }
  ~Phase() { timer_end(c, name, ev); }
};

}  // namespace fopis
