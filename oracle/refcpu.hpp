// ORACLE — test infrastructure only.
//
// CPU restatement of hsimyu/fopis's electrostatic PIC step (RootGrid::mainLoopES
// and everything it calls), written from the reference's algorithm, not copied.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it, and only to check results against or as the reported CPU baseline: the product
// (fopis_amd/libfopis_hip.so) never links, loads or calls anything here.
//
// Every function names the reference file:line it restates. Arithmetic order
// follows the reference operation by operation; the build uses
// -ffp-contract=off so no fused multiply-add changes a rounding.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <random>
#include <string>
#include <vector>

namespace refcpu {

// include/global.hpp:27-31
constexpr double kE = 1.60217733e-19;
constexpr double kEps0 = 8.854187817e-12;
constexpr double kMe = 9.1093818872e-31;

// include/normalizer.hpp:24-36 (units x=dx, t=dt, m=me, e=e; Initialize.cpp:33-36)
struct Units {
  double x = 1, t = 1, m = kMe, e = kE;
  double V = 1, A = 1, x3 = 1, eps0 = 1;
  void set(double dx, double dt);
  double normVelocity(double v) const { return t * v / x; }
  double normBfield(double b) const { return b * (t * e) / m; }
  double normDensity(double n) const { return n * x3; }
  double normCurrentDensity(double j) const { return j * (x * x) / A; }
  double normPotential(double p) const { return p / V; }
};

// include/position.hpp:11-99
struct Position {
  int i = 0, j = 0, k = 0;
  double x = 0, y = 0, z = 0;
  double dx1 = 0, dy1 = 0, dz1 = 0, dx2 = 0, dy2 = 0, dz2 = 0;
  Position() = default;
  Position(double x_, double y_, double z_) { setXYZ(x_, y_, z_); }
  static Position fromIJK(int i_, int j_, int k_);
  void setXYZ(double x_, double y_, double z_);
  void updateDelta();
};

// include/particle.hpp:11-113 — 56-byte AoS particle.
struct Particle {
  double x, y, z, vx, vy, vz;
  int typeId;
  int isValid;
};
static_assert(sizeof(Particle) == 56, "reference particle is 56 bytes");

struct SpeciesParams {
  char type[16];       // "ambient" | "photoelectron"
  double mass;         // in electron masses
  double charge;       // in elementary charges
  double density;      // m^-3
  double temperature;  // eV
  double drift_kms[3];
  int ppc;
  int injection_axis[6];
  double current_density;  // A/m^2 (photoelectron)
};

struct ObjectParams {
  char file_name[512];
  int is_potential_fixed;
  double fixed_potential;          // V
  double initial_potential_offset; // V
  int emit_species;                // photoelectron species id emitted by this object, or -1
};

struct Params {
  int nx, ny, nz;  // Environment nx/ny/nz (node counts, one rank)
  double dx, dt;   // m, s
  char boundary[8];
  double static_b_nT[3];
  double shine[3];
  int nspecies;
  SpeciesParams sp[8];
  int nobjects;
  ObjectParams obj[4];
  int post_loop;  // Options.maximum_poisson_post_loop
};

// One particle species with its RNG state (ParticleType.cpp:11-27).
struct Species {
  int id = 0;
  std::string type;
  double mass = 1, charge = -1, density = 0, temperature = 0;
  int ppc = 1, size = 1;
  double drift[3] = {0, 0, 0};
  bool inj[6] = {false, false, false, false, false, false};
  double current_density = 0;
  std::mt19937 mt_x, mt_y, mt_z, mt_vx, mt_vy, mt_vz;
  std::mt19937_64 mt_azimuth, mt_depression;  // random_distribution.hpp:10-29
  void setId(int id_);
  double chargeOfSuperParticle() const { return charge * size; }
  double thermalVelocity(const Units& u) const;
  double deviation(const Units& u) const;
};

// Output of ObjectUtils::getObjectNodesFromObjFile (ObjectUtils.cpp:9-333).
struct ObjData {
  std::map<unsigned, std::array<int, 3>> nodes;  // cmat -> absolute node ijk
  std::vector<std::array<int, 3>> cells;         // absolute cell ijk with count == 3
};

ObjData loadObjFile(const std::string& path, int nx, int ny, int nz);

struct Grid;

// Spacecraft.cpp — perfect-conductor branch with photoelectron emission.
struct Spacecraft {
  int nx = 0, ny = 0, nz = 0;
  unsigned num_cmat = 0;
  bool defined = false;
  bool is_potential_fixed = false;
  double fixed_potential = 0, initial_potential_offset = 0;
  int emit_species = -1;
  double potential = 0, total_charge = 0, total_cmat_value = 0;
  std::map<unsigned, Position> cmat_rel;  // cmat -> glue-relative node
  std::vector<uint8_t> cell_map;          // (nx+1)^3, glue-relative cells
  std::vector<double> cmat;               // num_cmat^2, element (col,row) at col*n+row
  std::vector<double> charge_map, temporary_charge_map;  // [S][ncmat]
  std::vector<double> current;                           // [S]
  // dense node -> cmat lookup; -1 where no cmat node
  std::vector<int> node_cmat;

  bool cell(int i, int j, int k) const { return cell_map[((size_t)i * (ny + 1) + j) * (nz + 1) + k] != 0; }
  bool isValidCell(int i, int j, int k) const;
  bool isContaining(const Position& p) const;
  int cmatNumber(int i, int j, int k) const;
  bool isCmatNode(int i, int j, int k) const { return cmatNumber(i, j, k) >= 0; }
  bool isXsurfaceMinus(int i, int j, int k) const { return cell(i, j, k) && !cell(i - 1, j, k); }
  bool isXsurfacePlus(int i, int j, int k) const { return !cell(i, j, k) && cell(i - 1, j, k); }
  bool isYsurfaceMinus(int i, int j, int k) const { return cell(i, j, k) && !cell(i, j - 1, k); }
  bool isYsurfacePlus(int i, int j, int k) const { return !cell(i, j, k) && cell(i, j - 1, k); }
  bool isZsurfaceMinus(int i, int j, int k) const { return cell(i, j, k) && !cell(i, j, k - 1); }
  bool isZsurfacePlus(int i, int j, int k) const { return !cell(i, j, k) && cell(i, j, k - 1); }
  bool isXsurfacePoint(const Position& p, int sign) const;
  bool isYsurfacePoint(const Position& p, int sign) const;
  bool isZsurfacePoint(const Position& p, int sign) const;
  void distributeInnerParticleCharge(Particle& p, const std::vector<Species>& sp);
  void addX(const Position& p, int id, double q, double sgn);
  void addY(const Position& p, int id, double q, double sgn);
  void addZ(const Position& p, int id, double q, double sgn);
  void sumWholeCharge(int S);
  double getTotalCharge(int S) const;
  void redistributeCharge(std::vector<double>& rho0, const std::vector<double>& phi, int cy, int cz, int S);
  void updateTotalCmatValue();
};

// RootGrid + Field for one rank (RootGrid.cpp, RootGridField.cpp, Field.cpp).
struct Grid {
  Params prm{};
  Units u;
  int nx = 0, ny = 0, nz = 0;  // RootGrid nx = Environment::cell_x
  int cx = 0, cy = 0, cz = 0;  // nx + 2
  double dx = 1, dt = 1;       // normalised (always 1)
  double b[3] = {0, 0, 0};
  double shine[3] = {0, 0, 0};
  std::string boundary;
  int S = 0;
  int timestep = 1;
  int post_loop = 1000;
  int last_sor_iterations = 0;
  std::vector<Species> sp;
  std::vector<std::vector<Particle>> particles;
  std::vector<double> phi, poisson_residual, poisson_error;
  std::vector<std::vector<double>> rho;      // S+1 arrays, cx*cy*cz
  std::vector<std::vector<double>> density;  // S arrays, (cx-1)^3
  std::vector<double> ex, ey, ez, exref, eyref, ezref;
  std::vector<Spacecraft> objects;
  std::vector<std::vector<double>> inj_residual;  // RootGridParticle.cpp:4 (function-static there)

  explicit Grid(const Params& p);
  size_t id3(int i, int j, int k) const { return ((size_t)i * cy + j) * cz + k; }
  char bc(int axis, int side) const { return boundary[2 * axis + side]; }
  bool isBoundary(int axis, int side) const { return bc(axis, side) != 'P'; }

  void loadInitialParticles();       // RootGrid.cpp:45-79
  void initialize();                 // Grid.cpp:23-36
  void mainLoopES();                 // RootGrid.cpp:82-143
  void updateParticleVelocityES();   // GridParticleTreatmentES.cpp:13-88
  void updateParticlePositionES();   // GridParticleTreatmentES.cpp:90-181
  void injectParticlesFromBoundary();  // RootGridParticle.cpp:3-130
  void emitParticlesFromObjects();     // RootGridObject.cpp:175-181
  void updateDensity();              // RootGridField.cpp:1129-1150
  void updateRho();                  // RootGridField.cpp:1153-1252
  void depositOnly();                // updateRho's containment+CIC loop without the conductor solve
  void solvePoisson();               // RootGridField.cpp:41-103
  int solvePoissonPSOR(int loopnum);
  int solvePoissonToRelres(double tol, int maxloop);
  void doPSOR();                     // RootGridField.cpp:105-214 (one rank: interior call only)
  double checkPhiResidual();         // RootGridField.cpp:267-317
  double relres() const;             // ||b - A phi||_2 / ||b||_2 over interior nodes
  void setBoundaryConditionPhi();    // Field.cpp:10-96
  void updateEfield();               // RootGridField.cpp:456-486
  void updateReferenceEfield();      // RootGridField.cpp:489-582
  void removeInvalidParticles();     // GridParticleTreatment.cpp:93-102
  double particleEnergy() const;     // GridParticleTreatment.cpp:104-118
  double efieldEnergy() const;       // Field.cpp:194-213
  void buildCmatrix();               // RootGridObject.cpp:77-148
  void initializeChargeMapOffset();  // Spacecraft.cpp:160-183
};

// flux of AmbientParticleType::calcFlux (AmbientParticleType.cpp:8-57)
std::vector<double> calcFlux(const Species& s, const Units& u);

}  // namespace refcpu
