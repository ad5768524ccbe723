// Host facade of the ES step: reference JSON input, unit normalisation,
// per-species mt19937 streams, .obj spacecraft geometry, boundary injection
// and photoemission on the host, and the device step through include/fopis.h.
// fopis::GpuRootGrid keeps the reference's RootGrid method names.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <dlfcn.h>
#include <sys/stat.h>

#include "../../include/fopis.h"
#include "../../include/fopis_host.h"
#include "../../include/fopis_io.h"

namespace fopis {

// ------------------------------------------------------------------ JSON
// Minimal reader for the reference's input files (objects, arrays, numbers,
// strings, booleans, null).
namespace json {

struct Value {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  bool has(const std::string& k) const { return kind == Obj && obj.count(k) > 0; }
  const Value& at(const std::string& k) const {
    auto it = obj.find(k);
    if (kind != Obj || it == obj.end()) throw std::invalid_argument("missing key '" + k + "' in input");
    return it->second;
  }
  double number() const {
    if (kind != Num) throw std::invalid_argument("expected a number");
    return num;
  }
  bool boolean() const {
    if (kind != Bool) throw std::invalid_argument("expected a boolean");
    return b;
  }
  const std::string& string() const {
    if (kind != Str) throw std::invalid_argument("expected a string");
    return str;
  }
};

struct Parser {
  const std::string& s;
  size_t i = 0;
  explicit Parser(const std::string& t) : s(t) {}
  [[noreturn]] void fail(const char* m) { throw std::invalid_argument(std::string("JSON: ") + m + " at offset " + std::to_string(i)); }
  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t' || s[i] == '\r')) ++i;
  }
  Value parse() {
    ws();
    if (i >= s.size()) fail("unexpected end");
    const char c = s[i];
    Value v;
    if (c == '{') {
      v.kind = Value::Obj;
      ++i;
      ws();
      if (i < s.size() && s[i] == '}') {
        ++i;
        return v;
      }
      for (;;) {
        ws();
        Value k = parse();
        if (k.kind != Value::Str) fail("object key must be a string");
        ws();
        if (i >= s.size() || s[i] != ':') fail("expected ':'");
        ++i;
        v.obj[k.str] = parse();
        ws();
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == '}') {
          ++i;
          return v;
        }
        fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = Value::Arr;
      ++i;
      ws();
      if (i < s.size() && s[i] == ']') {
        ++i;
        return v;
      }
      for (;;) {
        v.arr.push_back(parse());
        ws();
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == ']') {
          ++i;
          return v;
        }
        fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = Value::Str;
      ++i;
      while (i < s.size() && s[i] != '"') {
        if (s[i] == '\\' && i + 1 < s.size()) {
          ++i;
          const char e = s[i];
          v.str += e == 'n' ? '\n' : (e == 't' ? '\t' : e);
        } else {
          v.str += s[i];
        }
        ++i;
      }
      if (i >= s.size()) fail("unterminated string");
      ++i;
      return v;
    }
    if (s.compare(i, 4, "true") == 0) {
      v.kind = Value::Bool;
      v.b = true;
      i += 4;
      return v;
    }
    if (s.compare(i, 5, "false") == 0) {
      v.kind = Value::Bool;
      i += 5;
      return v;
    }
    if (s.compare(i, 4, "null") == 0) {
      i += 4;
      return v;
    }
    // number: strtod, as picojson does
    const char* b = s.c_str() + i;
    char* e = nullptr;
    v.num = std::strtod(b, &e);
    if (e == b) fail("bad value");
    v.kind = Value::Num;
    i += (size_t)(e - b);
    return v;
  }
};

Value parse_file(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::invalid_argument("cannot open " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string t = ss.str();
  Parser p(t);
  return p.parse();
}

}  // namespace json

// ------------------------------------------------------------- constants
// physical constants as in the reference (include/global.hpp:27-31)
constexpr double kE = 1.60217733e-19;
constexpr double kEps0 = 8.854187817e-12;
constexpr double kMe = 9.1093818872e-31;

// Normalizer units x = dx, t = dt, m = me, e = e (Initialize.cpp:33-36,
// normalizer.hpp:24-36).
struct Units {
  double x = 1, t = 1, m = kMe, e = kE, V = 1, A = 1, x3 = 1, eps0 = 1;
  void set(double dx, double dt) {
    x = dx;
    t = dt;
    x3 = std::pow(x, 3);
    V = std::pow(x, 2) * m / (e * std::pow(t, 2));
    A = e / t;
    eps0 = kEps0 * (x3 * m) / (std::pow(e, 2) * std::pow(t, 2));
  }
  double velocity(double v) const { return t * v / x; }
  double bfield(double b) const { return b * (t * e) / m; }
  double density(double n) const { return n * x3; }
  double current_density(double j) const { return j * (x * x) / A; }
  double potential(double p) const { return p / V; }
};

// One particle species with the reference's RNG streams.
struct Species {
  std::string name, type;
  int id = 0;
  double mass = 1, charge = -1, density = 0, temperature = 0;  // temperature as kB*T in J
  int ppc = 1, size = 1;
  double drift[3] = {0, 0, 0};
  bool inj[6] = {false, false, false, false, false, false};
  double current_density = 0;  // normalised
  std::mt19937 mt[6];           // x, y, z, vx, vy, vz
  std::mt19937_64 mt_azimuth, mt_depression;
  // ParticleType::generated_counts (particle_type.hpp:38-55): one count per
  // generated position (x, y, z) and velocity (vx, vy, vz) of an ambient
  // species, saved with the resume data
  uint64_t generated[6] = {0, 0, 0, 0, 0, 0};

  // ParticleType::setId (ParticleType.cpp:19-27), rank 0
  void set_id(int i) {
    id = i;
    mt[0].seed(10684930 + 106 * i);
    mt[1].seed(99881 + 171 * i);
    mt[2].seed(861200045 + 95 * i);
    mt[3].seed(930 + 2 * i);
    mt[4].seed(98076621 - 50 * i);
    mt[5].seed(7662566 + 4 * i);
  }
  double thermal_velocity(const Units& u) const { return u.velocity(std::sqrt(2.0 * temperature / (mass * kMe))); }
  double deviation(const Units& u) const { return u.velocity(std::sqrt(temperature / (mass * kMe))); }
  double super_charge() const { return charge * size; }
};

struct Particle6 {
  double v[6];
};

// ------------------------------------------------------- .obj geometry
// MagicaVoxel-style quads -> surface nodes numbered in order of first
// appearance, a face map per axis, and cells painted between pairs of faces
// along each axis; a cell is inside when all three axes paint it
// (ObjectUtils.cpp:9-261). Output is glue-relative like Spacecraft::construct.
struct ObjGeom {
  std::vector<uint8_t> mask;   // (nx+1)(ny+1)(nz+1)
  std::vector<int32_t> ijk;    // 3 per cmat node, glue-relative
};

static std::vector<std::string> split(const std::string& t, char d) {
  std::vector<std::string> r;
  std::stringstream ss(t);
  std::string b;
  while (std::getline(ss, b, d)) r.push_back(b);
  return r;
}

static ObjGeom load_obj(const std::string& path, int nx, int ny, int nz) {
  std::ifstream in(path);
  if (!in) throw std::invalid_argument("File " + path + " does not exist.");
  std::vector<std::array<int, 3>> verts;
  std::vector<std::array<int, 4>> faces;
  const std::regex rv("v (.*)"), rf("f (.*)");
  std::string line;
  while (std::getline(in, line)) {
    if (std::regex_match(line, rv)) {
      auto t = split(line, ' ');
      verts.push_back({std::stoi(t[1]) + nx / 2, std::stoi(t[2]) + ny / 2, std::stoi(t[3]) + nz / 2});
    } else if (std::regex_match(line, rf)) {
      auto t = split(line, ' ');
      std::array<int, 4> f{};
      for (int q = 0; q < 4; ++q) f[q] = std::stoi(split(t[q + 1], '/')[0]);
      faces.push_back(f);
    }
  }
  auto inb = [&](int i, int j, int k) { return i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz; };
  std::vector<int32_t> node((size_t)nx * ny * nz, -1);
  std::vector<std::array<int, 3>> order;
  // face maps: x-faces on (nx, ny-1, nz-1) etc.
  std::vector<uint8_t> fx((size_t)nx * (ny - 1) * (nz - 1)), fy((size_t)(nx - 1) * ny * (nz - 1)),
      fz((size_t)(nx - 1) * (ny - 1) * nz);
  for (const auto& f : faces) {
    std::array<int, 3> v[4];
    for (int q = 0; q < 4; ++q) {
      if (f[q] < 1 || f[q] > (int)verts.size()) throw std::invalid_argument("obj face refers to a missing vertex");
      v[q] = verts[f[q] - 1];
    }
    int axis = -1;
    for (int a = 0; a < 3 && axis < 0; ++a)
      if (v[0][a] == v[1][a] && v[0][a] == v[2][a] && v[0][a] == v[3][a]) axis = a;
    if (axis < 0) throw std::logic_error("Face type cannot be determinted by vertices position.");
    const int u = axis == 0 ? 1 : 0, w = axis == 2 ? 1 : 2;
    const int c = v[0][axis];
    const int u0 = std::min({v[0][u], v[1][u], v[2][u], v[3][u]}), u1 = std::max({v[0][u], v[1][u], v[2][u], v[3][u]});
    const int w0 = std::min({v[0][w], v[1][w], v[2][w], v[3][w]}), w1 = std::max({v[0][w], v[1][w], v[2][w], v[3][w]});
    for (int a = u0; a <= u1; ++a)
      for (int b = w0; b <= w1; ++b) {
        std::array<int, 3> p{};
        p[axis] = c;
        p[u] = a;
        p[w] = b;
        if (!inb(p[0], p[1], p[2])) throw std::invalid_argument("object node outside the grid");
        int32_t& id = node[((size_t)p[0] * ny + p[1]) * nz + p[2]];
        if (id < 0) {
          id = (int32_t)order.size();
          order.push_back(p);
        }
        if (a != u1 && b != w1) {
          if (axis == 0) fx[((size_t)p[0] * (ny - 1) + p[1]) * (nz - 1) + p[2]] = 1;
          if (axis == 1) fy[((size_t)p[0] * ny + p[1]) * (nz - 1) + p[2]] = 1;
          if (axis == 2) fz[((size_t)p[0] * (ny - 1) + p[1]) * nz + p[2]] = 1;
        }
      }
  }
  const int Cx = nx - 1, Cy = ny - 1, Cz = nz - 1;
  std::vector<uint8_t> count((size_t)Cx * Cy * Cz, 0);
  auto paint = [&](int n_along, auto face_at, auto cell_at) {
    // walk a line of faces, painting the cells between consecutive pairs
    for (int lo = 0; lo < n_along; ++lo) {
      if (!face_at(lo)) continue;
      for (int up = lo + 1; up < n_along; ++up)
        if (face_at(up)) {
          for (int q = lo; q < up; ++q) ++cell_at(q);
          lo = up;
          break;
        }
    }
  };
  for (int j = 0; j < Cy; ++j)
    for (int k = 0; k < Cz; ++k)
      paint(Cx, [&](int i) { return fx[((size_t)i * (ny - 1) + j) * (nz - 1) + k] != 0; },
            [&](int i) -> uint8_t& { return count[((size_t)i * Cy + j) * Cz + k]; });
  for (int i = 0; i < Cx; ++i)
    for (int k = 0; k < Cz; ++k)
      paint(Cy, [&](int j) { return fy[((size_t)i * ny + j) * (nz - 1) + k] != 0; },
            [&](int j) -> uint8_t& { return count[((size_t)i * Cy + j) * Cz + k]; });
  for (int i = 0; i < Cx; ++i)
    for (int j = 0; j < Cy; ++j)
      paint(Cz, [&](int k) { return fz[((size_t)i * (ny - 1) + j) * nz + k] != 0; },
            [&](int k) -> uint8_t& { return count[((size_t)i * Cy + j) * Cz + k]; });
  ObjGeom g;
  g.mask.assign((size_t)(nx + 1) * (ny + 1) * (nz + 1), 0);
  for (int i = 0; i < Cx; ++i)
    for (int j = 0; j < Cy; ++j)
      for (int k = 0; k < Cz; ++k)
        if (count[((size_t)i * Cy + j) * Cz + k] == 3)
          g.mask[((size_t)(i + 1) * (ny + 1) + (j + 1)) * (nz + 1) + (k + 1)] = 1;
  for (const auto& p : order) {
    g.ijk.push_back(p[0] + 1);
    g.ijk.push_back(p[1] + 1);
    g.ijk.push_back(p[2] + 1);
  }
  return g;
}

struct Object {
  std::string name, file;
  bool fixed = false;
  double fixed_potential = 0, offset = 0;
  std::vector<int> emit_species;
  ObjGeom geom;
  int ncmat = 0;
  // dense node lookup for emission bookkeeping
  std::vector<int32_t> node_cmat;
  bool cell(int nx, int ny, int nz, int i, int j, int k) const {
    if (i < 0 || j < 0 || k < 0 || i > nx || j > ny || k > nz) return false;
    return geom.mask[((size_t)i * (ny + 1) + j) * (nz + 1) + k] != 0;
  }
};

// One step's worth of host-drawn particles and emission charge.
struct StepBatch {
  std::vector<std::vector<Particle6>> particles;  // per species
  std::vector<std::vector<double>> dq, dcur;       // per object: S*ncmat, S
  std::vector<std::array<int64_t, 6>> inj_count;   // per species: particles the device draws per face
  uint64_t inj_seed = 0;
  struct Emit {
    int o, s;
    double amount, vth;
    uint64_t seed;
  };
  std::vector<Emit> emit;  // photoemission the device draws
};

// ------------------------------------------------------- HDF5 files
// The resume and capacity-matrix files go through libfopis_io.so (HDF5),
// loaded next to this library on first use, so that the step itself does not
// depend on HDF5.
struct IoApi {
  decltype(&fopis_io_open) open = nullptr;
  decltype(&fopis_io_close) close = nullptr;
  decltype(&fopis_io_write) write = nullptr;
  decltype(&fopis_io_exists) exists = nullptr;
  decltype(&fopis_io_info) info = nullptr;
  decltype(&fopis_io_read) read = nullptr;
  decltype(&fopis_io_last_error) last_error = nullptr;
  std::string why;
};

static void io_anchor() {}

static const IoApi& io_api() {
  static const IoApi api = [] {
    IoApi a;
    std::string dir = ".";
    Dl_info di;
    if (dladdr(reinterpret_cast<void*>(&io_anchor), &di) && di.dli_fname) {
      const std::string self = di.dli_fname;
      const size_t sl = self.rfind('/');
      if (sl != std::string::npos) dir = self.substr(0, sl);
    }
    const std::string path = dir + "/libfopis_io.so";
    void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) {
      const char* e = dlerror();
      a.why = path + ": " + (e ? e : "cannot load");
      return a;
    }
    a.open = reinterpret_cast<decltype(a.open)>(dlsym(h, "fopis_io_open"));
    a.close = reinterpret_cast<decltype(a.close)>(dlsym(h, "fopis_io_close"));
    a.write = reinterpret_cast<decltype(a.write)>(dlsym(h, "fopis_io_write"));
    a.exists = reinterpret_cast<decltype(a.exists)>(dlsym(h, "fopis_io_exists"));
    a.info = reinterpret_cast<decltype(a.info)>(dlsym(h, "fopis_io_info"));
    a.read = reinterpret_cast<decltype(a.read)>(dlsym(h, "fopis_io_read"));
    a.last_error = reinterpret_cast<decltype(a.last_error)>(dlsym(h, "fopis_io_last_error"));
    if (!a.open || !a.close || !a.write || !a.exists || !a.info || !a.read || !a.last_error) {
      a.open = nullptr;
      a.why = path + ": missing entry points";
    }
    return a;
  }();
  if (!api.open) throw std::runtime_error("HDF5 files need libfopis_io.so: " + api.why);
  return api;
}

// One HDF5 file, closed at scope exit.
class H5 {
 public:
  H5(const std::string& path, bool create) : io_(io_api()), path_(path) {
    if (io_.open(path.c_str(), create ? 1 : 0, &f_) != 0) throw std::runtime_error(io_.last_error());
  }
  ~H5() {
    if (f_) io_.close(f_);
  }
  H5(const H5&) = delete;
  H5& operator=(const H5&) = delete;
  bool exists(const std::string& name) const { return io_.exists(f_, name.c_str()) == 1; }
  std::vector<int64_t> dims(const std::string& name) const {
    int r = 0;
    int64_t d[8];
    if (io_.info(f_, name.c_str(), &r, d) != 0) throw std::runtime_error(path_ + ": " + io_.last_error());
    return std::vector<int64_t>(d, d + r);
  }
  void write(const std::string& name, int type, std::vector<int64_t> dims, const void* data) {
    if (io_.write(f_, name.c_str(), type, (int)dims.size(), dims.data(), data) != 0)
      throw std::runtime_error(path_ + ": " + io_.last_error());
  }
  void read(const std::string& name, int type, int64_t count, void* data) const {
    if (io_.read(f_, name.c_str(), type, count, data) != 0) throw std::runtime_error(path_ + ": " + io_.last_error());
  }
  double scalar(const std::string& name) const {
    double v = 0;
    read(name, FOPIS_IO_F64, 1, &v);
    return v;
  }

 private:
  const IoApi& io_;
  std::string path_;
  fopis_h5* f_ = nullptr;
};

static void make_dirs(const std::string& dir) {
  std::string cur;
  size_t b = 0;
  while (b <= dir.size()) {
    size_t e = dir.find('/', b);
    if (e == std::string::npos) e = dir.size();
    cur = dir.substr(0, e);
    if (!cur.empty()) mkdir(cur.c_str(), 0755);
    b = e + 1;
  }
}

static bool file_exists(const std::string& p) { return (bool)std::ifstream(p); }

static std::string base_name(const std::string& p) {
  const size_t sl = p.rfind('/');
  return sl == std::string::npos ? p : p.substr(sl + 1);
}

// ---------------------------------------------------------- the facade
class GpuRootGrid {
 public:
  GpuRootGrid(const std::string& path, const fopis_host_options& opt);
  ~GpuRootGrid() {
    if (worker_.joinable()) worker_.join();
    if (ctx_) fopis_destroy(ctx_);
  }
  // RootGrid surface (RootGrid.cpp, RootGridField.cpp, RootGridParticle.cpp)
  void initialize() { check(fopis_initialize(ctx_)); }
  void mainLoopES();
  void updateParticleVelocity() { check(fopis_update_particle_velocity_es(ctx_)); }
  void updateParticlePosition() { check(fopis_update_particle_position_es(ctx_)); }
  void updateDensity() { check(fopis_update_density(ctx_)); }
  void updateRho() { check(fopis_update_rho(ctx_)); }
  void solvePoisson() { check(fopis_solve_poisson(ctx_, nullptr, nullptr)); }
  void updateEfield() { check(fopis_update_efield(ctx_)); }
  void injectParticlesFromBoundary();
  void emitParticlesFromObjects();
  // RootGrid::saveResumeData / loadResumeData (RootGridIO.cpp:25-290) and
  // Environment::saveInfo / loadInfo (Environment.cpp:379-438) into dir
  void saveResumeData(const std::string& dir);
  void loadResumeData(const std::string& dir);
  // Grid::initializeForLoad (Grid.cpp:38-50): updateRho (which, with objects,
  // solves once for the redistribution), then the reference E from the
  // loaded field (here derived from the loaded phi, which is put back)
  void initializeForLoad() {
    updateRho();
    if (!resume_phi_.empty()) check(fopis_set_field(ctx_, FOPIS_PHI, 0, resume_phi_.data()));
    resume_phi_.clear();
    updateEfield();
  }
  std::vector<Particle6> generateInjection(int s);
  void emitInto(StepBatch& b);
  void emitHost(StepBatch& b);  // the reference's mt19937 draws on the host
  void injectInto(StepBatch& b);
  void stageBatch(StepBatch& b);
  StepBatch newBatch() const;

  fopis_ctx* ctx() { return ctx_; }
  const fopis_desc& desc() const { return d_; }
  int timestep() const { return timestep_; }
  double generate_ms() const { return gen_ms_; }
  const Object& object(int o) const { return objs_.at(o); }
  int nobjects() const { return (int)objs_.size(); }

 private:
  static void check(int rc) {
    if (rc != FOPIS_OK) throw std::runtime_error(fopis_last_error());
  }
  void load_initial();

 public:
  // the initial load again with another seed (device loader only)
  void reload(uint64_t seed) {
    if (!opt_.device_loader) throw std::invalid_argument("reload needs the device loader");
    opt_.device_seed = seed;
    load_initial();
  }
  void generateInitial(int s, std::vector<double> (&a)[6]);
  std::array<double, 6> flux(int s) const;

 private:
  void stage(int s, const std::vector<Particle6>& ps);
  bool inside_object(double x, double y, double z) const;

  fopis_desc d_{};
  Units u_;
  int nx_ = 0, ny_ = 0, nz_ = 0;
  std::vector<Species> sp_;
  std::vector<Object> objs_;
  std::vector<std::array<double, 6>> inj_residual_;  // RootGridParticle.cpp:4 (function-static there)
  uint64_t inj_calls_ = 0;                           // device injection: one seed per call
  uint64_t emit_calls_ = 0;                          // device photoemission: one seed per call
  double shine_[3] = {0, 0, 0};
  fopis_host_options opt_{};
  fopis_ctx* ctx_ = nullptr;
  int timestep_ = 1;
  double gen_ms_ = 0;
  std::thread worker_;   // draws the next step's batch while the device runs this one
  StepBatch next_;
  bool have_next_ = false;
  std::string base_dir_;
  double env_dx_ = 0, env_dt_ = 0;  // Environment::dx, dt as read from the input (SI)
  bool load_job_ = false;
  std::vector<double> resume_phi_;  // phi read by loadResumeData, restored by initializeForLoad
};

// AmbientParticleType::calcFlux (AmbientParticleType.cpp:8-57): rectangle
// rule over +-6 thermal widths of the drifting Maxwellian half fluxes.
static std::array<double, 6> calc_flux(const Species& s, const Units& u) {
  const double ddv = 1e-3;
  const double nvx = 12.0 / ddv;
  std::array<double, 6> f{0, 0, 0, 0, 0, 0};
  const double vth = s.thermal_velocity(u);
  const double nd = u.density(s.density);
  for (int a = 0; a < 3; ++a) {
    const double r = s.drift[a] / vth;
    for (int i = 0; i < nvx - 1; ++i) {
      const double a1 = i * ddv + (r - 6.0);
      const double e = std::exp(-std::pow(a1 - r, 2));
      if (a1 > 0.0)
        f[2 * a] += a1 * e * ddv;
      else
        f[2 * a + 1] += a1 * e * ddv;
    }
    f[2 * a] *= (vth / std::sqrt(M_PI)) * nd / s.size;
    f[2 * a + 1] *= -1.0 * (vth / std::sqrt(M_PI)) * nd / s.size;
  }
  for (int a = 0; a < 6; ++a)
    if (!s.inj[a]) f[a] = 0.0;
  return f;
}

GpuRootGrid::GpuRootGrid(const std::string& path, const fopis_host_options& opt) : opt_(opt) {
  const json::Value root = json::parse_file(path);
  const size_t slash = path.find_last_of('/');
  base_dir_ = slash == std::string::npos ? std::string(".") : path.substr(0, slash);
  const json::Value& env = root.at("Environment");
  nx_ = (int)env.at("nx").number();
  ny_ = (int)env.at("ny").number();
  nz_ = (int)env.at("nz").number();
  if (opt.nx_override > 0) nx_ = ny_ = nz_ = opt.nx_override;
  for (const char* k : {"proc_x", "proc_y", "proc_z"})
    if (env.has(k) && (int)env.at(k).number() != 1)
      throw std::invalid_argument("one GPU holds one rank's whole sub-domain: proc_x = proc_y = proc_z = 1");
  if (env.has("solver_type") && env.at("solver_type").string() != "ES")
    throw std::invalid_argument("only the electrostatic step (solver_type \"ES\") runs on this path");
  if (env.has("job_type")) {  // "new", or "load": continue from the resume files (Initialize.cpp:44-47,83-95)
    const std::string jt = env.at("job_type").string();
    if (jt != "new" && jt != "load") throw std::invalid_argument("unknown job_type: " + jt);
    load_job_ = jt == "load";
  }
  const double dx = env.at("dx").number(), dt = env.at("dt").number();
  env_dx_ = dx;
  env_dt_ = dt;
  u_.set(dx, dt);
  const std::string boundary = env.has("boundary") ? env.at("boundary").string() : "DDDDDD";
  if (boundary.size() != 6) throw std::invalid_argument("boundary must have 6 letters");
  for (char c : boundary) {  // Field::setBoundaryConditionPhi (Field.cpp:25-29)
    if (c == 'P') throw std::invalid_argument("periodic faces are not supported by this build");
    if (c != 'D' && c != 'N')
      throw std::invalid_argument(std::string("Unknown boundary condition type was used: boundary = ") + c);
  }

  // static fields (Initialize.cpp:306-327)
  double b[3] = {0, 0, 0};
  if (root.has("Field")) {
    const json::Value& f = root.at("Field");
    if (f.has("static_bfield")) {
      const auto& a = f.at("static_bfield").arr;
      const double bn = u_.bfield(1.0);
      for (int q = 0; q < 3 && q < (int)a.size(); ++q) b[q] = a[q].number() * (bn * 1e-9);
    }
    if (f.has("shine_vector")) {
      const auto& a = f.at("shine_vector").arr;
      for (int q = 0; q < 3 && q < (int)a.size(); ++q) shine_[q] = a[q].number();
    }
  }

  // species (Initialize.cpp:329-421): ids by type group then name order
  const json::Value& plasma = root.at("Plasma");
  std::vector<std::string> amb, photo;
  for (const auto& kv : plasma.obj) {
    const std::string& t = kv.second.at("type").string();
    if (t == "ambient")
      amb.push_back(kv.first);
    else if (t == "photoelectron")
      photo.push_back(kv.first);
    else
      throw std::invalid_argument("species type '" + t + "' is not on this path (ambient, photoelectron)");
  }
  std::vector<std::string> names = amb;
  names.insert(names.end(), photo.begin(), photo.end());
  if (names.empty() || names.size() > FOPIS_MAX_SPECIES) throw std::invalid_argument("1..8 species required");
  for (size_t q = 0; q < names.size(); ++q) {
    const json::Value& p = plasma.at(names[q]);
    Species s;
    s.name = names[q];
    s.type = p.at("type").string();
    s.set_id((int)q);
    s.mass = p.at("mass").number();
    s.charge = p.at("charge").number();
    s.temperature = kE * p.at("temperature").number();
    s.density = p.at("density").number();
    s.ppc = (int)p.at("particle_per_cell").number();
    if (opt.ppc_override > 0 && s.type == "ambient") s.ppc = opt.ppc_override;
    if (p.has("drift_velocity")) {
      const auto& a = p.at("drift_velocity").arr;
      for (int k = 0; k < 3 && k < (int)a.size(); ++k) s.drift[k] = a[k].number() * (1e3 * u_.velocity(1.0));
    }
    if (p.has("injection_axis")) {
      const auto& a = p.at("injection_axis").arr;
      if (a.size() != 6) throw std::invalid_argument("injection_axis needs 6 entries");
      for (int k = 0; k < 6; ++k) s.inj[k] = a[k].boolean();
    }
    if (p.has("current_density")) s.current_density = u_.current_density(p.at("current_density").number());
    s.size = static_cast<int>(std::pow(dx, 3) * s.density / static_cast<double>(s.ppc));  // ParticleType.cpp:33-36
    s.mt_azimuth.seed(586936);  // random_distribution.hpp:19 (rank 0)
    s.mt_depression.seed(8785997666ULL);
    sp_.push_back(std::move(s));
  }
  // Environment::checkPlasmaInfo (Environment.cpp:205-277)
  for (const auto& s : sp_) {
    if (s.type != "ambient") continue;
    const double debye = std::sqrt(s.temperature * kEps0 / (std::pow(kE, 2) * s.density));
    if (s.thermal_velocity(u_) >= 0.4 || debye / dx < 1.0)
      throw std::invalid_argument("There are some critical Plasma constraints violations. Check your input.");
  }

  // device description
  d_.nx = nx_;
  d_.ny = ny_;
  d_.nz = nz_;
  std::snprintf(d_.boundary, sizeof(d_.boundary), "%s", boundary.c_str());
  d_.nspecies = (int)sp_.size();
  for (size_t q = 0; q < sp_.size(); ++q) {
    d_.species[q].charge = sp_[q].charge;
    d_.species[q].mass = sp_[q].mass;
    d_.species[q].size = sp_[q].size;
  }
  for (int q = 0; q < 3; ++q) d_.static_b[q] = b[q];
  d_.eps0 = u_.eps0;
  d_.precision = FOPIS_FP64;
  d_.mg_precision = FOPIS_FP64;
  if (root.has("Build")) {
    const json::Value& bd = root.at("Build");
    if (bd.has("precision")) d_.precision = bd.at("precision").string() == "fp32" ? FOPIS_FP32 : FOPIS_FP64;
    if (bd.has("mg_precision")) d_.mg_precision = bd.at("mg_precision").string() == "fp32" ? FOPIS_FP32 : FOPIS_FP64;
    if (bd.has("mg_omega")) d_.mg_omega = bd.at("mg_omega").number();
    if (bd.has("mg_pre")) d_.mg_pre = (int)bd.at("mg_pre").number();
    if (bd.has("mg_post")) d_.mg_post = (int)bd.at("mg_post").number();
    if (bd.has("mg_max_cycles")) d_.mg_max_cycles = (int)bd.at("mg_max_cycles").number();
    if (bd.has("poisson_tol")) d_.poisson_tol = bd.at("poisson_tol").number();
  }
  const bool dev = !opt.no_device;
  if (dev) check(fopis_create(&d_, &ctx_));
  inj_residual_.assign(sp_.size(), std::array<double, 6>{0, 0, 0, 0, 0, 0});

  // objects (RootGridObject.cpp:6-75, Spacecraft.cpp:16-152)
  if (root.has("Object")) {
    int o = 0;
    for (const auto& kv : root.at("Object").obj) {
      const json::Value& v = kv.second;
      Object ob;
      ob.name = kv.first;
      ob.file = v.at("file_name").string();
      if (v.has("surface_type") && v.at("surface_type").string() != "perfect")
        throw std::invalid_argument("only perfect-conductor objects are on this path");
      if (v.has("is_potential_fixed")) ob.fixed = v.at("is_potential_fixed").boolean();
      if (v.has("fixed_potential")) ob.fixed_potential = u_.potential(v.at("fixed_potential").number());
      if (v.has("initial_potential_offset")) ob.offset = u_.potential(v.at("initial_potential_offset").number());
      if (v.has("emit_particles"))
        for (const auto& e : v.at("emit_particles").obj)
          for (const auto& s : sp_)
            if (s.name == e.first && s.type == "photoelectron") ob.emit_species.push_back(s.id);
      std::string file = ob.file;
      if (!std::ifstream(file)) file = base_dir_ + "/" + ob.file;
      ob.geom = load_obj(file, nx_, ny_, nz_);
      ob.ncmat = (int)ob.geom.ijk.size() / 3;
      ob.node_cmat.assign((size_t)(nx_ + 2) * (ny_ + 2) * (nz_ + 2), -1);
      for (int q = 0; q < ob.ncmat; ++q)
        ob.node_cmat[((size_t)ob.geom.ijk[3 * q] * (ny_ + 2) + ob.geom.ijk[3 * q + 1]) * (nz_ + 2) +
                     ob.geom.ijk[3 * q + 2]] = q;
      if (dev)
        check(fopis_set_object(ctx_, o, ob.geom.mask.data(), ob.ncmat, ob.geom.ijk.data(), ob.fixed ? 1 : 0,
                               ob.fixed_potential));
      objs_.push_back(std::move(ob));
      ++o;
    }
  }
  if (!dev) return;
  for (int o = 0; o < (int)objs_.size(); ++o) {
    const Object& ob = objs_[o];
    bool loaded = false;
    std::string cache;
    if (opt.cmatrix_cache[0]) {
      // IO::loadCmatrixData (IO.cpp:59-120): <cache>/objects/<obj file name>.h5 is
      // used when its capacity_matrix is ncmat x ncmat and its grid_spacing and
      // time_spacing equal the input's dx and dt
      cache = std::string(opt.cmatrix_cache) + "/objects/" + base_name(ob.file) + ".h5";
      if (file_exists(cache)) {
        H5 f(cache, false);
        if (f.exists("capacity_matrix")) {
          const std::vector<int64_t> d = f.dims("capacity_matrix");
          const bool valid = d.size() == 2 && d[0] == ob.ncmat && d[1] == ob.ncmat &&
                             f.scalar("grid_spacing") == env_dx_ && f.scalar("time_spacing") == env_dt_;
          if (valid) {
            std::vector<double> C((size_t)ob.ncmat * ob.ncmat);
            f.read("capacity_matrix", FOPIS_IO_F64, (int64_t)C.size(), C.data());
            check(fopis_set_cmatrix(ctx_, o, C.data()));
            loaded = true;
          }
        }
      }
    }
    if (!loaded && opt.build_cmatrix) {
      check(fopis_build_cmatrix(ctx_, o));
      if (!cache.empty()) {
        // IO::writeCmatrixData (IO.cpp:28-57): element [i][j] = Spacecraft::getCmatValue(i, j)
        std::vector<double> C((size_t)ob.ncmat * ob.ncmat);
        check(fopis_get_cmatrix(ctx_, o, C.data()));
        make_dirs(std::string(opt.cmatrix_cache) + "/objects");
        H5 f(cache, true);
        f.write("capacity_matrix", FOPIS_IO_F64, {ob.ncmat, ob.ncmat}, C.data());
        f.write("grid_spacing", FOPIS_IO_F64, {}, &env_dx_);
        f.write("time_spacing", FOPIS_IO_F64, {}, &env_dt_);
      }
    }
    // Spacecraft::initializeChargeMapOffset (Spacecraft.cpp:160-183)
    if (ob.offset != 0.0 && (loaded || opt.build_cmatrix)) {
      std::vector<double> C((size_t)ob.ncmat * ob.ncmat), cmap((size_t)sp_.size() * ob.ncmat, 0.0);
      check(fopis_get_cmatrix(ctx_, o, C.data()));
      for (int i = 0; i < ob.ncmat; ++i) {
        double dr = 0.0;
        for (int j = 0; j < ob.ncmat; ++j) dr += C[(size_t)i * ob.ncmat + j] * ob.offset;
        cmap[i] += dr;
      }
      check(fopis_set_object_charge_map(ctx_, o, cmap.data()));
    }
  }
  if (opt.load_particles) load_initial();
  if (load_job_) {
    // RootGrid::loadResumeData, then Grid::initializeForLoad (Grid.cpp:38-50):
    // rho from the loaded particles, reference E from the loaded phi
    loadResumeData(opt.resume_dir[0] ? std::string(opt.resume_dir) : base_dir_ + "/resume");
    initializeForLoad();
  } else if (opt.initialize) {
    initialize();
  }
}

void GpuRootGrid::saveResumeData(const std::string& dir) {
  make_dirs(dir);
  {  // Environment::saveInfo (Environment.cpp:421-438)
    H5 f(dir + "/environment.h5", true);
    f.write("dx", FOPIS_IO_F64, {}, &env_dx_);
    f.write("dt", FOPIS_IO_F64, {}, &env_dt_);
    const int32_t ts = timestep_;
    f.write("timestep", FOPIS_IO_I32, {}, &ts);
  }
  // RootGrid::saveResumeData (RootGridIO.cpp:42-53); only valid particles exist on the device
  H5 f(dir + "/grid_id_0.h5", true);
  for (const Species& s : sp_) {  // saveResumeParticleData (:133-181)
    int64_t n = 0;
    check(fopis_num_particles(ctx_, s.id, &n));
    std::vector<double> a[6];
    for (auto& v : a) v.resize((size_t)n);
    int64_t got = 0;
    check(fopis_get_particles(ctx_, s.id, n, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(),
                              a[5].data(), &got));
    const double size = (double)got;  // createDataSet<double>("size", ...) of a size_t
    f.write(s.name + "/size", FOPIS_IO_F64, {}, &size);
    f.write(s.name + "/generated_counts", FOPIS_IO_U64, {6}, s.generated);
    const char* comp[6] = {"x", "y", "z", "vx", "vy", "vz"};
    for (int c = 0; c < 6; ++c) f.write(s.name + "/" + comp[c], FOPIS_IO_F64, {got}, a[c].data());
  }
  {  // saveResumeFieldData (:227-262): phi, the edge E, and B (never sized in ES mode)
    const int64_t cx = nx_ + 2, cy = ny_ + 2, cz = nz_ + 2;
    std::vector<double> buf((size_t)(cx * cy * cz));
    check(fopis_get_field(ctx_, FOPIS_PHI, 0, buf.data()));
    f.write("Potential/phi", FOPIS_IO_F64, {cx, cy, cz}, buf.data());
    const std::vector<int64_t> ed[3] = {{cx - 1, cy, cz}, {cx, cy - 1, cz}, {cx, cy, cz - 1}};
    const char* en[3] = {"Efield/ex", "Efield/ey", "Efield/ez"};
    for (int q = 0; q < 3; ++q) {
      check(fopis_get_field(ctx_, FOPIS_EX + q, 0, buf.data()));
      f.write(en[q], FOPIS_IO_F64, ed[q], buf.data());
    }
    for (const char* b : {"Bfield/bx", "Bfield/by", "Bfield/bz"}) f.write(b, FOPIS_IO_F64, {0, 0, 0}, nullptr);
  }
  for (int o = 0; o < (int)objs_.size(); ++o) {  // saveResumeObjectData (:273-280)
    const int nc = objs_[o].ncmat;
    std::vector<double> cmap(sp_.size() * (size_t)nc);
    check(fopis_get_object_state(ctx_, o, nullptr, nullptr, nullptr, cmap.data()));
    f.write(objs_[o].name + "/charge_map", FOPIS_IO_F64, {(int64_t)sp_.size(), nc}, cmap.data());
  }
}

void GpuRootGrid::loadResumeData(const std::string& dir) {
  {  // Environment::loadInfo (Environment.cpp:379-419)
    const std::string fn = dir + "/environment.h5";
    if (!file_exists(fn)) throw std::invalid_argument("[ERROR] resume file_name " + fn + " does not exist.");
    H5 f(fn, false);
    if (f.scalar("dx") != env_dx_) throw std::invalid_argument("dx in the resume data does not match the input's dx");
    if (f.scalar("dt") != env_dt_) throw std::invalid_argument("dt in the resume data does not match the input's dt");
    int32_t ts = 0;
    f.read("timestep", FOPIS_IO_I32, 1, &ts);
    timestep_ = ts;
  }
  const std::string fn = dir + "/grid_id_0.h5";
  if (!file_exists(fn)) return;  // RootGrid::loadResumeData reads only an existing file
  H5 f(fn, false);
  for (Species& s : sp_) {  // loadResumeParticleData (:55-131)
    const int64_t n = (int64_t)f.scalar(s.name + "/size");
    std::vector<double> a[6];
    if (n > 0) {
      // ParticleType::proceedGeneratedCounts (ParticleType.cpp:70-103): each
      // engine skips (saved count - own count) outputs
      uint64_t target[6];
      f.read(s.name + "/generated_counts", FOPIS_IO_U64, 6, target);
      for (int c = 0; c < 6; ++c) {
        if (target[c] < s.generated[c])
          throw std::invalid_argument("resume generated_counts are behind this run's streams");
        s.mt[c].discard(target[c] - s.generated[c]);
        s.generated[c] = target[c];
      }
      const char* comp[6] = {"x", "y", "z", "vx", "vy", "vz"};
      for (int c = 0; c < 6; ++c) {
        a[c].resize((size_t)n);
        f.read(s.name + "/" + comp[c], FOPIS_IO_F64, n, a[c].data());
      }
    }
    check(fopis_set_particles(ctx_, s.id, n, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(),
                              a[5].data()));
  }
  {  // loadResumeFieldData (:183-225): phi; E and B are checked for presence (E is derived from phi here)
    const int64_t cx = nx_ + 2, cy = ny_ + 2, cz = nz_ + 2;
    std::vector<double> phi((size_t)(cx * cy * cz));
    f.read("Potential/phi", FOPIS_IO_F64, (int64_t)phi.size(), phi.data());
    check(fopis_set_field(ctx_, FOPIS_PHI, 0, phi.data()));
    resume_phi_ = std::move(phi);
    for (const char* e : {"Efield/ex", "Efield/ey", "Efield/ez"})
      if (!f.exists(e)) throw std::invalid_argument(std::string("resume data lacks ") + e);
  }
  for (int o = 0; o < (int)objs_.size(); ++o) {  // loadResumeObjectData (:264-271)
    const int nc = objs_[o].ncmat;
    std::vector<double> cmap(sp_.size() * (size_t)nc);
    f.read(objs_[o].name + "/charge_map", FOPIS_IO_F64, (int64_t)cmap.size(), cmap.data());
    check(fopis_set_object_charge_map(ctx_, o, cmap.data()));
  }
}

bool GpuRootGrid::inside_object(double x, double y, double z) const {
  const int i = (int)std::floor(x) + 1, j = (int)std::floor(y) + 1, k = (int)std::floor(z) + 1;
  for (const auto& o : objs_)
    if (o.cell(nx_, ny_, nz_, i, j, k)) return true;
  return false;
}

// RootGrid.cpp:45-79 with AmbientParticleType::generateNew{Position,Velocity}
// (AmbientParticleType.cpp:76-93): every component has its own engine, so
// the six streams are drawn on separate threads with identical results.
void GpuRootGrid::generateInitial(int sid, std::vector<double> (&a)[6]) {
  Species& s = sp_[sid];
  const double mx[3] = {(double)nx_ - 1.0, (double)ny_ - 1.0, (double)nz_ - 1.0};  // upper faces are physical
  const size_t total = s.type == "ambient" ? (size_t)(nx_ - 1) * (ny_ - 1) * (nz_ - 1) * s.ppc : 0;
  for (auto& v : a) v.resize(total);
  const double dev = s.deviation(u_);
  auto gen = [&](int c) {
    if (c < 3) {
      std::uniform_real_distribution<> d(0.0, mx[c]);
      for (size_t n = 0; n < total; ++n) a[c][n] = d(s.mt[c]);
    } else {
      for (size_t n = 0; n < total; ++n) {
        std::normal_distribution<> d(s.drift[c - 3], dev);  // a fresh distribution per draw
        a[c][n] = d(s.mt[c]);
      }
    }
  };
  std::vector<std::thread> th;
  for (int c = 0; c < 6; ++c) th.emplace_back(gen, c);
  for (auto& t : th) t.join();
  for (int c = 0; c < 6; ++c) s.generated[c] += total;  // AmbientParticleType.cpp:86,97
  // Spacecraft::removeInnerParticle: drop the particles born inside an object
  size_t w = 0;
  for (size_t n = 0; n < total; ++n) {
    if (!objs_.empty() && inside_object(a[0][n], a[1][n], a[2][n])) continue;
    for (int c = 0; c < 6; ++c) a[c][w] = a[c][n];
    ++w;
  }
  for (auto& v : a) v.resize(w);
}

void GpuRootGrid::load_initial() {
  const double mx[3] = {(double)nx_ - 1.0, (double)ny_ - 1.0, (double)nz_ - 1.0};
  for (auto& s : sp_) {
    if (s.type != "ambient") continue;
    const size_t total = (size_t)(nx_ - 1) * (ny_ - 1) * (nz_ - 1) * s.ppc;  // ParticleType.cpp:39-47
    if (opt_.device_loader) {
      const double lo[3] = {0, 0, 0};
      check(fopis_load_maxwellian(ctx_, s.id, (int64_t)total, opt_.device_seed + 7919ull * s.id, lo, mx, s.drift,
                                  s.deviation(u_)));
      continue;
    }
    std::vector<double> a[6];
    generateInitial(s.id, a);
    check(fopis_set_particles(ctx_, s.id, (int64_t)a[0].size(), a[0].data(), a[1].data(), a[2].data(),
                              a[3].data(), a[4].data(), a[5].data()));
  }
}

std::array<double, 6> GpuRootGrid::flux(int s) const { return calc_flux(sp_.at(s), u_); }

void GpuRootGrid::stage(int s, const std::vector<Particle6>& ps) {
  if (ps.empty()) return;
  std::vector<double> a[6];
  for (auto& v : a) v.resize(ps.size());
  for (size_t n = 0; n < ps.size(); ++n)
    for (int c = 0; c < 6; ++c) a[c][n] = ps[n].v[c];
  check(fopis_add_particles(ctx_, s, (int64_t)ps.size(), a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                            a[4].data(), a[5].data()));
}

// RootGrid::injectParticlesFromBoundary (RootGridParticle.cpp:3-130) for one
// species: per face, floor(dt*A*flux + residue) particles; normal velocity
// N(u, vth) redrawn until it points inward, tangential N(u, sigma); position
// uniform across the face and within [0, v*dt] of it.
std::vector<Particle6> GpuRootGrid::generateInjection(int sid) {
  std::vector<Particle6> out;
  Species& s = sp_[sid];
  if (s.type != "ambient") return out;
  const std::array<double, 6> flux = calc_flux(s, u_);
  const double mx[3] = {(double)nx_ - 1.0, (double)ny_ - 1.0, (double)nz_ - 1.0};
  const int ns[3] = {nx_, ny_, nz_};
  const double dev = s.deviation(u_), vth = s.thermal_velocity(u_);
  for (int face = 0; face < 6; ++face) {
    const int axis = face / 2, side = face % 2;
    if (d_.boundary[face] == 'P') continue;
    const int a1 = axis == 0 ? 2 : (axis == 1 ? 2 : 0), a2 = axis == 0 ? 1 : (axis == 1 ? 0 : 1);
    const double area = 1.0 * 1.0 * (ns[a1] - 1) * (ns[a2] - 1);
    const double nd = 1.0 * area * flux[face];
    const int num = static_cast<int>(std::floor(nd + inj_residual_[sid][face]));
    inj_residual_[sid][face] += nd - num;
    const double sign = side == 0 ? 1.0 : -1.0;
    for (int q = 0; q < num; ++q) {
      std::normal_distribution<> dx_(s.drift[0], axis == 0 ? vth : dev), dy_(s.drift[1], axis == 1 ? vth : dev),
          dz_(s.drift[2], axis == 2 ? vth : dev);
      double v[3] = {dx_(s.mt[3]), dy_(s.mt[4]), dz_(s.mt[5])};
      while (v[axis] * sign < 0.0) {
        v[0] = dx_(s.mt[3]);
        v[1] = dy_(s.mt[4]);
        v[2] = dz_(s.mt[5]);
      }
      double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {mx[0], mx[1], mx[2]};
      if (side == 0)
        hi[axis] = v[axis] * 1.0;
      else
        lo[axis] = mx[axis] + v[axis] * 1.0;
      Particle6 p;
      for (int c = 0; c < 3; ++c) {
        std::uniform_real_distribution<> d(lo[c], hi[c]);
        p.v[c] = d(s.mt[c]);
        ++s.generated[c];  // generateNewPosition counts; the injection velocity does not
      }
      p.v[3] = v[0];
      p.v[4] = v[1];
      p.v[5] = v[2];
      out.push_back(p);
    }
  }
  return out;
}

StepBatch GpuRootGrid::newBatch() const {
  StepBatch b;
  b.particles.resize(sp_.size());
  b.dq.resize(objs_.size());
  b.dcur.resize(objs_.size());
  for (size_t o = 0; o < objs_.size(); ++o) {
    b.dq[o].assign(sp_.size() * (size_t)objs_[o].ncmat, 0.0);
    b.dcur[o].assign(sp_.size(), 0.0);
  }
  return b;
}

// Each species draws from its own engines, so species run on separate threads
// and the draws stay in the reference's per-stream order.
void GpuRootGrid::injectInto(StepBatch& b) {
  if (opt_.device_injection) {
    // the per-face counts and residues exactly as the reference keeps them;
    // the particles themselves are drawn on the device when the batch is staged
    b.inj_count.assign(sp_.size(), std::array<int64_t, 6>{0, 0, 0, 0, 0, 0});
    b.inj_seed = opt_.device_seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(inj_calls_ + 1));
    ++inj_calls_;
    const int ns[3] = {nx_, ny_, nz_};
    for (size_t sid = 0; sid < sp_.size(); ++sid) {
      const Species& s = sp_[sid];
      if (s.type != "ambient") continue;
      const std::array<double, 6> flux = calc_flux(s, u_);
      for (int face = 0; face < 6; ++face) {
        const int axis = face / 2;
        if (d_.boundary[face] == 'P') continue;
        const int a1 = axis == 0 ? 2 : (axis == 1 ? 2 : 0), a2 = axis == 0 ? 1 : (axis == 1 ? 0 : 1);
        const double nd = 1.0 * (ns[a1] - 1) * (ns[a2] - 1) * flux[face];
        const int num = static_cast<int>(std::floor(nd + inj_residual_[sid][face]));
        inj_residual_[sid][face] += nd - num;
        b.inj_count[sid][face] = num;
      }
    }
    return;
  }
  std::vector<std::vector<Particle6>> out(sp_.size());
  std::vector<std::thread> th;
  for (size_t s = 0; s < sp_.size(); ++s)
    if (sp_[s].type == "ambient") th.emplace_back([this, s, &out] { out[s] = generateInjection((int)s); });
  for (auto& t : th) t.join();
  for (size_t s = 0; s < sp_.size(); ++s) b.particles[s].insert(b.particles[s].end(), out[s].begin(), out[s].end());
}

void GpuRootGrid::stageBatch(StepBatch& b) {
  for (size_t s = 0; s < sp_.size(); ++s) stage((int)s, b.particles[s]);
  for (size_t s = 0; s < b.inj_count.size(); ++s) {
    const Species& sp = sp_[s];
    if (sp.type != "ambient") continue;
    check(fopis_inject_maxwellian(ctx_, (int)s, b.inj_count[s].data(), b.inj_seed + 7919ull * s, sp.drift,
                                  sp.thermal_velocity(u_), sp.deviation(u_)));
  }
  for (const StepBatch::Emit& e : b.emit)
    check(fopis_emit_photoelectrons(ctx_, e.o, e.s, shine_, e.amount, e.vth, e.seed));
  if (!opt_.device_injection)
    for (size_t o = 0; o < objs_.size(); ++o)
      if (!objs_[o].emit_species.empty())
        check(fopis_add_object_charge(ctx_, (int)o, b.dq[o].data(), b.dcur[o].data()));
}

void GpuRootGrid::injectParticlesFromBoundary() {
  StepBatch b = newBatch();
  injectInto(b);
  stageBatch(b);
}

void GpuRootGrid::emitParticlesFromObjects() {
  StepBatch b = newBatch();
  emitInto(b);
  stageBatch(b);
}

// Spacecraft::emitParticles, photoelectron branch (Spacecraft.cpp:829-900),
// PhotoElectronParticleType (PhotoElectronParticleType.cpp:9-106) and the
// charge taken from the emitting face (Spacecraft.cpp:945-970). The emitted
// particles then get this step's position update, as in the reference where
// emission precedes updateParticlePosition.
void GpuRootGrid::emitInto(StepBatch& b) {
  if (opt_.device_injection) {
    // the device draws the particles, their surface charge and the current
    for (int o = 0; o < (int)objs_.size(); ++o)
      for (int sid : objs_[o].emit_species) {
        const Species& s = sp_[sid];
        const double amount = (1.0 * 1.0) * (s.current_density / (std::fabs(s.charge) * s.size));
        const uint64_t seed = (opt_.device_seed ^ (0xD1B54A32D192ED03ull * (emit_calls_ + 1))) + 1009ull * o + 31ull * sid;
        b.emit.push_back({o, sid, amount, s.thermal_velocity(u_), seed});
      }
    ++emit_calls_;
    return;
  }
  emitHost(b);
}

void GpuRootGrid::emitHost(StepBatch& b) {
  for (int o = 0; o < (int)objs_.size(); ++o) {
    Object& ob = objs_[o];
    if (ob.emit_species.empty()) continue;
    int ax = -1;
    for (int a = 0; a < 3; ++a)
      if (std::fabs(shine_[a]) == 1.0) {
        ax = a;
        break;
      }
    if (ax < 0) throw std::invalid_argument("At Spacecraft::emitParticles: Now 'shine_vector' must be [+-1,0,0] or [0,+-1,0] or [0,0,+-1].");
    const int ssign = shine_[ax] > 0.0 ? 1 : -1;
    std::vector<double>& dq = b.dq[o];
    std::vector<double>& dcur = b.dcur[o];
    auto cell = [&](int i, int j, int k) { return ob.cell(nx_, ny_, nz_, i, j, k); };
    auto cmat = [&](int i, int j, int k) -> int {
      if (i < 0 || j < 0 || k < 0 || i > nx_ + 1 || j > ny_ + 1 || k > nz_ + 1) return -1;
      return ob.node_cmat[((size_t)i * (ny_ + 2) + j) * (nz_ + 2) + k];
    };
    auto valid_cell = [&](int i, int j, int k) {
      return i >= 0 && j >= 0 && k >= 0 && i < nx_ + 1 && j < ny_ + 1 && k < nz_ + 1;
    };
    for (int sid : ob.emit_species) {
      Species& s = sp_[sid];
      const double q = s.super_charge();
      const double amount = (1.0 * 1.0) * (s.current_density / (std::fabs(s.charge) * s.size));
      const double ev[3] = {-shine_[0], -shine_[1], -shine_[2]};
      const double vth = s.thermal_velocity(u_);
      std::uniform_real_distribution<> unit(0.0, 1.0);
      std::vector<Particle6>& born = b.particles[sid];
      for (int c = 0; c < ob.ncmat; ++c) {
        const int I = ob.geom.ijk[3 * c], J = ob.geom.ijk[3 * c + 1], K = ob.geom.ijk[3 * c + 2];
        // isXsurfacePoint(node, sign) for the lit face (Spacecraft.cpp:461-504)
        int pi[3] = {I, J, K};
        bool surf;
        int dI = ax == 0, dJ = ax == 1, dK = ax == 2;
        if (valid_cell(I, J, K) && valid_cell(I - dI, J - dJ, K - dK))
          surf = ssign > 0 ? (cell(I, J, K) && !cell(I - dI, J - dJ, K - dK))
                           : (!cell(I, J, K) && cell(I - dI, J - dJ, K - dK));
        else {
          int e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
          e1[ax == 0 ? 1 : 0] = 1;
          e2[ax == 2 ? 1 : 2] = 1;
          surf = cmat(I, J, K) >= 0 && cmat(I + e1[0], J + e1[1], K + e1[2]) >= 0 &&
                 cmat(I + e2[0], J + e2[1], K + e2[2]) >= 0 &&
                 cmat(I + e1[0] + e2[0], J + e1[1] + e2[1], K + e1[2] + e2[2]) >= 0;
        }
        (void)pi;
        if (!surf) continue;
        const double nxp = I - 1.0, nyp = J - 1.0, nzp = K - 1.0;
        for (int n = 0; n < amount; ++n) {
          const double dep = std::asin(2.0 * unit(s.mt_depression) - 1.0);
          const double azi = 2.0 * M_PI * unit(s.mt_azimuth);
          double v[3];
          if (ax == 0) {
            v[0] = vth * std::cos(dep) * ev[0];
            v[1] = vth * std::sin(dep) * std::cos(azi);
            v[2] = vth * std::sin(dep) * std::sin(azi);
          } else if (ax == 1) {
            v[0] = vth * std::sin(dep) * std::sin(azi);
            v[1] = vth * std::cos(dep) * ev[1];
            v[2] = vth * std::sin(dep) * std::cos(azi);
          } else {
            v[0] = vth * std::sin(dep) * std::cos(azi);
            v[1] = vth * std::sin(dep) * std::sin(azi);
            v[2] = vth * std::cos(dep) * ev[2];
          }
          std::uniform_real_distribution<> d0(0.0, 1.0), d1(0.0, 1.0), d2(0.0, 1.0);
          const double rt = d0(s.mt[0]);
          double p[3];
          if (ax == 0) {
            const double ry = d1(s.mt[1]), rz = d2(s.mt[2]);
            p[0] = nxp + (rt - 1.0) * v[0];
            p[1] = nyp + (rt - 1.0) * v[1] + ry;
            p[2] = nzp + (rt - 1.0) * v[2] + rz;
          } else if (ax == 1) {
            const double rx = d1(s.mt[0]), rz = d2(s.mt[2]);
            p[0] = nxp + (rt - 1.0) * v[0] + rx;
            p[1] = nyp + (rt - 1.0) * v[1];
            p[2] = nzp + (rt - 1.0) * v[2] + rz;
          } else {
            const double rx = d1(s.mt[0]), ry = d2(s.mt[1]);
            p[0] = nxp + (rt - 1.0) * v[0] + rx;
            p[1] = nyp + (rt - 1.0) * v[1] + ry;
            p[2] = nzp + (rt - 1.0) * v[2];
          }
          // getNext{X,Y,Z}CrossPoint (particle.cpp:206-255), then the
          // bilinear share of -q on the four face nodes
          double c0[3] = {p[0], p[1], p[2]};
          const double fl[3] = {std::floor(c0[0]), std::floor(c0[1]), std::floor(c0[2])};
          double d1_[3] = {c0[0] - fl[0], c0[1] - fl[1], c0[2] - fl[2]};
          double d2_[3] = {1.0 - d1_[0], 1.0 - d1_[1], 1.0 - d1_[2]};
          if (v[ax] > 0.0) {
            const double m = d2_[ax] / v[ax];
            for (int a = 0; a < 3; ++a) c0[a] = a == ax ? std::ceil(p[a]) : p[a] + v[a] * m;
          } else if (v[ax] < 0.0) {
            const double m = d1_[ax] / (-v[ax]);
            for (int a = 0; a < 3; ++a) c0[a] = a == ax ? std::floor(p[a]) : p[a] + v[a] * m;
          }
          int ci[3];
          double w1[3], w2[3];
          for (int a = 0; a < 3; ++a) {
            ci[a] = (int)std::floor(c0[a]) + 1;
            w1[a] = c0[a] - (ci[a] - 1);
            w2[a] = 1.0 - w1[a];
          }
          const int ua = ax == 0 ? 1 : 0, wa = ax == 2 ? 1 : 2;
          for (int du = 0; du < 2; ++du)
            for (int dw = 0; dw < 2; ++dw) {
              int nn[3] = {ci[0], ci[1], ci[2]};
              nn[ua] += du;
              nn[wa] += dw;
              const int cm = cmat(nn[0], nn[1], nn[2]);
              const double wu = du ? w1[ua] : w2[ua], ww = dw ? w1[wa] : w2[wa];
              if (cm >= 0) dq[(size_t)sid * ob.ncmat + cm] -= q * wu * ww;
            }
          dcur[sid] -= q;
          // this step's position update and absorbing faces (grid.hpp:309-395)
          Particle6 np;
          for (int a = 0; a < 3; ++a) {
            np.v[a] = p[a] + v[a];
            np.v[3 + a] = v[a];
          }
          bool ok = true;
          for (int a = 0; a < 3; ++a) {
            const double lim = (a == 0 ? nx_ : (a == 1 ? ny_ : nz_)) - 1.0;
            if (np.v[a] < 0.0 || np.v[a] > lim) ok = false;
          }
          if (ok) born.push_back(np);
        }
      }
    }
  }
}

// RootGrid::mainLoopES (RootGrid.cpp:82-143): the host draws this step's
// emitted and injected particles (they need the reference's RNG streams),
// stages them, and the device runs the fused step.
//
// The host streams do not depend on the plasma state, so the batch for step
// n+1 is drawn on a worker thread while the device runs step n; the draws
// happen in the reference's order either way.
void GpuRootGrid::mainLoopES() {
  const auto t0 = std::chrono::steady_clock::now();
  StepBatch cur;
  if (have_next_) {
    worker_.join();
    cur = std::move(next_);
    have_next_ = false;
  } else {
    cur = newBatch();
    emitInto(cur);
    injectInto(cur);
  }
  stageBatch(cur);
  const auto t1 = std::chrono::steady_clock::now();
  gen_ms_ = std::chrono::duration<double, std::milli>(t1 - t0).count();
  next_ = newBatch();
  worker_ = std::thread([this] {
    emitInto(next_);
    injectInto(next_);
  });
  have_next_ = true;
  try {
    check(fopis_main_loop_es(ctx_));
  } catch (...) {
    worker_.join();
    have_next_ = false;
    throw;
  }
  ++timestep_;
}

}  // namespace fopis

// ------------------------------------------------------------- C entry points
using fopis::GpuRootGrid;

struct fopis_host {
  std::unique_ptr<GpuRootGrid> g;
};

static thread_local std::string g_host_err;

template <class F>
static int host_guard(F&& f) {
  try {
    f();
    return FOPIS_OK;
  } catch (const std::invalid_argument& e) {
    g_host_err = e.what();
    return FOPIS_E_ARG;
  } catch (const std::logic_error& e) {
    g_host_err = e.what();
    return FOPIS_E_STATE;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return FOPIS_E_DEVICE;
  }
}

extern "C" {

const char* fopis_host_last_error(void) { return g_host_err.c_str(); }

int fopis_host_open(const char* json_path, const fopis_host_options* opt, fopis_host** out) {
  if (!json_path || !out) return FOPIS_E_ARG;
  *out = nullptr;
  fopis_host_options o{};
  o.load_particles = 1;
  o.initialize = 1;
  o.build_cmatrix = 1;
  if (opt) o = *opt;
  auto h = std::make_unique<fopis_host>();
  const int rc = host_guard([&] { h->g = std::make_unique<GpuRootGrid>(json_path, o); });
  if (rc == FOPIS_OK) *out = h.release();
  return rc;
}

int fopis_host_close(fopis_host* h) {
  delete h;
  return FOPIS_OK;
}

fopis_ctx* fopis_host_ctx(fopis_host* h) { return h ? h->g->ctx() : nullptr; }

int fopis_host_save_resume(fopis_host* h, const char* dir) {
  if (!h || !dir) return FOPIS_E_ARG;
  return host_guard([&] { h->g->saveResumeData(dir); });
}

int fopis_host_load_resume(fopis_host* h, const char* dir) {
  if (!h || !dir) return FOPIS_E_ARG;
  return host_guard([&] {
    h->g->loadResumeData(dir);
    h->g->initializeForLoad();
  });
}

int fopis_host_desc(fopis_host* h, fopis_desc* out) {
  if (!h || !out) return FOPIS_E_ARG;
  *out = h->g->desc();
  return FOPIS_OK;
}

int fopis_host_main_loop_es(fopis_host* h) {
  if (!h) return FOPIS_E_ARG;
  return host_guard([&] { h->g->mainLoopES(); });
}

int fopis_host_timestep(fopis_host* h) { return h ? h->g->timestep() : -1; }

int fopis_host_reload(fopis_host* h, uint64_t seed) {
  if (!h) return FOPIS_E_ARG;
  return host_guard([&] { h->g->reload(seed); });
}

double fopis_host_generate_ms(fopis_host* h) { return h ? h->g->generate_ms() : 0.0; }

int64_t fopis_host_generate_injection(fopis_host* h, int s, int64_t cap, double* x, double* y, double* z, double* vx,
                                      double* vy, double* vz) {
  if (!h) return FOPIS_E_ARG;
  int64_t n = -1;
  const int rc = host_guard([&] {
    if (s < 0 || s >= h->g->desc().nspecies) throw std::invalid_argument("species out of range");
    auto ps = h->g->generateInjection(s);
    n = (int64_t)ps.size();
    double* o[6] = {x, y, z, vx, vy, vz};
    for (int64_t q = 0; q < n && q < cap; ++q)
      for (int c = 0; c < 6; ++c)
        if (o[c]) o[c][q] = ps[q].v[c];
  });
  return rc == FOPIS_OK ? n : rc;
}

int64_t fopis_host_generate_emission(fopis_host* h, int s, int64_t cap, double* x, double* y, double* z, double* vx,
                                     double* vy, double* vz, double* dq, double* dcurrent) {
  if (!h) return FOPIS_E_ARG;
  int64_t n = -1;
  const int rc = host_guard([&] {
    if (s < 0 || s >= h->g->desc().nspecies) throw std::invalid_argument("species out of range");
    fopis::StepBatch b = h->g->newBatch();
    h->g->emitHost(b);
    const auto& ps = b.particles[s];
    n = (int64_t)ps.size();
    double* o[6] = {x, y, z, vx, vy, vz};
    for (int64_t q = 0; q < n && q < cap; ++q)
      for (int c = 0; c < 6; ++c)
        if (o[c]) o[c][q] = ps[q].v[c];
    if (h->g->nobjects() > 0) {
      const int nc = h->g->object(0).ncmat;
      if (dq)
        for (int q = 0; q < nc; ++q) dq[q] = b.dq[0][(size_t)s * nc + q];
      if (dcurrent) *dcurrent = b.dcur[0][s];
    }
  });
  return rc == FOPIS_OK ? n : rc;
}

int64_t fopis_host_generate_initial(fopis_host* h, int s, int64_t cap, double* x, double* y, double* z, double* vx,
                                    double* vy, double* vz) {
  if (!h) return FOPIS_E_ARG;
  int64_t n = -1;
  const int rc = host_guard([&] {
    if (s < 0 || s >= h->g->desc().nspecies) throw std::invalid_argument("species out of range");
    std::vector<double> a[6];
    h->g->generateInitial(s, a);
    n = (int64_t)a[0].size();
    double* o[6] = {x, y, z, vx, vy, vz};
    for (int c = 0; c < 6; ++c)
      if (o[c]) std::memcpy(o[c], a[c].data(), sizeof(double) * (size_t)std::min<int64_t>(n, cap));
  });
  return rc == FOPIS_OK ? n : rc;
}

int fopis_host_flux(fopis_host* h, int s, double* flux6) {
  if (!h || !flux6) return FOPIS_E_ARG;
  return host_guard([&] {
    if (s < 0 || s >= h->g->desc().nspecies) throw std::invalid_argument("species out of range");
    const auto f = h->g->flux(s);
    for (int a = 0; a < 6; ++a) flux6[a] = f[a];
  });
}

int fopis_host_object(fopis_host* h, int obj, uint8_t* mask, int32_t* ncmat, int32_t* cmat_ijk, int32_t cap) {
  if (!h) return FOPIS_E_ARG;
  return host_guard([&] {
    if (obj < 0 || obj >= h->g->nobjects()) throw std::invalid_argument("object index out of range");
    const auto& o = h->g->object(obj);
    if (ncmat) *ncmat = o.ncmat;
    if (mask) std::memcpy(mask, o.geom.mask.data(), o.geom.mask.size());
    if (cmat_ijk)
      for (int q = 0; q < 3 * o.ncmat && q < 3 * cap; ++q) cmat_ijk[q] = o.geom.ijk[q];
  });
}

}  // extern "C"
