// ORACLE — test infrastructure only (see refcpu.hpp). CPU restatement of the
// reference's ES step; each function cites the reference file:line it follows.
#include "refcpu.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <regex>
#include <sstream>
#include <stdexcept>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace refcpu {

// ---------------------------------------------------------------- units
// include/normalizer.hpp:24-36,189-191
void Units::set(double dx, double dt) {
  x = dx;
  t = dt;
  m = kMe;
  e = kE;
  x3 = std::pow(x, 3);
  V = std::pow(x, 2) * m / (e * std::pow(t, 2));
  A = e / t;
  eps0 = kEps0 * (x3 * m) / (std::pow(e, 2) * std::pow(t, 2));
}

// ---------------------------------------------------------------- position
// include/position.hpp:13-21
void Position::updateDelta() {
  dx1 = x - (i - 1);
  dy1 = y - (j - 1);
  dz1 = z - (k - 1);
  dx2 = 1.0 - dx1;
  dy2 = 1.0 - dy1;
  dz2 = 1.0 - dz1;
}

// include/position.hpp:59-69
void Position::setXYZ(double x_, double y_, double z_) {
  x = x_;
  y = y_;
  z = z_;
  i = static_cast<int>(std::floor(x)) + 1;
  j = static_cast<int>(std::floor(y)) + 1;
  k = static_cast<int>(std::floor(z)) + 1;
  updateDelta();
}

// include/position.hpp:71-82
Position Position::fromIJK(int i_, int j_, int k_) {
  Position p;
  p.i = i_;
  p.j = j_;
  p.k = k_;
  p.x = static_cast<double>(i_ - 1);
  p.y = static_cast<double>(j_ - 1);
  p.z = static_cast<double>(k_ - 1);
  p.updateDelta();
  return p;
}

// ---------------------------------------------------------------- species
// src/Particle/ParticleType.cpp:19-27 (rank 0); random_distribution.hpp:19
void Species::setId(int id_) {
  id = id_;
  mt_x.seed(10684930 + 106 * id_);
  mt_y.seed(99881 + 171 * id_);
  mt_z.seed(861200045 + 95 * id_);
  mt_vx.seed(930 + 2 * id_);
  mt_vy.seed(98076621 - 50 * id_);
  mt_vz.seed(7662566 + 4 * id_);
}

// ParticleType.cpp:49-55 (temperature held as kB*T in J, particle_type.hpp:75-79)
double Species::thermalVelocity(const Units& u) const {
  return u.normVelocity(std::sqrt(2.0 * (kE * temperature) / (mass * kMe)));
}
double Species::deviation(const Units& u) const {
  return u.normVelocity(std::sqrt((kE * temperature) / (mass * kMe)));
}

// AmbientParticleType.cpp:8-57
std::vector<double> calcFlux(const Species& s, const Units& u) {
  const double ddv = 1e-3;
  const double nvx = 12.0 / ddv;
  std::vector<double> flux(6, 0.0);
  const double vth = s.thermalVelocity(u);
  const double nd = u.normDensity(s.density);
  const double size = s.size;
  for (int axis = 0; axis < 3; ++axis) {
    const double r = s.drift[axis] / vth;
    for (int i = 0; i < nvx - 1; ++i) {
      const double a1 = i * ddv + (r - 6.0);
      const double ax1 = std::pow(a1 - r, 2);
      if (a1 > 0.0) {
        flux[2 * axis] += a1 * std::exp(-ax1) * ddv;
      } else {
        flux[2 * axis + 1] += a1 * std::exp(-ax1) * ddv;
      }
    }
    flux[2 * axis] *= (vth / std::sqrt(M_PI)) * nd / size;
    flux[2 * axis + 1] *= -1.0 * (vth / std::sqrt(M_PI)) * nd / size;
  }
  for (int a = 0; a < 6; ++a)
    if (!s.inj[a]) flux[a] = 0.0;
  return flux;
}

// AmbientParticleType.cpp:76-83
static Position genPosition(Species& s, double x0, double x1, double y0, double y1, double z0, double z1) {
  std::uniform_real_distribution<> dx(x0, x1), dy(y0, y1), dz(z0, z1);
  const double a = dx(s.mt_x);
  const double b = dy(s.mt_y);
  const double c = dz(s.mt_z);
  return Position(a, b, c);
}

// AmbientParticleType.cpp:85-93 — a fresh normal_distribution per call
static void genVelocity(Species& s, const Units& u, double v[3]) {
#if REFCPU_MUTANT == 6  // planted mistake: the injection's sqrt(2T/m) in place of sqrt(T/m)
  const double dev = s.thermalVelocity(u);
#else
  const double dev = s.deviation(u);
#endif
  std::normal_distribution<> nx(s.drift[0], dev), ny(s.drift[1], dev), nz(s.drift[2], dev);
  v[0] = nx(s.mt_vx);
  v[1] = ny(s.mt_vy);
  v[2] = nz(s.mt_vz);
}

// AmbientParticleType.cpp:101-157
static void genInjectionVelocity(Species& s, const Units& u, int axis, int side, double v[3]) {
  const double dev = s.deviation(u);
  const double vth = s.thermalVelocity(u);
  const double sign = (side == 0) ? 1.0 : -1.0;
  const double sx = axis == 0 ? vth : dev, sy = axis == 1 ? vth : dev, sz = axis == 2 ? vth : dev;
  std::normal_distribution<> nx(s.drift[0], sx), ny(s.drift[1], sy), nz(s.drift[2], sz);
  v[0] = nx(s.mt_vx);
  v[1] = ny(s.mt_vy);
  v[2] = nz(s.mt_vz);
  while (v[axis] * sign < 0.0) {
    v[0] = nx(s.mt_vx);
    v[1] = ny(s.mt_vy);
    v[2] = nz(s.mt_vz);
  }
}

// ---------------------------------------------------------------- obj loader
static std::vector<std::string> split(const std::string& t, char d) {
  std::stringstream ss(t);
  std::string b;
  std::vector<std::string> r;
  while (std::getline(ss, b, d)) r.push_back(b);
  return r;
}

// src/Spacecraft/ObjectUtils.cpp:9-261: surface nodes numbered by first
// appearance, face maps, then per-axis ray fill; a cell is inside iff all
// three axes paint it (count == 3).
ObjData loadObjFile(const std::string& path, int nx, int ny, int nz) {
  std::ifstream in(path);
  if (!in) throw std::invalid_argument("File " + path + " does not exist.");
  ObjData od;
  std::vector<int> node_map((size_t)nx * ny * nz, -1);
  auto nm = [&](int i, int j, int k) -> int& { return node_map[((size_t)i * ny + j) * nz + k]; };
  std::vector<uint8_t> xf((size_t)nx * (ny - 1) * (nz - 1), 0), yf((size_t)(nx - 1) * ny * (nz - 1), 0),
      zf((size_t)(nx - 1) * (ny - 1) * nz, 0);
  auto XF = [&](int i, int j, int k) -> uint8_t& { return xf[((size_t)i * (ny - 1) + j) * (nz - 1) + k]; };
  auto YF = [&](int i, int j, int k) -> uint8_t& { return yf[((size_t)i * ny + j) * (nz - 1) + k]; };
  auto ZF = [&](int i, int j, int k) -> uint8_t& { return zf[((size_t)i * (ny - 1) + j) * nz + k]; };

  std::regex re_v("v (.*)"), re_f("f (.*)");
  std::vector<std::array<int, 3>> verts;
  std::vector<std::array<int, 4>> faces;
  std::string buf;
  while (!in.eof()) {
    std::getline(in, buf);
    if (std::regex_match(buf, re_v)) {
      auto t = split(buf, ' ');
      verts.push_back({std::stoi(t[1]) + nx / 2, std::stoi(t[2]) + ny / 2, std::stoi(t[3]) + nz / 2});
    } else if (std::regex_match(buf, re_f)) {
      auto t = split(buf, ' ');
      std::array<int, 4> f;
      for (int q = 1; q < 5; ++q) f[q - 1] = std::stoi(split(t[q], '/')[0]);
      faces.push_back(f);
    }
  }
  unsigned num = 0;
  auto touch = [&](int i, int j, int k) {
    if (nm(i, j, k) < 0) {
      od.nodes[num] = {i, j, k};
      nm(i, j, k) = (int)num;
      ++num;
    }
  };
  for (const auto& f : faces) {
    const auto &a = verts[f[0] - 1], &b = verts[f[1] - 1], &c = verts[f[2] - 1], &d = verts[f[3] - 1];
    if (a[0] == b[0] && a[0] == c[0] && a[0] == d[0]) {
      const int i = a[0];
      const int y0 = std::min({a[1], b[1], c[1], d[1]}), y1 = std::max({a[1], b[1], c[1], d[1]});
      const int z0 = std::min({a[2], b[2], c[2], d[2]}), z1 = std::max({a[2], b[2], c[2], d[2]});
      for (int j = y0; j < y1 + 1; ++j)
        for (int k = z0; k < z1 + 1; ++k) {
          touch(i, j, k);
          if (j != y1 && k != z1) XF(i, j, k) = 1;
        }
    } else if (a[1] == b[1] && a[1] == c[1] && a[1] == d[1]) {
      const int j = a[1];
      const int x0 = std::min({a[0], b[0], c[0], d[0]}), x1 = std::max({a[0], b[0], c[0], d[0]});
      const int z0 = std::min({a[2], b[2], c[2], d[2]}), z1 = std::max({a[2], b[2], c[2], d[2]});
      for (int i = x0; i < x1 + 1; ++i)
        for (int k = z0; k < z1 + 1; ++k) {
          touch(i, j, k);
          if (i != x1 && k != z1) YF(i, j, k) = 1;
        }
    } else if (a[2] == b[2] && a[2] == c[2] && a[2] == d[2]) {
      const int k = a[2];
      const int x0 = std::min({a[0], b[0], c[0], d[0]}), x1 = std::max({a[0], b[0], c[0], d[0]});
      const int y0 = std::min({a[1], b[1], c[1], d[1]}), y1 = std::max({a[1], b[1], c[1], d[1]});
      for (int i = x0; i < x1 + 1; ++i)
        for (int j = y0; j < y1 + 1; ++j) {
          touch(i, j, k);
          if (i != x1 && j != y1) ZF(i, j, k) = 1;
        }
    } else {
      throw std::logic_error("Face type cannot be determinted by vertices position.");
    }
  }
  std::vector<int> cnt((size_t)(nx - 1) * (ny - 1) * (nz - 1), 0);
  auto C = [&](int i, int j, int k) -> int& { return cnt[((size_t)i * (ny - 1) + j) * (nz - 1) + k]; };
  for (int j = 0; j < ny - 1; ++j)
    for (int k = 0; k < nz - 1; ++k)
      for (int lo = 0; lo < nx - 1; ++lo)
        if (XF(lo, j, k))
          for (int up = lo + 1; up < nx - 1; ++up)
            if (XF(up, j, k)) {
              for (int i = lo; i < up; ++i) C(i, j, k) += 1;
              lo = up;
              break;
            }
  for (int i = 0; i < nx - 1; ++i)
    for (int k = 0; k < nz - 1; ++k)
      for (int lo = 0; lo < ny - 1; ++lo)
        if (YF(i, lo, k))
          for (int up = lo + 1; up < ny - 1; ++up)
            if (YF(i, up, k)) {
              for (int j = lo; j < up; ++j) C(i, j, k) += 1;
              lo = up;
              break;
            }
  for (int i = 0; i < nx - 1; ++i)
    for (int j = 0; j < ny - 1; ++j)
      for (int lo = 0; lo < nz - 1; ++lo)
        if (ZF(i, j, lo))
          for (int up = lo + 1; up < nz - 1; ++up)
            if (ZF(i, j, up)) {
              for (int k = lo; k < up; ++k) C(i, j, k) += 1;
              lo = up;
              break;
            }
  for (int i = 0; i < nx - 1; ++i)
    for (int j = 0; j < ny - 1; ++j)
      for (int k = 0; k < nz - 1; ++k)
        if (C(i, j, k) == 3) od.cells.push_back({i, j, k});
  return od;
}

// ---------------------------------------------------------------- spacecraft
// Environment.cpp:291-293
bool Spacecraft::isValidCell(int i, int j, int k) const {
  return i >= 0 && i < nx + 1 && j >= 0 && j < ny + 1 && k >= 0 && k < nz + 1;
}

// Spacecraft.cpp:320-332
bool Spacecraft::isContaining(const Position& p) const {
  if (!defined) return false;
  if (isValidCell(p.i, p.j, p.k)) return cell(p.i, p.j, p.k);
  return false;
}

// Spacecraft.cpp:210-235 (the linear search becomes a dense lookup; one rank
// holds every node, so inner/glue/whole searches reduce to one table)
int Spacecraft::cmatNumber(int i, int j, int k) const {
  if (i < 0 || j < 0 || k < 0 || i >= nx + 2 || j >= ny + 2 || k >= nz + 2) return -1;
  return node_cmat[((size_t)i * (ny + 2) + j) * (nz + 2) + k];
}

// Spacecraft.cpp:461-504
bool Spacecraft::isXsurfacePoint(const Position& p, int sign) const {
  constexpr double eps = 1e-10;
  if (p.dx1 >= eps) return false;
  if (isValidCell(p.i, p.j, p.k) && isValidCell(p.i - 1, p.j, p.k)) {
    return sign > 0 ? isXsurfaceMinus(p.i, p.j, p.k) : isXsurfacePlus(p.i, p.j, p.k);
  }
  return isCmatNode(p.i, p.j, p.k) && isCmatNode(p.i, p.j + 1, p.k) && isCmatNode(p.i, p.j, p.k + 1) &&
         isCmatNode(p.i, p.j + 1, p.k + 1);
}
bool Spacecraft::isYsurfacePoint(const Position& p, int sign) const {
  constexpr double eps = 1e-10;
  if (p.dy1 >= eps) return false;
  if (isValidCell(p.i, p.j, p.k) && isValidCell(p.i, p.j - 1, p.k)) {
    return sign > 0 ? isYsurfaceMinus(p.i, p.j, p.k) : isYsurfacePlus(p.i, p.j, p.k);
  }
  return isCmatNode(p.i, p.j, p.k) && isCmatNode(p.i + 1, p.j, p.k) && isCmatNode(p.i, p.j, p.k + 1) &&
         isCmatNode(p.i + 1, p.j, p.k + 1);
}
bool Spacecraft::isZsurfacePoint(const Position& p, int sign) const {
  constexpr double eps = 1e-10;
  if (p.dz1 >= eps) return false;
  if (isValidCell(p.i, p.j, p.k) && isValidCell(p.i, p.j, p.k - 1)) {
    return sign > 0 ? isZsurfaceMinus(p.i, p.j, p.k) : isZsurfacePlus(p.i, p.j, p.k);
  }
  return isCmatNode(p.i, p.j, p.k) && isCmatNode(p.i + 1, p.j, p.k) && isCmatNode(p.i, p.j + 1, p.k) &&
         isCmatNode(p.i + 1, p.j + 1, p.k);
}

// Spacecraft.cpp:594-613 (sgn = +1) and 945-970 (sgn = -1, emission)
void Spacecraft::addX(const Position& p, int id, double q, double sgn) {
  double* t = &temporary_charge_map[(size_t)id * num_cmat];
#if REFCPU_MUTANT == 1  // planted mistake for the pin tests: dy1 and dy2 swapped
  t[cmatNumber(p.i, p.j, p.k)] += sgn * (q * p.dy1 * p.dz2);
  t[cmatNumber(p.i, p.j + 1, p.k)] += sgn * (q * p.dy2 * p.dz2);
  t[cmatNumber(p.i, p.j, p.k + 1)] += sgn * (q * p.dy1 * p.dz1);
  t[cmatNumber(p.i, p.j + 1, p.k + 1)] += sgn * (q * p.dy2 * p.dz1);
#else
  t[cmatNumber(p.i, p.j, p.k)] += sgn * (q * p.dy2 * p.dz2);
  t[cmatNumber(p.i, p.j + 1, p.k)] += sgn * (q * p.dy1 * p.dz2);
  t[cmatNumber(p.i, p.j, p.k + 1)] += sgn * (q * p.dy2 * p.dz1);
  t[cmatNumber(p.i, p.j + 1, p.k + 1)] += sgn * (q * p.dy1 * p.dz1);
#endif
}
void Spacecraft::addY(const Position& p, int id, double q, double sgn) {
  double* t = &temporary_charge_map[(size_t)id * num_cmat];
  t[cmatNumber(p.i, p.j, p.k)] += sgn * (q * p.dx2 * p.dz2);
  t[cmatNumber(p.i + 1, p.j, p.k)] += sgn * (q * p.dx1 * p.dz2);
  t[cmatNumber(p.i, p.j, p.k + 1)] += sgn * (q * p.dx2 * p.dz1);
  t[cmatNumber(p.i + 1, p.j, p.k + 1)] += sgn * (q * p.dx1 * p.dz1);
}
void Spacecraft::addZ(const Position& p, int id, double q, double sgn) {
  double* t = &temporary_charge_map[(size_t)id * num_cmat];
  t[cmatNumber(p.i, p.j, p.k)] += sgn * (q * p.dx2 * p.dy2);
  t[cmatNumber(p.i + 1, p.j, p.k)] += sgn * (q * p.dx1 * p.dy2);
  t[cmatNumber(p.i, p.j + 1, p.k)] += sgn * (q * p.dx2 * p.dy1);
  t[cmatNumber(p.i + 1, p.j + 1, p.k)] += sgn * (q * p.dx1 * p.dy1);
}

// particle.cpp:46-80
static double moveRatio(double d1, double d2, double v) {
  if (v == 0.0) return 0.0;
  return v > 0.0 ? d1 / std::fabs(v) : d2 / std::fabs(v);
}

// particle.cpp:82-204 — crossing points in crossing order (incl. the
// `mvx <= mvx` comparison at :175, which always puts x before z)
static std::vector<Position> computeCrossPoints(const Particle& p) {
  std::vector<Position> out;
  const Position pos(p.x, p.y, p.z);
  const Position old(p.x - p.vx, p.y - p.vy, p.z - p.vz);
  const double mvx = 1.0 - moveRatio(pos.dx1, pos.dx2, p.vx);
  const double mvy = 1.0 - moveRatio(pos.dy1, pos.dy2, p.vy);
  const double mvz = 1.0 - moveRatio(pos.dz1, pos.dz2, p.vz);
  int count = 0;
  Position xc, yc, zc;
  const bool mx = pos.i != old.i, my = pos.j != old.j, mz = pos.k != old.k;
  if (mx) {
    xc.setXYZ(p.vx > 0.0 ? std::ceil(old.x) : std::floor(old.x), old.y + p.vy * mvx, old.z + p.vz * mvx);
    ++count;
  }
  if (my) {
    yc.setXYZ(old.x + p.vx * mvy, p.vy > 0.0 ? std::ceil(old.y) : std::floor(old.y), old.z + p.vz * mvy);
    ++count;
  }
  if (mz) {
    zc.setXYZ(old.x + p.vx * mvz, old.y + p.vy * mvz, p.vz > 0.0 ? std::ceil(old.z) : std::floor(old.z));
    ++count;
  }
  if (count == 3) {
    if (mvx <= mvy) {
      if (mvx <= mvz) {
        out.push_back(xc);
        if (mvy <= mvz) {
          out.push_back(yc);
          out.push_back(zc);
        } else {
          out.push_back(zc);
          out.push_back(yc);
        }
      } else {
        out.push_back(zc);
        out.push_back(xc);
        out.push_back(yc);
      }
    } else {
      if (mvy <= mvz) {
        out.push_back(yc);
        if (mvx <= mvz) {
          out.push_back(xc);
          out.push_back(zc);
        } else {
          out.push_back(zc);
          out.push_back(xc);
        }
      } else {
        out.push_back(zc);
        out.push_back(yc);
        out.push_back(xc);
      }
    }
  } else if (count == 2) {
    if (mx) {
      if (my) {
        if (mvx <= mvy) {
          out.push_back(xc);
          out.push_back(yc);
        } else {
          out.push_back(yc);
          out.push_back(xc);
        }
      } else {
        out.push_back(xc);  // `mvx <= mvx` is always true
        out.push_back(zc);
      }
    } else {
      if (mvy <= mvz) {
        out.push_back(yc);
        out.push_back(zc);
      } else {
        out.push_back(zc);
        out.push_back(yc);
      }
    }
  } else if (count == 1) {
    out.push_back(mx ? xc : (my ? yc : zc));
  }
  return out;
}

// Spacecraft.cpp:512-592 (no secondary emission, no force computation)
void Spacecraft::distributeInnerParticleCharge(Particle& p, const std::vector<Species>& sp) {
  const Position pos(p.x, p.y, p.z);
  if (!isContaining(pos)) return;
  const int id = p.typeId;
  const double q = sp[id].chargeOfSuperParticle();
  const int is = p.vx > 0.0 ? 1 : -1, js = p.vy > 0.0 ? 1 : -1, ks = p.vz > 0.0 ? 1 : -1;
  for (const auto& c : computeCrossPoints(p)) {
    if (isXsurfacePoint(c, is)) {
      addX(c, id, q, 1.0);
      break;
    } else if (isYsurfacePoint(c, js)) {
      addY(c, id, q, 1.0);
      break;
    } else if (isZsurfacePoint(c, ks)) {
      addZ(c, id, q, 1.0);
      break;
    }
  }
  current[id] += q;
  p.isValid = 0;
}

// Spacecraft.cpp:615-626
void Spacecraft::sumWholeCharge(int S) {
  for (unsigned c = 0; c < num_cmat; ++c)
    for (int s = 0; s < S; ++s) {
      charge_map[(size_t)s * num_cmat + c] += temporary_charge_map[(size_t)s * num_cmat + c];
      temporary_charge_map[(size_t)s * num_cmat + c] = 0.0;
    }
}

// Spacecraft.cpp:248-259
double Spacecraft::getTotalCharge(int S) const {
  double q = 0.0;
  for (unsigned c = 0; c < num_cmat; ++c)
    for (int s = 0; s < S; ++s) q += charge_map[(size_t)s * num_cmat + c];
  return q;
}

// include/Spacecraft/Cmatrix.hpp:36-44
void Spacecraft::updateTotalCmatValue() {
  double r = 0.0;
  for (unsigned col = 0; col < num_cmat; ++col)
    for (unsigned row = 0; row < num_cmat; ++row) r += cmat[(size_t)col * num_cmat + row];
  total_cmat_value = r;
}

// Spacecraft.cpp:628-639 + 708-758 (perfect conductor)
void Spacecraft::redistributeCharge(std::vector<double>& rho0, const std::vector<double>& phi, int cy, int cz, int S) {
  total_charge = getTotalCharge(S);
  auto P = [&](const Position& q) { return phi[((size_t)q.i * cy + q.j) * cz + q.k]; };
  double ctp = 0.0;
  for (const auto& nd : cmat_rel) {
    const unsigned j = nd.first;
    for (unsigned i = 0; i < num_cmat; ++i) ctp += cmat[(size_t)i * num_cmat + j] * P(nd.second);
  }
  potential = is_potential_fixed ? fixed_potential : (ctp + total_charge) / total_cmat_value;
  for (unsigned i = 0; i < num_cmat; ++i) {
    double d = 0.0;
    for (unsigned j = 0; j < num_cmat; ++j) d += cmat[(size_t)i * num_cmat + j] * (potential - P(cmat_rel.at(j)));
    const Position& t = cmat_rel.at(i);
    rho0[((size_t)t.i * cy + t.j) * cz + t.k] += d;
  }
}

// ---------------------------------------------------------------- grid
Grid::Grid(const Params& p) : prm(p) {
  u.set(p.dx, p.dt);
  nx = p.nx;  // one rank: cell_x = nx / proc_x with proc_x = 1 (Initialize.cpp:214-218)
  ny = p.ny;
  nz = p.nz;
  cx = nx + 2;
  cy = ny + 2;
  cz = nz + 2;
  boundary = p.boundary;
  for (char c : boundary)
    if (c != 'D' && c != 'N') throw std::invalid_argument("oracle supports D and N boundaries only");
  post_loop = p.post_loop > 0 ? p.post_loop : 1000;
  // Initialize.cpp:306-327 (static B in nT)
  const double bn = u.normBfield(1.0);
  for (int a = 0; a < 3; ++a) {
    b[a] = p.static_b_nT[a] * (bn * 1e-9);
    shine[a] = p.shine[a];
  }
  // Initialize.cpp:329-421 (species arrive already in reference id order)
  S = p.nspecies;
  sp.resize(S);
  for (int s = 0; s < S; ++s) {
    const SpeciesParams& q = p.sp[s];
    Species& t = sp[s];
    t.setId(s);
    t.type = q.type;
    t.mass = q.mass;
    t.charge = q.charge;
    t.temperature = q.temperature;
    t.density = q.density;
    t.ppc = q.ppc;
    for (int a = 0; a < 3; ++a) t.drift[a] = q.drift_kms[a] * (1e3 * u.normVelocity(1.0));
    for (int a = 0; a < 6; ++a) t.inj[a] = q.injection_axis[a] != 0;
    t.size = static_cast<int>(std::pow(p.dx, 3) * t.density / static_cast<double>(t.ppc));  // ParticleType.cpp:33-36
    t.current_density = u.normCurrentDensity(q.current_density);
    t.mt_azimuth.seed(586936);         // random_distribution.hpp:19 (rank 0)
    t.mt_depression.seed(8785997666ULL);
  }
  particles.resize(S);
  inj_residual.assign(S, std::vector<double>(6, 0.0));
  const size_t n3 = (size_t)cx * cy * cz;
  phi.assign(n3, 0.0);
  poisson_residual.assign(n3, 0.0);
  poisson_error.assign(n3, 0.0);
  rho.assign(S + 1, std::vector<double>(n3, 0.0));
  density.assign(S, std::vector<double>((size_t)(cx - 1) * (cy - 1) * (cz - 1), 0.0));
  ex.assign((size_t)(cx - 1) * cy * cz, 0.0);
  ey.assign((size_t)cx * (cy - 1) * cz, 0.0);
  ez.assign((size_t)cx * cy * (cz - 1), 0.0);
  exref.assign(n3, 0.0);
  eyref.assign(n3, 0.0);
  ezref.assign(n3, 0.0);

  // RootGridObject.cpp:6-75 + Spacecraft.cpp:16-152 (one rank)
  for (int o = 0; o < p.nobjects; ++o) {
    const ObjectParams& op = p.obj[o];
    ObjData od = loadObjFile(op.file_name, nx, ny, nz);
    Spacecraft sc;
    sc.nx = nx;
    sc.ny = ny;
    sc.nz = nz;
    sc.num_cmat = (unsigned)od.nodes.size();
    sc.is_potential_fixed = op.is_potential_fixed != 0;
    sc.fixed_potential = u.normPotential(op.fixed_potential);
    sc.initial_potential_offset = u.normPotential(op.initial_potential_offset);
    sc.emit_species = op.emit_species;
    sc.node_cmat.assign((size_t)cx * cy * cz, -1);
    for (const auto& nd : od.nodes) {
      const int i = nd.second[0], j = nd.second[1], k = nd.second[2];
      // isInnerNode (grid.hpp:408-416): 0 <= i < cell_x on one rank
      if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
        sc.defined = true;
        sc.cmat_rel[nd.first] = Position::fromIJK(i + 1, j + 1, k + 1);
        sc.node_cmat[((size_t)(i + 1) * cy + (j + 1)) * cz + (k + 1)] = (int)nd.first;
      } else if (i >= -1 && i <= nx && j >= -1 && j <= ny && k >= -1 && k <= nz) {
        sc.node_cmat[((size_t)(i + 1) * cy + (j + 1)) * cz + (k + 1)] = (int)nd.first;
      }
    }
    sc.cell_map.assign((size_t)(nx + 1) * (ny + 1) * (nz + 1), 0);
    for (const auto& c : od.cells) {
      const int i = c[0] + 1, j = c[1] + 1, k = c[2] + 1;
      if (i >= 0 && i <= nx && j >= 0 && j <= ny && k >= 0 && k <= nz)
        sc.cell_map[((size_t)i * (ny + 1) + j) * (nz + 1) + k] = 1;
    }
    sc.cmat.assign((size_t)sc.num_cmat * sc.num_cmat, 0.0);
    sc.charge_map.assign((size_t)S * sc.num_cmat, 0.0);
    sc.temporary_charge_map.assign((size_t)S * sc.num_cmat, 0.0);
    sc.current.assign(S, 0.0);
    objects.push_back(std::move(sc));
  }
}

// RootGrid.cpp:45-79
void Grid::loadInitialParticles() {
  double mx = nx, my = ny, mz = nz;
  if (isBoundary(0, 1)) mx -= 1.0;
  if (isBoundary(1, 1)) my -= 1.0;
  if (isBoundary(2, 1)) mz -= 1.0;
  for (int s = 0; s < S; ++s) {
    if (sp[s].type != "ambient") continue;
    // ParticleType.cpp:39-47
    const size_t cxn = isBoundary(0, 1) ? nx - 1 : nx, cyn = isBoundary(1, 1) ? ny - 1 : ny,
                 czn = isBoundary(2, 1) ? nz - 1 : nz;
    const size_t total = cxn * cyn * czn * sp[s].ppc;
    particles[s].reserve(total * 2);
    for (size_t n = 0; n < total; ++n) {
      Position ps = genPosition(sp[s], 0.0, mx, 0.0, my, 0.0, mz);
      double v[3];
      genVelocity(sp[s], u, v);
      Particle pt{ps.x, ps.y, ps.z, v[0], v[1], v[2], s, 1};
      for (const auto& o : objects)
        if (o.defined && o.isContaining(Position(pt.x, pt.y, pt.z))) pt.isValid = 0;
      if (pt.isValid) particles[s].push_back(pt);
    }
  }
}

// Grid.cpp:23-36
void Grid::initialize() {
  updateRho();
  solvePoisson();
  updateEfield();
}

// RootGrid.cpp:82-143 (timestep advanced by the caller, main.cpp:33)
void Grid::mainLoopES() {
  for (auto& o : objects) std::fill(o.current.begin(), o.current.end(), 0.0);  // resetObjects
  updateParticleVelocityES();
  emitParticlesFromObjects();
  updateParticlePositionES();
  injectParticlesFromBoundary();
  updateDensity();
  updateRho();
  solvePoisson();
  updateEfield();
  if (timestep % 50 == 0) removeInvalidParticles();
}

// GridParticleTreatmentES.cpp:13-88 (level 0: time_factor = 1)
void Grid::updateParticleVelocityES() {
  const double tf = 1.0 / std::pow(2, 0);
  const double bx = b[0], by = b[1], bz = b[2];
  for (int s = 0; s < S; ++s) {
    const double qm = 0.5 * sp[s].charge / sp[s].mass;
    auto& pa = particles[s];
    const long n = (long)pa.size();
#pragma omp parallel for
    for (long q = 0; q < n; ++q) {
      Particle& p = pa[q];
      if (!p.isValid) continue;
      const Position pos(p.x, p.y, p.z);
      const int i = pos.i, j = pos.j, k = pos.k;
      const double v1 = qm * pos.dx2 * pos.dy2 * pos.dz2;
      const double v2 = qm * pos.dx1 * pos.dy2 * pos.dz2;
      const double v3 = qm * pos.dx2 * pos.dy1 * pos.dz2;
      const double v4 = qm * pos.dx1 * pos.dy1 * pos.dz2;
      const double v5 = qm * pos.dx2 * pos.dy2 * pos.dz1;
      const double v6 = qm * pos.dx1 * pos.dy2 * pos.dz1;
      const double v7 = qm * pos.dx2 * pos.dy1 * pos.dz1;
      const double v8 = qm * pos.dx1 * pos.dy1 * pos.dz1;
      auto g = [&](const std::vector<double>& E) {
        return v1 * E[id3(i, j, k)] + v2 * E[id3(i + 1, j, k)] + v3 * E[id3(i, j + 1, k)] +
               v4 * E[id3(i + 1, j + 1, k)] + v5 * E[id3(i, j, k + 1)] + v6 * E[id3(i + 1, j, k + 1)] +
               v7 * E[id3(i, j + 1, k + 1)] + v8 * E[id3(i + 1, j + 1, k + 1)];
      };
      const double exv = g(exref), eyv = g(eyref), ezv = g(ezref);
      const double boris = 2.0 / (1.0 + (bx * bx + by * by + bz * bz));
      const double vx1 = p.vx + exv, vy1 = p.vy + eyv, vz1 = p.vz + ezv;
      const double vxt = vx1 + vy1 * bz - vz1 * by;
      const double vyt = vy1 + vz1 * bx - vx1 * bz;
      const double vzt = vz1 + vx1 * by - vy1 * bx;
      p.vx = vx1 + (exv + boris * (vyt * bz - vzt * by)) * tf;
      p.vy = vy1 + (eyv + boris * (vzt * bx - vxt * bz)) * tf;
      p.vz = vz1 + (ezv + boris * (vxt * by - vyt * bx)) * tf;
    }
  }
}

// GridParticleTreatmentES.cpp:90-113 + grid.hpp:309-395 on one rank where
// every face is a physical (non-periodic) boundary: leaving means invalid.
void Grid::updateParticlePositionES() {
  const double slx = dx * nx, sly = dx * ny, slz = dx * nz;
  for (int s = 0; s < S; ++s) {
    auto& pa = particles[s];
    const long n = (long)pa.size();
#pragma omp parallel for
    for (long q = 0; q < n; ++q) {
      Particle& p = pa[q];
      if (!p.isValid) continue;
      p.x += p.vx;
      p.y += p.vy;
      p.z += p.vz;
#if REFCPU_MUTANT == 5  // planted mistake: a particle exactly on the upper x face is absorbed
      if (p.x < 0.0 || p.x >= (slx - dx)) p.isValid = 0;
#else
      if (p.x < 0.0 || p.x > (slx - dx)) p.isValid = 0;
#endif
      if (p.y < 0.0 || p.y > (sly - dx)) p.isValid = 0;
      if (p.z < 0.0 || p.z > (slz - dx)) p.isValid = 0;
    }
  }
}

// RootGridParticle.cpp:3-130 (node sizes RootGrid.cpp:192-204)
void Grid::injectParticlesFromBoundary() {
  double mx = nx, my = ny, mz = nz;
  if (isBoundary(0, 1)) mx -= 1.0;
  if (isBoundary(1, 1)) my -= 1.0;
  if (isBoundary(2, 1)) mz -= 1.0;
  const int xn = nx, yn = ny, zn = nz;  // getXNodeSize etc. with non-periodic faces
  for (int s = 0; s < S; ++s) {
    if (sp[s].type != "ambient") continue;
    Species& t = sp[s];
    const std::vector<double> flux = calcFlux(t, u);
    for (int face = 0; face < 6; ++face) {
      const int axis = face / 2, side = face % 2;
      if (!isBoundary(axis, side)) continue;
      double area;
      if (axis == 0) area = dx * dx * (zn - 1) * (yn - 1);
      else if (axis == 1) area = dx * dx * (zn - 1) * (xn - 1);
      else area = dx * dx * (xn - 1) * (yn - 1);
      const double nd = dt * area * flux[face];
      const int num = static_cast<int>(std::floor(nd + inj_residual[s][face]));
      inj_residual[s][face] += nd - num;
      for (int q = 0; q < num; ++q) {
        double v[3];
        genInjectionVelocity(t, u, axis, side, v);
        double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {mx, my, mz};
        if (side == 0) {
          hi[axis] = v[axis] * dt;
        } else {
          lo[axis] = hi[axis] + v[axis] * dt;
        }
        Position ps = genPosition(t, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
        particles[s].push_back(Particle{ps.x, ps.y, ps.z, v[0], v[1], v[2], s, 1});
      }
    }
  }
}

// RootGridObject.cpp:175-181 → Spacecraft.cpp:829-900 (photoelectron branch),
// PhotoElectronParticleType.cpp:9-106, particle.cpp:206-255
void Grid::emitParticlesFromObjects() {
  for (auto& o : objects) {
    if (!o.defined || o.emit_species < 0) continue;
    Species& t = sp[o.emit_species];
    const int id = o.emit_species;
    int ax = -1;
    for (int a = 0; a < 3; ++a)
      if (std::fabs(shine[a]) == 1.0) {
        ax = a;
        break;
      }
    if (ax < 0) throw std::invalid_argument("shine_vector must be an axis unit vector");
    const int ssign = shine[ax] > 0.0 ? 1 : -1;
    const double q = t.chargeOfSuperParticle();
    const double amount = (dx * dx) * (t.current_density / (std::fabs(t.charge) * t.size));
    const double ev[3] = {-shine[0], -shine[1], -shine[2]};
    const double vth = t.thermalVelocity(u);
    std::uniform_real_distribution<> unit(0.0, 1.0);
    for (const auto& nd : o.cmat_rel) {
      const Position& pos = nd.second;
      const bool surf = ax == 0 ? o.isXsurfacePoint(pos, ssign)
                                : (ax == 1 ? o.isYsurfacePoint(pos, ssign) : o.isZsurfacePoint(pos, ssign));
      if (!surf) continue;
      for (int n = 0; n < amount; ++n) {
        // generateNewVelocity: depression first, then azimuth
        const double dep = std::asin(2.0 * unit(t.mt_depression) - 1.0);
        const double azi = 2.0 * M_PI * unit(t.mt_azimuth);
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
        // generateNewPosition
        std::uniform_real_distribution<> dt_(0.0, 1.0), da(0.0, 1.0), db(0.0, 1.0);
        const double rt = dt_(t.mt_x);
        double px, py, pz;
        if (ax == 0) {
          const double ry = da(t.mt_y), rz = db(t.mt_z);
          px = pos.x + (rt - 1.0) * v[0];
          py = pos.y + (rt - 1.0) * v[1] + ry;
          pz = pos.z + (rt - 1.0) * v[2] + rz;
        } else if (ax == 1) {
          const double rx = da(t.mt_x), rz = db(t.mt_z);
          px = pos.x + (rt - 1.0) * v[0] + rx;
          py = pos.y + (rt - 1.0) * v[1];
          pz = pos.z + (rt - 1.0) * v[2] + rz;
        } else {
          const double rx = da(t.mt_x), ry = db(t.mt_y);
          px = pos.x + (rt - 1.0) * v[0] + rx;
          py = pos.y + (rt - 1.0) * v[1] + ry;
          pz = pos.z + (rt - 1.0) * v[2];
        }
        Particle p{px, py, pz, v[0], v[1], v[2], id, 1};
        // getNext{X,Y,Z}CrossPoint then subtractChargeOfParticleFrom{X,Y,Z}surface
        Position c(p.x, p.y, p.z);
        const double pv[3] = {p.vx, p.vy, p.vz};
        if (ax == 0) {
          if (pv[0] > 0.0) c.setXYZ(std::ceil(c.x), c.y + pv[1] * (c.dx2 / pv[0]), c.z + pv[2] * (c.dx2 / pv[0]));
          else if (pv[0] < 0.0)
            c.setXYZ(std::floor(c.x), c.y + pv[1] * (c.dx1 / (-pv[0])), c.z + pv[2] * (c.dx1 / (-pv[0])));
          o.addX(c, id, q, -1.0);
        } else if (ax == 1) {
          if (pv[1] > 0.0) c.setXYZ(c.x + pv[0] * (c.dy2 / pv[1]), std::ceil(c.y), c.z + pv[2] * (c.dy2 / pv[1]));
          else if (pv[1] < 0.0)
            c.setXYZ(c.x + pv[0] * (c.dy1 / (-pv[1])), std::floor(c.y), c.z + pv[2] * (c.dy1 / (-pv[1])));
          o.addY(c, id, q, -1.0);
        } else {
          if (pv[2] > 0.0) c.setXYZ(c.x + pv[0] * (c.dz2 / pv[2]), c.y + pv[1] * (c.dz2 / pv[2]), std::ceil(c.z));
          else if (pv[2] < 0.0)
            c.setXYZ(c.x + pv[0] * (c.dz1 / (-pv[2])), c.y + pv[1] * (c.dz1 / (-pv[2])), std::floor(c.z));
          o.addZ(c, id, q, -1.0);
        }
        o.current[id] -= q;
        particles[id].push_back(p);
      }
    }
  }
}

// RootGridField.cpp:1129-1150
void Grid::updateDensity() {
  const int dy_ = cy - 1, dz_ = cz - 1;
  for (int s = 0; s < S; ++s) {
    std::fill(density[s].begin(), density[s].end(), 0.0);
    const double size = sp[s].size;
    for (const auto& p : particles[s]) {
      if (!p.isValid) continue;
      const Position pos(p.x, p.y, p.z);
#if REFCPU_MUTANT == 4  // planted mistake: the cell index without the glue offset along x
      density[s][((size_t)(pos.i - 1) * dy_ + pos.j) * dz_ + pos.k] += size;
#else
      density[s][((size_t)pos.i * dy_ + pos.j) * dz_ + pos.k] += size;
#endif
    }
  }
}

// RootGridField.cpp:1161-1214: zero rho, per species containment + CIC, sum
void Grid::depositOnly() {
  for (auto& r : rho) std::fill(r.begin(), r.end(), 0.0);
#pragma omp parallel for
  for (int s = 0; s < S; ++s) {
    const double q = sp[s].chargeOfSuperParticle();
    auto& R = rho[s + 1];
    for (auto& p : particles[s]) {
      if (!p.isValid) continue;
      for (auto& o : objects) o.distributeInnerParticleCharge(p, sp);
      if (!p.isValid) continue;
      const Position pos(p.x, p.y, p.z);
      const int i = pos.i, j = pos.j, k = pos.k;
      R[id3(i, j, k)] += pos.dx2 * pos.dy2 * pos.dz2 * q;
      R[id3(i + 1, j, k)] += pos.dx1 * pos.dy2 * pos.dz2 * q;
      R[id3(i, j + 1, k)] += pos.dx2 * pos.dy1 * pos.dz2 * q;
      R[id3(i + 1, j + 1, k)] += pos.dx1 * pos.dy1 * pos.dz2 * q;
      R[id3(i, j, k + 1)] += pos.dx2 * pos.dy2 * pos.dz1 * q;
      R[id3(i + 1, j, k + 1)] += pos.dx1 * pos.dy2 * pos.dz1 * q;
      R[id3(i, j + 1, k + 1)] += pos.dx2 * pos.dy1 * pos.dz1 * q;
      R[id3(i + 1, j + 1, k + 1)] += pos.dx1 * pos.dy1 * pos.dz1 * q;
    }
  }
  for (auto& o : objects)
    if (o.defined) o.sumWholeCharge(S);
  for (int s = 1; s < S + 1; ++s)
    for (size_t n = 0; n < rho[0].size(); ++n) rho[0][n] += rho[s][n];
}

// RootGridField.cpp:1153-1252
void Grid::updateRho() {
  depositOnly();
  if (!objects.empty()) {
    solvePoisson();
    for (auto& o : objects)
      if (o.defined) o.redistributeCharge(rho[0], phi, cy, cz, S);
  }
}

// RootGridField.cpp:41-63 (no AMR children)
void Grid::solvePoisson() { last_sor_iterations = solvePoissonPSOR(post_loop); }

// RootGridField.cpp:65-103
int Grid::solvePoissonPSOR(int loopnum) {
  constexpr double required = 1.0e-7;
  poisson_error = phi;
  int loop = 1;
  for (; loop <= loopnum; ++loop) {
    doPSOR();
    if ((loop % 10 == 0) && (checkPhiResidual() < required)) break;
  }
  setBoundaryConditionPhi();
  for (int i = 1; i < cx - 1; ++i)
    for (int j = 1; j < cy - 1; ++j)
      for (int k = 1; k < cz - 1; ++k) poisson_error[id3(i, j, k)] = phi[id3(i, j, k)] - poisson_error[id3(i, j, k)];
  return loop - 1;
}

// The reference SOR (same sweeps) stopped on this project's criterion
// ||b - A phi||_2 / ||b||_2 <= tol, checked every 10 sweeps, for the CPU baseline.
int Grid::solvePoissonToRelres(double tol, int maxloop) {
  int loop = 1;
  for (; loop <= maxloop; ++loop) {
    doPSOR();
    if ((loop % 10 == 0) && (relres() <= tol)) break;
  }
  setBoundaryConditionPhi();
  return std::min(loop, maxloop);
}

// RootGridField.cpp:105-264 on one rank with physical faces: every partial
// call but the interior one is empty; zebra over i, lexicographic GS in j,k.
void Grid::doPSOR() {
  const int minr = std::min({prm.nx, prm.ny, prm.nz});
  const double omega = 2.0 / (1.0 + std::sin(M_PI / minr));
  const double rc = std::pow(dx, 2) / u.eps0;
  // The interior call is doSORPartial(2, cx-2, 2, cy-2, 2, cz-2): an end on a
  // physical face is trimmed by one, so the sweep covers nodes 2..n-1.
  const int i0 = 2, j0 = 2, k0 = 2;
  const int i1 = isBoundary(0, 1) ? cx - 3 : cx - 2;
  const int j1 = isBoundary(1, 1) ? cy - 3 : cy - 2;
  const int k1 = isBoundary(2, 1) ? cz - 3 : cz - 2;
  const double* R = rho[0].data();
  double* P = phi.data();
  auto upd = [&](int i) {
    for (int j = j0; j <= j1; ++j)
      for (int k = k0; k <= k1; ++k) {
        const size_t c = id3(i, j, k);
        P[c] = (1.0 - omega) * P[c] + omega * (P[id3(i + 1, j, k)] + P[id3(i - 1, j, k)] + P[id3(i, j + 1, k)] +
                                               P[id3(i, j - 1, k)] + P[id3(i, j, k + 1)] + P[id3(i, j, k - 1)] +
                                               rc * R[c]) /
                                          6.0;
      }
  };
  const int isz = i1 - i0 + 1;
  if (isz > 2) {
#pragma omp parallel for
    for (int i = i0; i <= i1; i += 2) upd(i);
#pragma omp parallel for
    for (int i = i0 + 1; i <= i1; i += 2) upd(i);
  } else {
    for (int i = i0; i <= i1; ++i) upd(i);
  }
}

// RootGridField.cpp:267-317
double Grid::checkPhiResidual() {
  double res = 0.0;
  const double ne = u.eps0;
#pragma omp parallel for reduction(max : res)
  for (int i = 1; i < cx - 1; ++i) {
    if (!((i != 1 || !isBoundary(0, 0)) && (i != cx - 2 || !isBoundary(0, 1)))) continue;
    for (int j = 1; j < cy - 1; ++j) {
      if (!((j != 1 || !isBoundary(1, 0)) && (j != cy - 2 || !isBoundary(1, 1)))) continue;
      for (int k = 1; k < cz - 1; ++k) {
        if (!((k != 1 || !isBoundary(2, 0)) && (k != cz - 2 || !isBoundary(2, 1)))) continue;
        const double src = rho[0][id3(i, j, k)] / ne;
        if (std::fabs(src) > 0.0) {
          const double t = (phi[id3(i - 1, j, k)] + phi[id3(i + 1, j, k)] + phi[id3(i, j - 1, k)] +
                            phi[id3(i, j + 1, k)] + phi[id3(i, j, k - 1)] + phi[id3(i, j, k + 1)] -
                            6.0 * phi[id3(i, j, k)]) +
                           src;
          poisson_residual[id3(i, j, k)] = t;
          if (std::fabs(t / src) > res) res = std::fabs(t / src);
        }
      }
    }
  }
  return res;
}

double Grid::relres() const {
  const double rc = std::pow(dx, 2) / u.eps0;
  double rr = 0.0, bb = 0.0;
#pragma omp parallel for reduction(+ : rr, bb)
  for (int i = 2; i < cx - 2; ++i)
    for (int j = 2; j < cy - 2; ++j)
      for (int k = 2; k < cz - 2; ++k) {
        const double bv = rc * rho[0][id3(i, j, k)];
        const double ap = 6.0 * phi[id3(i, j, k)] - (phi[id3(i - 1, j, k)] + phi[id3(i + 1, j, k)] +
                                                     phi[id3(i, j - 1, k)] + phi[id3(i, j + 1, k)] +
                                                     phi[id3(i, j, k - 1)] + phi[id3(i, j, k + 1)]);
        rr += (bv - ap) * (bv - ap);
        bb += bv * bv;
      }
  return bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

// Field.cpp:10-96 (one rank: every face is on the edge)
void Grid::setBoundaryConditionPhi() {
  for (int axis = 0; axis < 3; ++axis)
    for (int side = 0; side < 2; ++side) {
      const char c = bc(axis, side);
      const int nmax = axis == 0 ? cx : (axis == 1 ? cy : cz);
      const int bi = side == 0 ? 1 : nmax - 2;
      const int ni = side == 0 ? 2 : nmax - 3;
      if (c == 'P') continue;
      for (int a = 1; a < (axis == 0 ? cy : cx) - 1; ++a)
        for (int bb = 1; bb < (axis == 2 ? cy : cz) - 1; ++bb) {
          size_t t, src;
          if (axis == 0) {
            t = id3(bi, a, bb);
            src = id3(ni, a, bb);
          } else if (axis == 1) {
            t = id3(a, bi, bb);
            src = id3(a, ni, bb);
          } else {
            t = id3(a, bb, bi);
            src = id3(a, bb, ni);
          }
          phi[t] = (c == 'D') ? 0.0 : phi[src];
        }
    }
}

// RootGridField.cpp:456-486
void Grid::updateEfield() {
  const double pdx = 1.0 / dx;
  for (int i = 0; i < cx; ++i)
    for (int j = 0; j < cy; ++j)
      for (int k = 0; k < cz; ++k) {
        if (i < cx - 1) ex[((size_t)i * cy + j) * cz + k] = (phi[id3(i, j, k)] - phi[id3(i + 1, j, k)]) * pdx;
        if (j < cy - 1) ey[((size_t)i * (cy - 1) + j) * cz + k] = (phi[id3(i, j, k)] - phi[id3(i, j + 1, k)]) * pdx;
        if (k < cz - 1) ez[((size_t)i * cy + j) * (cz - 1) + k] = (phi[id3(i, j, k)] - phi[id3(i, j, k + 1)]) * pdx;
      }
  updateReferenceEfield();
}

// RootGridField.cpp:489-582
void Grid::updateReferenceEfield() {
  auto EX = [&](int i, int j, int k) { return ex[((size_t)i * cy + j) * cz + k]; };
  auto EY = [&](int i, int j, int k) { return ey[((size_t)i * (cy - 1) + j) * cz + k]; };
  auto EZ = [&](int i, int j, int k) { return ez[((size_t)i * cy + j) * (cz - 1) + k]; };
  auto P = [&](int i, int j, int k) { return phi[id3(i, j, k)]; };
  for (int i = 1; i < cx - 1; ++i)
    for (int j = 1; j < cy - 1; ++j)
      for (int k = 1; k < cz - 1; ++k) {
        exref[id3(i, j, k)] = 0.5 * (EX(i - 1, j, k) + EX(i, j, k));
        eyref[id3(i, j, k)] = 0.5 * (EY(i, j - 1, k) + EY(i, j, k));
        ezref[id3(i, j, k)] = 0.5 * (EZ(i, j, k - 1) + EZ(i, j, k));
      }
  if (isBoundary(0, 0))
    for (int j = 1; j < cy - 1; ++j)
      for (int k = 1; k < cz - 1; ++k) {
        exref[id3(0, j, k)] = 0.0;
        exref[id3(1, j, k)] = 0.5 * (P(3, j, k) - 4.0 * P(2, j, k) + 3.0 * P(1, j, k));
      }
  if (isBoundary(0, 1))
    for (int j = 1; j < cy - 1; ++j)
      for (int k = 1; k < cz - 1; ++k) {
        exref[id3(cx - 1, j, k)] = 0.0;
        exref[id3(cx - 2, j, k)] = 0.5 * (-P(cx - 4, j, k) + 4.0 * P(cx - 3, j, k) - 3.0 * P(cx - 2, j, k));
      }
  if (isBoundary(1, 0))
    for (int i = 1; i < cx - 1; ++i)
      for (int k = 1; k < cz - 1; ++k) {
        eyref[id3(i, 0, k)] = 0.0;
        eyref[id3(i, 1, k)] = 0.5 * (P(i, 3, k) - 4.0 * P(i, 2, k) + 3.0 * P(i, 1, k));
      }
  if (isBoundary(1, 1))
    for (int i = 1; i < cx - 1; ++i)
      for (int k = 1; k < cz - 1; ++k) {
        eyref[id3(i, cy - 1, k)] = 0.0;
        eyref[id3(i, cy - 2, k)] = 0.5 * (-P(i, cy - 4, k) + 4.0 * P(i, cy - 3, k) - 3.0 * P(i, cy - 2, k));
      }
  if (isBoundary(2, 0))
    for (int i = 1; i < cx - 1; ++i)
      for (int j = 1; j < cy - 1; ++j) {
        ezref[id3(i, j, 0)] = 0.0;
        ezref[id3(i, j, 1)] = 0.5 * (P(i, j, 3) - 4.0 * P(i, j, 2) + 3.0 * P(i, j, 1));
      }
  if (isBoundary(2, 1))
    for (int i = 1; i < cx - 1; ++i)
      for (int j = 1; j < cy - 1; ++j) {
        ezref[id3(i, j, cz - 1)] = 0.0;
        ezref[id3(i, j, cz - 2)] = 0.5 * (-P(i, j, cz - 4) + 4.0 * P(i, j, cz - 3) - 3.0 * P(i, j, cz - 2));
      }
}

// GridParticleTreatment.cpp:93-102
void Grid::removeInvalidParticles() {
  for (auto& pa : particles)
    pa.erase(std::remove_if(pa.begin(), pa.end(), [](const Particle& p) { return p.isValid == 0; }), pa.end());
}

// GridParticleTreatment.cpp:104-118
double Grid::particleEnergy() const {
  double res = 0.0;
  for (int s = 0; s < S; ++s) {
    double e = 0.0;
    for (const auto& p : particles[s])
#if REFCPU_MUTANT == 2  // planted mistake: the vz term dropped
      if (p.isValid) e += (p.vx * p.vx + p.vy * p.vy);
#else
      if (p.isValid) e += (p.vx * p.vx + p.vy * p.vy + p.vz * p.vz);
#endif
    res += 0.5 * sp[s].mass * e * sp[s].size;
  }
  return res;
}

// Field.cpp:194-213
double Grid::efieldEnergy() const {
  double e = 0.0;
  for (int i = 1; i < cx - 1; ++i)
    for (int j = 1; j < cy - 1; ++j)
      for (int k = 1; k < cz - 1; ++k) {
        if (i < cx - 2) e += std::pow(ex[((size_t)i * cy + j) * cz + k], 2);
        if (j < cy - 2) e += std::pow(ey[((size_t)i * (cy - 1) + j) * cz + k], 2);
#if REFCPU_MUTANT != 3  // planted mistake (mutant 3): the ez edges dropped
        if (k < cz - 2) e += std::pow(ez[((size_t)i * cy + j) * (cz - 1) + k], 2);
#endif
      }
  return 0.5 * u.eps0 * e;
}

// FP64 inverse by Gauss-Jordan with partial pivoting, in place of the
// reference's Eigen MatrixXd::inverse (include/Spacecraft/Cmatrix.hpp:31-34).
static void invertInPlace(std::vector<double>& a, unsigned n) {
  std::vector<double> inv((size_t)n * n, 0.0);
  for (unsigned i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (unsigned c = 0; c < n; ++c) {
    unsigned piv = c;
    double best = std::fabs(a[(size_t)c * n + c]);
    for (unsigned r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > best) {
        best = std::fabs(a[(size_t)r * n + c]);
        piv = r;
      }
    if (best == 0.0) throw std::runtime_error("singular capacitance matrix");
    if (piv != c)
      for (unsigned q = 0; q < n; ++q) {
        std::swap(a[(size_t)c * n + q], a[(size_t)piv * n + q]);
        std::swap(inv[(size_t)c * n + q], inv[(size_t)piv * n + q]);
      }
    const double d = a[(size_t)c * n + c];
    for (unsigned q = 0; q < n; ++q) {
      a[(size_t)c * n + q] /= d;
      inv[(size_t)c * n + q] /= d;
    }
    for (unsigned r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (unsigned q = 0; q < n; ++q) {
        a[(size_t)r * n + q] -= f * a[(size_t)c * n + q];
        inv[(size_t)r * n + q] -= f * inv[(size_t)c * n + q];
      }
    }
  }
  a.swap(inv);
}

// RootGridObject.cpp:77-148: one unit-charge solve per surface node, then
// C = B^-1 (element (col,row) = phi at row for the charge at col).
void Grid::buildCmatrix() {
  for (auto& o : objects) {
    for (auto& r : rho) std::fill(r.begin(), r.end(), 0.0);
    std::fill(phi.begin(), phi.end(), 0.0);
    const unsigned n = o.num_cmat;
    for (unsigned col = 0; col < n; ++col) {
      const Position& cp = o.cmat_rel.at(col);
      rho[0][id3(cp.i, cp.j, cp.k)] = 1.0;
      solvePoisson();
      for (unsigned row = 0; row < n; ++row) {
        const Position& tp = o.cmat_rel.at(row);
        o.cmat[(size_t)col * n + row] = phi[id3(tp.i, tp.j, tp.k)];
      }
      rho[0][id3(cp.i, cp.j, cp.k)] = 0.0;
    }
    invertInPlace(o.cmat, n);
    o.updateTotalCmatValue();
  }
  initializeChargeMapOffset();
}

// Spacecraft.cpp:160-183
void Grid::initializeChargeMapOffset() {
  for (auto& o : objects) {
    if (o.initial_potential_offset == 0.0) continue;
    o.potential = o.initial_potential_offset;
    for (unsigned i = 0; i < o.num_cmat; ++i) {
      double d = 0.0;
      for (unsigned j = 0; j < o.num_cmat; ++j) d += o.cmat[(size_t)i * o.num_cmat + j] * o.potential;
      o.charge_map[i] += d;
    }
  }
}

}  // namespace refcpu
