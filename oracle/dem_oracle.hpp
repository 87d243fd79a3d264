// dem_oracle.hpp -- CPU restatement of DemSo's 2D DEM time step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it restates the
// reference kernels as plain host loops so the HIP kernels can be checked
// against it on identical inputs.  Only tests, the smoke check and the CPU
// baseline leg of the benchmark may load it; the product library never links
// or calls it.
//
// Every function follows one reference function, cited file:line
// (paths relative to the DemSo source tree):
//   calc_grid_pos / calc_grid_hash   src/particles_kernel.cuh:113-130
//   calc_hash                        src/particles_kernel.cuh:132-151
//   sort_particles                   src/functions.cu:336-341
//   reorder_and_find_cell_start      src/particles_kernel.cuh:154-233, src/functions.cu:371
//   series                           src/particles_kernel.cuh:36-41
//   collide_spheres                  src/particles_kernel.cuh:235-307
//   attract_spheres                  src/particles_kernel.cuh:310-342
//   collide_cell                     src/particles_kernel.cuh:345-385
//   collide_boundary                 src/particles_kernel.cuh:388-456
//   collide                          src/particles_kernel.cuh:459-549, src/functions.cu:441
//   integrate                        src/particles_kernel.cuh:580-627
//   step                             src/main.cu:396-507
//
// The arithmetic is written operation by operation in the reference's order
// (float2 operators are component-wise: includes/cutil_math.h:279-295,
// 474-490, 663-679, 856-874; dot = a.x*b.x + a.y*b.y, length = sqrtf(dot)).
// Expressions that involve the double literals M_PI, 2*M_PI and 1.2 are
// evaluated in double and rounded back, as C++ promotion rules prescribe.
// Compile with -ffp-contract=off: no fused multiply-add anywhere.
//
// Template parameter T is the state/arithmetic type: float reproduces the
// FP32 kernels; double gives the FP64 run used to state trajectory error.
//
// Extensions beyond the reference (each also implemented by the HIP path):
//   * per-particle radius (radius array non-null): m = (areal*r)*r and
//     I = 2*m*r*r/5 (the reference's sphere inertia, src/main.cu:193);
//   * clamp-to-edge grid mode for non-power-of-two grids (flag GRID_CLAMP):
//     own cell clamped, out-of-range neighbour cells contribute zero;
//   * open top (flag OPEN_TOP): the top wall clamp of
//     src/particles_kernel.cuh:611-613 is skipped;
//   * a hashing extent separate from the wall box, and several independent
//     instances with their own box/grid whose cells occupy disjoint key
//     ranges [cell_base, cell_base + Gx*Gy) (a material's `inst` names it).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace orc {

constexpr uint32_t EMPTY = 0xFFFFFFFFu;
enum : uint32_t { GRID_CLAMP = 1u, OPEN_TOP = 2u };

// Per-type material row: ParticleProperties (src/main.cuh:39-54) plus the
// areal density used with per-particle radii and the owning instance.
struct Material {
  float radius, mass, inertia, kn, kt, cn, bdamp, mu, attract, areal;
  uint32_t inst, pad;
};

// Per-instance geometry: SystemProperties' cubeDimension/gridSize
// (src/main.cuh:112-114) plus hashing extent and key offset.
struct Instance {
  float box[2];
  float extent[2];
  uint32_t grid[2];
  uint32_t cell_base, pad;
};

// Global scalars: SystemProperties' timeStep, gravity, boundary stiffness
// (src/main.cuh:116-121).
struct System {
  float dt;
  float g[2];
  float bkn, bkt;
  uint32_t flags;
  uint32_t num_types, num_instances;
  const Material* mat;
  const Instance* inst;
  // Type and instance indices are clamped to the table sizes (out-of-range
  // values would read past the reference's constant tables).
  const Material& material(uint32_t t) const { return mat[t < num_types ? t : num_types - 1u]; }
  const Instance& instance_of(uint32_t t) const {
    const uint32_t k = material(t).inst;
    return inst[k < num_instances ? k : num_instances - 1u];
  }
};

template <class T> struct V2 { T x, y; };

// The controlled ("big") particle of the reference's default build
// (USE_BIG_PARTICLE, src/main.cuh:29): a disc of material `type` at pos with
// velocity vel and spin omega that every particle collides with
// (src/particles_kernel.cuh:525-543), moved kinematically after each step
// (src/main.cu:510-514: pos += move, computed in double as the reference's
// double literal; y < wrap_below_y -> reset_y; x > wrap_above_x -> reset_x).
struct Control {
  float pos[2], vel[2];
  float theta, omega;
  uint32_t type, enabled;
  double move[2];
  double wrap_below_y, wrap_above_x;
  float reset_y, reset_x;
};
template <class T> inline V2<T> mk(T x, T y) { return V2<T>{x, y}; }
template <class T> inline V2<T> operator+(V2<T> a, V2<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <class T> inline V2<T> operator-(V2<T> a, V2<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <class T> inline V2<T> operator*(T s, V2<T> a) { return mk(s * a.x, s * a.y); }
template <class T> inline V2<T> operator*(V2<T> a, T s) { return mk(a.x * s, a.y * s); }
template <class T> inline V2<T> operator/(V2<T> a, T s) { return mk(a.x / s, a.y / s); }
template <class T> inline void operator+=(V2<T>& a, V2<T> b) { a.x += b.x; a.y += b.y; }
template <class T> inline T dot(V2<T> a, V2<T> b) { return a.x * b.x + a.y * b.y; }
template <class T> inline T length(V2<T> v) { return std::sqrt(dot(v, v)); }

// float -> int for a floor()ed grid coordinate, saturating outside int range
// and mapping NaN to 0 (the reference's conversion is undefined there).
template <class T> inline int to_int_sat(T g) {
  if (!(g == g)) return 0;
  if (g >= T(2147483648.0)) return 2147483647;
  if (g < T(-2147483648.0)) return (-2147483647 - 1);
  return int(g);
}

// calcGridPos (src/particles_kernel.cuh:114-121): floor(p * G / L) per axis,
// with L the hashing extent.
template <class T> inline void calc_grid_pos(const Instance& in, V2<T> p, int& gx, int& gy) {
  gx = to_int_sat<T>(std::floor(p.x * T(float(in.grid[0])) / T(in.extent[0])));
  gy = to_int_sat<T>(std::floor(p.y * T(float(in.grid[1])) / T(in.extent[1])));
}

// Cell coordinates of a particle's own cell: wrap (src/particles_kernel.cuh:127-128)
// or clamp.
inline void own_cell(const Instance& in, bool clamp, int gx, int gy, int& cx, int& cy) {
  if (clamp) {
    cx = std::min(std::max(gx, 0), int(in.grid[0]) - 1);
    cy = std::min(std::max(gy, 0), int(in.grid[1]) - 1);
  } else {
    cx = int(uint32_t(gx) & (in.grid[0] - 1u));
    cy = int(uint32_t(gy) & (in.grid[1] - 1u));
  }
}

// calcGridHash (src/particles_kernel.cuh:125-130) for in-range cell coords.
inline uint32_t cell_key(const Instance& in, int cx, int cy) {
  return in.cell_base + uint32_t(cy) * in.grid[0] + uint32_t(cx);
}

// calcHashD (src/particles_kernel.cuh:133-151).
template <class T>
void calc_hash(const System& s, const V2<T>* pos, const uint32_t* type, uint32_t* hash,
               uint32_t* index, uint32_t n) {
  const bool clamp = s.flags & GRID_CLAMP;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    const Instance& in = type ? s.instance_of(type[i]) : s.inst[0];
    int gx, gy, cx, cy;
    calc_grid_pos(in, pos[i], gx, gy);
    own_cell(in, clamp, gx, gy, cx, cy);
    hash[i] = cell_key(in, cx, cy);
    index[i] = uint32_t(i);
  }
}

// thrust::sort_by_key (src/functions.cu:336-341), defined here as the stable
// ascending sort of (key, value) pairs: ties keep their input order.
inline void sort_particles(uint32_t* hash, uint32_t* index, uint32_t n) {
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  std::stable_sort(perm.begin(), perm.end(),
                   [&](uint32_t a, uint32_t b) { return hash[a] < hash[b]; });
  std::vector<uint32_t> k(n), v(n);
  for (uint32_t i = 0; i < n; ++i) { k[i] = hash[perm[i]]; v[i] = index[perm[i]]; }
  if (n) {
    std::memcpy(hash, k.data(), sizeof(uint32_t) * n);
    std::memcpy(index, v.data(), sizeof(uint32_t) * n);
  }
}

// Particle state arrays (one SoA set).  radius may be null (type radii).
template <class T> struct State {
  V2<T>* pos; V2<T>* vel; T* theta; T* omega; uint32_t* id; uint32_t* type; uint32_t* loc; T* radius;
};

// cudaMemset(cellStart, 0xff) (src/functions.cu:371) followed by
// reorderDataAndFindCellStartD (src/particles_kernel.cuh:156-233).
// Keys >= num_cells, indices >= n and IDs >= n are skipped (the reference
// would access out of bounds).
template <class T>
void reorder_and_find_cell_start(uint32_t* cell_start, uint32_t* cell_end, const State<T>& srt,
                                 const uint32_t* hash, const uint32_t* index, const State<T>& old,
                                 uint32_t n, uint32_t num_cells) {
  std::fill(cell_start, cell_start + num_cells, EMPTY);
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < (long long)n; ++ii) {
    const uint32_t i = uint32_t(ii);
    const uint32_t h = hash[i];
    if (i == 0 || h != hash[i - 1]) {
      if (h < num_cells) cell_start[h] = i;
      if (i > 0 && hash[i - 1] < num_cells) cell_end[hash[i - 1]] = i;
    }
    if (i == n - 1 && h < num_cells) cell_end[h] = i + 1;
    const uint32_t k = index[i];
    if (k >= n) continue;
    srt.pos[i] = old.pos[k];
    srt.vel[i] = old.vel[k];
    srt.theta[i] = old.theta[k];
    srt.omega[i] = old.omega[k];
    const uint32_t id = old.id[k];
    srt.id[i] = id;
    if (id < n) srt.loc[id] = i;
    srt.type[i] = old.type[k];
    if (srt.radius) srt.radius[i] = old.radius[k];
  }
}

// seriesAssociation (src/particles_kernel.cuh:36-41), in the run's precision
// (float for the FP32 run, as the reference; double for the FP64 run).
template <class T> inline T series(T a, T b) {
  if (a == T(0) || b == T(0)) return T(0);
  return (a * b) / (a + b);
}

template <class T> struct Body {
  T r, m, inertia;
};

template <class T> inline Body<T> body(const System& s, const T* radius, uint32_t i, uint32_t type) {
  const Material& M = s.material(type);
  Body<T> b;
  if (radius) {
    b.r = radius[i];
    b.m = (T(M.areal) * b.r) * b.r;
    b.inertia = (((T(2) * b.m) * b.r) * b.r) / T(5);
  } else {
    b.r = T(M.radius);
    b.m = T(M.mass);
    b.inertia = T(M.inertia);
  }
  return b;
}

// pressure += len/(4*M_PI*r*r), evaluated in double (src/particles_kernel.cuh:284).
template <class T> inline void add_pressure(T& p, T len, T r) {
  p = T(double(p) + double(len) / (4 * M_PI * double(r) * double(r)));
}

// collideSpheres (src/particles_kernel.cuh:236-307): force on A from B.
template <class T>
V2<T> collide_spheres(const System& s, V2<T> posA, V2<T> posB, V2<T> velA, V2<T> velB, T omegaA,
                      T omegaB, uint32_t typeA, uint32_t typeB, T rA, T rB, T& moment, T& pressure) {
  const Material& A = s.material(typeA);
  const Material& B = s.material(typeB);
  V2<T> relPos = posB - posA;
  T dist = length(relPos);
  T collideDist = rA + rB;
  V2<T> force = mk(T(0), T(0));
  if (dist < collideDist) {
    V2<T> norm = relPos / dist;
    V2<T> tang = mk(-norm.y, norm.x);
    V2<T> relVel = velB - velA;
    T relVel_n = dot(relVel, norm);
    V2<T> relVel_t = relVel - relVel_n * norm;
    T kn = series<T>(A.kn, B.kn);
    T kt = series<T>(A.kt, B.kt);
    T cn = series<T>(A.cn, B.cn);
    T mu = (T(A.mu) + T(B.mu)) / T(2);
    force = (-kn * (collideDist - dist)) * norm;
    if (relVel_n < T(0)) force += (cn * relVel_n) * norm;
    add_pressure(pressure, length(force), rA);
    if (mu != T(0)) {
      V2<T> contactVel_t = relVel_t - (rA * omegaA + rB * omegaB) * tang;
      V2<T> Ftrial = (kt * contactVel_t) * T(s.dt);
      T Ftmax = mu * length(force);
      V2<T> Ft;
      if (length(Ftrial) < Ftmax) Ft = Ftrial;
      else Ft = (Ftmax * contactVel_t) / length(contactVel_t);
      force += Ft;
      moment += rA * dot(Ft, tang);
    }
  }
  return force;
}

// attractSpheres (src/particles_kernel.cuh:311-342).
template <class T>
V2<T> attract_spheres(const System& s, V2<T> posA, V2<T> posB, uint32_t typeA, uint32_t typeB, T rA,
                      T rB, T& pressure) {
  const Material& A = s.material(typeA);
  const Material& B = s.material(typeB);
  V2<T> relPos = posB - posA;
  T dist = length(relPos);
  T attractDistA = T(double(rA * T(2)) * 1.2);
  T attractDistB = T(double(rB * T(2)) * 1.2);
  V2<T> norm = relPos / dist;
  V2<T> force = mk(T(0), T(0));
  if (dist < attractDistA) force += (T(A.attract) * (attractDistA - dist)) * norm;
  if (dist < attractDistB) force += (T(B.attract) * (attractDistB - dist)) * norm;
  add_pressure(pressure, length(force), rA);
  return force;
}

// collideD (src/particles_kernel.cuh:459-549, without the controlled-particle
// block) over collideCell (:345-385), including the pressure reset of
// src/functions.cu:441.  Outputs acc = F/m and alpha = M/I per particle.
template <class T>
void collide(const System& s, const V2<T>* pos, const V2<T>* vel, V2<T>* acc, const T* omega,
             T* alpha, const uint32_t* type, const T* radius, const uint32_t* cell_start,
             const uint32_t* cell_end, uint32_t n, T* pressure, const Control* ctl = nullptr,
             T* react_x = nullptr, T* react_y = nullptr, T* react_m = nullptr) {
  const bool clamp = s.flags & GRID_CLAMP;
#pragma omp parallel for schedule(dynamic, 1024)
  for (long long ii = 0; ii < (long long)n; ++ii) {
    const uint32_t i = uint32_t(ii);
    const uint32_t tA = type[i];
    const Instance& in = s.instance_of(tA);
    const Body<T> bA = body<T>(s, radius, i, tA);
    const V2<T> p = pos[i], v = vel[i];
    const T w = omega[i];
    int gx, gy;
    calc_grid_pos(in, p, gx, gy);
    if (clamp) own_cell(in, true, gx, gy, gx, gy);
    V2<T> force = mk(T(0), T(0));
    T moment = T(0), pres = T(0);
    for (int y = -1; y <= 1; ++y) {
      for (int x = -1; x <= 1; ++x) {
        V2<T> cf = mk(T(0), T(0));
        int cx = gx + x, cy = gy + y;
        bool valid = true;
        if (clamp) valid = cx >= 0 && cy >= 0 && cx < int(in.grid[0]) && cy < int(in.grid[1]);
        else own_cell(in, false, int(uint32_t(gx) + uint32_t(x)), int(uint32_t(gy) + uint32_t(y)), cx, cy);
        if (valid) {
          const uint32_t h = cell_key(in, cx, cy);
          const uint32_t st = cell_start[h];
          if (st != EMPTY) {
            const uint32_t en = std::min(cell_end[h], n);
            for (uint32_t j = st; j < en; ++j) {
              if (j == i) continue;
              const uint32_t tB = type[j];
              const T rB = body<T>(s, radius, j, tB).r;
              cf += collide_spheres<T>(s, p, pos[j], v, vel[j], w, omega[j], tA, tB, bA.r, rB, moment, pres);
              cf += attract_spheres<T>(s, p, pos[j], tA, tB, bA.r, rB, pres);
            }
          }
        }
        force += cf;
      }
    }
    if (ctl && ctl->enabled) {
      // src/particles_kernel.cuh:527-541 (USE_ATOMIC 0): the pair with the
      // controlled particle, summed on its own, then added; the reaction on
      // the controlled particle goes to per-particle vectors
      const uint32_t tC = ctl->type;
      const T rC = T(s.material(tC).radius);
      const V2<T> pC = mk(T(ctl->pos[0]), T(ctl->pos[1])), vC = mk(T(ctl->vel[0]), T(ctl->vel[1]));
      T m = T(0);
      V2<T> f = collide_spheres<T>(s, p, pC, v, vC, w, T(ctl->omega), tA, tC, bA.r, rC, m, pres);
      f += attract_spheres<T>(s, p, pC, tA, tC, bA.r, rC, pres);
      force += f;
      moment += m;
      react_x[i] = -f.x;
      react_y[i] = -f.y;
      react_m[i] = -m / bA.r * rC;
    }
    acc[i] = force / bA.m;
    alpha[i] = moment / bA.inertia;
    pressure[i] = pres;
  }
}

// Sum of the per-particle reactions as the reference reduces them
// (reduceVecD, src/particles_kernel.cuh:552-578, launched with 512 threads by
// src/functions.cu:475-513): a pairwise tree over each block of 512
// consecutive values (the last block padded with zeros), then the block sums
// added in order.  (For n < 512 the reference launches n threads and its tree
// drops values when n is not a power of two; padding to 512 sums them all.)
template <class T> T reduce_reaction(const T* v, uint32_t n) {
  const uint32_t B = 512;
  T total = T(0);
  std::vector<T> cache(B);
  for (uint32_t b0 = 0; b0 < n; b0 += B) {
    for (uint32_t t = 0; t < B; ++t) cache[t] = b0 + t < n ? v[b0 + t] : T(0);
    for (uint32_t i = B / 2; i != 0; i /= 2)
      for (uint32_t t = 0; t < i; ++t) cache[t] += cache[t + i];
    total += cache[0];
  }
  return total;
}

// The kinematic move of the controlled particle after a step
// (src/main.cu:511-514).
inline void move_control(Control& c) {
  c.pos[0] = float(double(c.pos[0]) + c.move[0]);
  c.pos[1] = float(double(c.pos[1]) + c.move[1]);
  if (double(c.pos[1]) < c.wrap_below_y) c.pos[1] = c.reset_y;
  if (double(c.pos[0]) > c.wrap_above_x) c.pos[0] = c.reset_x;
}

// collideBoundary (src/particles_kernel.cuh:389-456): the bottom-wall law.
template <class T>
V2<T> collide_boundary(const System& s, V2<T> pos, V2<T> vel, T omega, uint32_t type, const Body<T>& b,
                       V2<T> boundPos, T& alpha, T& pressure) {
  const Material& M = s.material(type);
  V2<T> relPos = boundPos - pos;
  T dist = length(relPos);
  V2<T> force = mk(T(0), T(0));
  if (dist < b.r) {
    V2<T> norm = relPos / dist;
    V2<T> tang = mk(-norm.y, norm.x);
    V2<T> relVel = mk(-vel.x, -vel.y);
    T relVel_n = dot(relVel, norm);
    V2<T> relVel_t = relVel - relVel_n * norm;
    T kt = series<T>(M.kt, s.bkt);
    T cn = T(M.cn);
    force = (-b.m) * mk(T(s.g[0]), T(s.g[1]));
    if (relVel_n < T(0)) force += (cn * relVel_n) * norm;
    add_pressure(pressure, length(force), b.r);
    if (M.mu != 0.0f) {
      V2<T> contactVel_t = relVel_t - (b.r * omega) * tang;
      V2<T> Ftrial = (kt * contactVel_t) * T(s.dt);
      T Ftmax = T(M.mu) * length(force);
      V2<T> Ft;
      if (length(Ftrial) < Ftmax) Ft = Ftrial;
      else Ft = (Ftmax * contactVel_t) / length(contactVel_t);
      force = Ft;
      alpha += (b.r * dot(Ft, tang)) / b.inertia;
    }
  }
  return force / b.m;
}

// theta wrap of src/particles_kernel.cuh:594-595, in double.
template <class T> inline T wrap_theta(T th) {
  if (double(th) > 2 * M_PI) th = T(double(th) - 2 * M_PI * std::floor(double(th) / (2 * M_PI)));
  if (th < T(0)) th = T(double(th) - 2 * M_PI * std::floor(double(th) / (2 * M_PI)));
  return th;
}

// integrateSystemD (src/particles_kernel.cuh:581-627), in place.
template <class T>
void integrate(const System& s, V2<T>* pos, V2<T>* vel, const V2<T>* acc, T* theta, T* omega,
               const T* alpha, const uint32_t* type, const T* radius, T* pressure, uint32_t n) {
  const T dt = T(s.dt);
  const V2<T> g = mk(T(s.g[0]), T(s.g[1]));
  const bool open_top = s.flags & OPEN_TOP;
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < (long long)n; ++ii) {
    const uint32_t i = uint32_t(ii);
    const uint32_t t = type[i];
    const Material& M = s.material(t);
    const Instance& in = s.instance_of(t);
    const Body<T> b = body<T>(s, radius, i, t);
    V2<T> v = vel[i] + (g + acc[i]) * dt;
    V2<T> p = pos[i] + v * dt;
    T w = omega[i] + alpha[i] * dt;
    T th = theta[i] + w * dt;
    th = wrap_theta(th);
    const T bd = -T(M.bdamp);
    const T Lx = T(in.box[0]), Ly = T(in.box[1]);
    if (p.x > Lx - b.r) { p.x = Lx - b.r; v.x = v.x * bd; }
    if (p.x < b.r) { p.x = b.r; v.x = v.x * bd; }
    if (!open_top && p.y > Ly - b.r) { p.y = Ly - b.r; v.y = v.y * bd; }
    if (p.y < b.r) {
      T wallAlpha = T(0);
      T pr = pressure[i];
      V2<T> wallAcc = collide_boundary<T>(s, p, v, w, t, b, mk(p.x, T(0)), wallAlpha, pr);
      v += wallAcc * dt;
      p += (wallAcc * dt) * dt;
      w += wallAlpha * dt;
      th += (wallAlpha * dt) * dt;
      p.y = b.r;
      v.y = v.y * bd;
      pressure[i] = pr;
    }
    pos[i] = p;
    vel[i] = v;
    omega[i] = w;
    theta[i] = th;
  }
}

// Owned double-buffered state and one step in the reference's stage order
// (src/main.cu:396-507): hash -> sort -> reorder -> collide -> integrate;
// the sorted buffer of step k is the unsorted input of step k+1.
template <class T> struct Sim {
  System sys;
  uint32_t n = 0, num_cells = 0;
  bool has_radius = false;
  std::vector<V2<T>> pos[2], vel[2], acc;
  std::vector<T> theta[2], omega[2], alpha, pressure, radius[2];
  std::vector<uint32_t> id[2], type[2], loc[2], hash, index, cell_start, cell_end;
  int cur = 0;
  Control ctl{};
  std::vector<T> react_x, react_y, react_m;
  T reaction[3] = {T(0), T(0), T(0)};  // force on the controlled particle (x, y) and its moment

  State<T> st(int b) {
    return State<T>{pos[b].data(), vel[b].data(), theta[b].data(), omega[b].data(), id[b].data(),
                    type[b].data(), loc[b].data(), has_radius ? radius[b].data() : nullptr};
  }
  void resize(uint32_t n_, uint32_t cells, bool rad) {
    n = n_; num_cells = cells; has_radius = rad;
    for (int b = 0; b < 2; ++b) {
      pos[b].assign(n, V2<T>{0, 0}); vel[b].assign(n, V2<T>{0, 0});
      theta[b].assign(n, 0); omega[b].assign(n, 0); radius[b].assign(rad ? n : 0, 0);
      id[b].assign(n, 0); type[b].assign(n, 0); loc[b].assign(n, 0);
    }
    acc.assign(n, V2<T>{0, 0}); alpha.assign(n, 0); pressure.assign(n, 0);
    hash.assign(n, 0); index.assign(n, 0);
    cell_start.assign(num_cells, EMPTY); cell_end.assign(num_cells, 0);
  }
  void step() {
    State<T> o = st(cur), s = st(cur ^ 1);
    calc_hash<T>(sys, o.pos, o.type, hash.data(), index.data(), n);
    sort_particles(hash.data(), index.data(), n);
    reorder_and_find_cell_start<T>(cell_start.data(), cell_end.data(), s, hash.data(), index.data(), o,
                                   n, num_cells);
    if (ctl.enabled) {
      react_x.resize(n);
      react_y.resize(n);
      react_m.resize(n);
    }
    collide<T>(sys, s.pos, s.vel, acc.data(), s.omega, alpha.data(), s.type, s.radius,
               cell_start.data(), cell_end.data(), n, pressure.data(), &ctl, react_x.data(), react_y.data(),
               react_m.data());
    if (ctl.enabled) {
      reaction[0] = reduce_reaction(react_x.data(), n);
      reaction[1] = reduce_reaction(react_y.data(), n);
      reaction[2] = reduce_reaction(react_m.data(), n);
    }
    integrate<T>(sys, s.pos, s.vel, acc.data(), s.theta, s.omega, alpha.data(), s.type, s.radius,
                 pressure.data(), n);
    if (ctl.enabled) move_control(ctl);
    cur ^= 1;
  }
};

}  // namespace orc
