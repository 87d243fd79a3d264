// kernels_step.hip -- HIP kernels for hash, cell fill, reorder/cellStart,
// collide and integrate on gfx950.
//
// One lane per particle throughout.  Arithmetic follows the reference kernel
// bodies operation by operation (no FMA contraction in this file: the pragma
// below and -ffp-contract=off), so FP32 results equal the IEEE sequence of
// src/particles_kernel.cuh.  Terms the reference writes with the double
// literals M_PI / 2*M_PI / 1.2 are evaluated in double as C++ promotes them.
#include "dem_internal.h"

#include <math.h>

#pragma clang fp contract(off)

namespace dem {
namespace {

__device__ __forceinline__ int to_int_sat(float g) {
  // NaN -> 0, saturation at the int range, plain truncation inside it
  int r = (int)fminf(fmaxf(g, -2147483648.0f), 2147483520.0f);
  r = (g >= 2147483648.0f) ? 2147483647 : r;
  return (g == g) ? r : 0;
}

__device__ __forceinline__ uint32_t clamp_idx(uint32_t t, uint32_t count) {
  return t < count ? t : count - 1u;
}

// calcGridPos (src/particles_kernel.cuh:114-121): floor(p * G / extent).
__device__ __forceinline__ void grid_pos(const Instance& in, float2 p, int& gx, int& gy) {
  gx = to_int_sat(floorf(p.x * (float)in.grid[0] / in.extent[0]));
  gy = to_int_sat(floorf(p.y * (float)in.grid[1] / in.extent[1]));
}

__device__ __forceinline__ void clamp_cell(const Instance& in, int& gx, int& gy) {
  gx = min(max(gx, 0), (int)in.grid[0] - 1);
  gy = min(max(gy, 0), (int)in.grid[1] - 1);
}

// seriesAssociation (src/particles_kernel.cuh:36-41).
__device__ __forceinline__ float series(float a, float b) {
  if (a == 0.0f || b == 0.0f) return 0.0f;
  return (a * b) / (a + b);
}

// Radius, mass, inertia of particle i: from its type row, or from the
// per-particle radius with m = (areal*r)*r and I = 2*m*r*r/5 (src/main.cu:193).
template <bool kRadius>
__device__ __forceinline__ void body(const Material& m, const float* __restrict__ radius,
                                     uint32_t i, float& r, float& mass, float& inertia) {
  if (kRadius) {
    r = radius[i];
    mass = (m.areal * r) * r;
    inertia = (((2.0f * mass) * r) * r) / 5.0f;
  } else {
    r = m.radius;
    mass = m.mass;
    inertia = m.inertia;
  }
}

// pressure += len/(4*M_PI*r*r) in double (src/particles_kernel.cuh:284, 339, 433);
// den is 4*M_PI*r*r evaluated once per particle in the same order.
__device__ __forceinline__ float add_pressure(float p, float len, double den) {
  return (float)((double)p + (double)len / den);
}

// RN(sqrt(x)) < m, the Coulomb test of the shear law (length(Ftrial) < Ftmax,
// src/particles_kernel.cuh:295), for x a sum of two
// float squares (+0 or more, or NaN).  RN is monotone and the root of a
// float is never a rounding midpoint, so the test is x < (m-)^2, m- the
// midpoint below m, which lies in [m^2 (1 - 2^-23), m^2).  For m in
// [2^-63, 2^63] t = RN(m m) is a normal float within 2^-24 of m^2, so
// x < RN(t (1 - 2^-20)) proves the test true and x > RN(t (1 + 2^-20))
// proves it false; only between the two (or for other m, or NaN) is the
// correctly rounded root taken.
__device__ __forceinline__ bool sqrt_below(float x, float m) {
  if (m >= 0x1p-63f && m <= 0x1p63f) {
    const float t = m * m;
    if (x < t * (1.0f - 0x1p-20f)) return true;
    if (x > t * (1.0f + 0x1p-20f)) return false;
  }
  return sqrtf(x) < m;
}

__device__ __forceinline__ double pressure_den(float r) {
  return 4 * M_PI * (double)r * (double)r;
}

// ---------------------------------------------------------------- calcHash --
// Cell key of one particle: calcGridHash (src/particles_kernel.cuh:123-130)
// with the instance's key base; clamped or wrapped as the flags say.
__device__ __forceinline__ uint32_t cell_key_in(const StepConst& c, const Instance& in, float2 p) {
  int gx, gy;
  grid_pos(in, p, gx, gy);
  uint32_t cx, cy;
  if (c.flags & DEM_GRID_CLAMP) {
    clamp_cell(in, gx, gy);
    cx = (uint32_t)gx;
    cy = (uint32_t)gy;
  } else {
    cx = (uint32_t)gx & (in.grid[0] - 1u);
    cy = (uint32_t)gy & (in.grid[1] - 1u);
  }
  return in.cell_base + cy * in.grid[0] + cx;
}

template <bool kMulti>
__device__ __forceinline__ uint32_t cell_key(const StepConst& c, float2 p, const uint32_t* __restrict__ type,
                                             uint32_t i) {
  const Instance in =
      kMulti ? c.inst[clamp_idx(c.mat[clamp_idx(type[i], c.num_types)].inst, c.num_instances)] : c.inst0;
  return cell_key_in(c, in, p);
}

// calcHashD (src/particles_kernel.cuh:133-151).  kPer particles a lane: a
// block covers kPer * blockDim consecutive particles, lane t taking t,
// t + blockDim, ... (each pass over the block's lanes stays coalesced).
template <bool kMulti, int kPer>
__global__ __launch_bounds__(256) void k_hash(StepConst c, const float2* __restrict__ pos,
                                              const uint32_t* __restrict__ type,
                                              uint32_t* __restrict__ hash,
                                              uint32_t* __restrict__ index, uint32_t n) {
  const uint32_t base = blockIdx.x * blockDim.x * kPer + threadIdx.x;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const uint32_t i = base + (uint32_t)k * blockDim.x;
    if (i < n) {
      hash[i] = cell_key<kMulti>(c, pos[i], type, i);
      index[i] = i;
    }
  }
}

// -------------------------------------------------------------- cell fill --
// cudaMemset(cellStart, 0xff) of src/functions.cu:371 as a kernel (a memset
// node in a captured graph is avoided on purpose).
__global__ __launch_bounds__(256) void k_fill(uint32_t* __restrict__ p, uint32_t v, uint32_t n, int vec) {
  const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) * 4u;
  if (i >= n) return;
  if (vec && i + 3u < n) {
    *reinterpret_cast<uint4*>(p + i) = make_uint4(v, v, v, v);
  } else {
    for (uint32_t k = i; k < n && k < i + 4u; ++k) p[k] = v;
  }
}

// --------------------------------------------------------------- reorder --
// reorderDataAndFindCellStartD (src/particles_kernel.cuh:156-233).  The
// previous key is read from global memory (it is in the same cache line for
// all lanes but lane 0).  Keys >= num_cells, indices >= n and IDs >= n are
// skipped instead of being written out of bounds.  kFull = false gathers the
// positions and the per-particle constants only: the owned simulation's step
// reads velocity, angle and spin through the sorted index instead (collide
// for its own particle and its contacts, integrate for every particle).
template <bool kRadius, bool kFull, int kPer>
__global__ __launch_bounds__(256) void k_reorder(
    uint32_t* __restrict__ cell_start, uint32_t* __restrict__ cell_end, float2* __restrict__ s_pos,
    float2* __restrict__ s_vel, float* __restrict__ s_theta, float* __restrict__ s_omega,
    uint32_t* __restrict__ s_id, uint32_t* __restrict__ s_loc, uint32_t* __restrict__ s_type,
    float* __restrict__ s_radius, const uint32_t* __restrict__ hash,
    const uint32_t* __restrict__ index, const float2* __restrict__ o_pos,
    const float2* __restrict__ o_vel, const float* __restrict__ o_theta,
    const float* __restrict__ o_omega, const uint32_t* __restrict__ o_id,
    const uint32_t* __restrict__ o_type, const float* __restrict__ o_radius, uint32_t n,
    uint32_t num_cells) {
  // kPer particles a lane, as k_hash
  const uint32_t base = blockIdx.x * blockDim.x * kPer + threadIdx.x;
#pragma unroll
  for (int kk = 0; kk < kPer; ++kk) {
    const uint32_t i = base + (uint32_t)kk * blockDim.x;
    if (i >= n) continue;
    const uint32_t h = hash[i];
    if (i == 0) {
      if (h < num_cells) cell_start[h] = 0;
    } else {
      const uint32_t hp = hash[i - 1];
      if (h != hp) {
        if (h < num_cells) cell_start[h] = i;
        if (hp < num_cells) cell_end[hp] = i;
      }
    }
    if (i == n - 1 && h < num_cells) cell_end[h] = n;

    const uint32_t k = index[i];
    if (k >= n) continue;
    s_pos[i] = o_pos[k];
    if (kFull) {
      s_vel[i] = o_vel[k];
      s_theta[i] = o_theta[k];
      s_omega[i] = o_omega[k];
    }
    const uint32_t id = o_id[k];
    s_id[i] = id;
    if (id < n) s_loc[id] = i;
    s_type[i] = o_type[k];
    if (kRadius) s_radius[i] = o_radius[k];
  }
}

// --------------------------------------------------------------- collide --
// collideD (src/particles_kernel.cuh:459-549) with collideCell (:345-385),
// collideSpheres (:235-307) and attractSpheres (:310-342) inlined.
//
// The kernel's cost is per candidate (instructions and the latency of the
// dependent loads), so the candidate loop is kept lean:
//  * the three cells of a row are consecutive keys, so their candidates are
//    one contiguous range of the sorted arrays; collideCell returns one
//    partial float2 per cell that collideD adds up, so the running partial
//    sum is flushed into the total when a candidate that may exert a force
//    lies in a later cell (RowSum; flushing only there changes nothing, as
//    candidates without force add +0 and the sums start at +0 and never
//    become -0).  Where a wrapping grid makes the row non-contiguous (edge
//    columns), the row's three cells are walked one by one in the reference
//    order;
//  * most candidates are not in contact: d2 = rx*rx+ry*ry is the float the
//    reference takes sqrtf of, and if d2 > fl(R*R)*(1+2^-20) then
//    sqrtf(d2) >= R for certain (fl(R*R)*(1+2^-20) exceeds the square of
//    the midpoint between R and its predecessor while R*R is a normal float,
//    i.e. R >= 2^-63; smaller R always takes the exact path), so the correctly rounded
//    sqrt and the contact law run only for candidates that may touch; the
//    contact decision itself is the reference's dist < R;
//  * the neighbour's velocity, spin and material are loaded only for actual
//    contacts, except in the plain mode, which loads them with the position
//    (the reference loads them for every candidate; the values used are the
//    same).
constexpr uint32_t kNoBoundary = 0xFFFFFFFFu;

struct PairCtx {
  float2 p, v;
  float w, rA, dt;
  double den;
  double rden;  // RN(1/den), for the contact-list mode's pressure terms
  Material mA;
  uint32_t tA;
};

// The pressure term with the particle's rounded reciprocal y = RN(1/den) and
// two fused corrections: q1 = RN(a y) (off by more than an ulp for about 1%
// of inputs), q2 = RN(q1 + RN(a - q1 den) y) (within a tiny fraction of an
// ulp of a/den, so faithful), q3 = RN(q2 + (a - q2 den) y), where the
// remainder of the faithful q2 is exact.  q3 is the correctly rounded a/den
// (Markstein's theorem for y = RN(1/den) and a faithful quotient: the
// remainder and den * ulp/2 lie on the grid ulp(den) ulp/2 and differ from
// ulp/(2y) by less than one step of it, so q2 + r y and a/den lie on the same
// side of every rounding midpoint, and a/den is never one).  a is a float
// length and den = 4 pi r^2 of a float r, so no step overflows or underflows
// in double; a non-finite length divides as the reference does.  Five FP64
// operations per contact instead of a division sequence.
__device__ __forceinline__ double pressure_quotient(float len, double den, double rden) {
  const double a = (double)len;
  double qd = a * rden;
  const double rem = fma(-qd, den, a);
  if (__builtin_expect(__builtin_isfinite(rem), 1)) {
    qd = fma(rem, rden, qd);
    qd = fma(fma(-qd, den, a), rden, qd);
  } else {
    qd = a / den;
  }
  return qd;
}

__device__ __forceinline__ float add_pressure_recip(float p, float len, double den, double rden) {
  return (float)((double)p + pressure_quotient(len, den, rden));
}

// collideSpheres' contact law (src/particles_kernel.cuh:255-303) for a
// partner B of material tB at offset (rx, ry) and distance dist < R: the
// force on A as collideSpheres returns it (ffx, ffy), the length of its
// normal part (the pressure term's numerator) as the return value, and the
// friction moment term mterm (+0 without friction; the moment is a sum that
// starts at +0 and is never -0, so adding +0 leaves it unchanged).
__device__ __forceinline__ float contact_terms(const StepConst& c, const PairCtx& q, float rx, float ry, float dist,
                                               float R, float rB, uint32_t tB, float2 vB, float wB, float& ffx,
                                               float& ffy, float& mterm) {
  const float4 pr = c.pair[q.tA * c.num_types + tB];
  const float kn = pr.x, kt = pr.y, cn = pr.z, mu = pr.w;
  const float nx = rx / dist, ny = ry / dist;
  const float tx = -ny, ty = nx;
  const float rvx = vB.x - q.v.x, rvy = vB.y - q.v.y;
  const float vn = rvx * nx + rvy * ny;
  const float vtx = rvx - vn * nx, vty = rvy - vn * ny;
  const float s = -kn * (R - dist);
  ffx = s * nx;
  ffy = s * ny;
  if (vn < 0.0f) {
    const float s2 = cn * vn;
    ffx += s2 * nx;
    ffy += s2 * ny;
  }
  const float flen = sqrtf(ffx * ffx + ffy * ffy);
  mterm = 0.0f;
  if (mu != 0.0f) {
    const float sr = q.rA * q.w + rB * wB;
    const float cvx = vtx - sr * tx, cvy = vty - sr * ty;
    float ftx = (kt * cvx) * q.dt, fty = (kt * cvy) * q.dt;
    const float ftmax = mu * flen;
    if (!sqrt_below(ftx * ftx + fty * fty, ftmax)) {
      const float cl = sqrtf(cvx * cvx + cvy * cvy);
      ftx = (ftmax * cvx) / cl;
      fty = (ftmax * cvy) / cl;
    }
    ffx += ftx;
    ffy += fty;
    mterm = q.rA * (ftx * tx + fty * ty);
  }
  return flen;
}

// The contact law with its pressure term added to pres and its moment term
// to moment.
template <bool kRcp = false>
__device__ __forceinline__ void contact_force(const StepConst& c, const PairCtx& q, float rx, float ry, float dist,
                                              float R, float rB, uint32_t tB, float2 vB, float wB, float& ffx,
                                              float& ffy, float& moment, float& pres) {
  float mterm;
  const float flen = contact_terms(c, q, rx, ry, dist, R, rB, tB, vB, wB, ffx, ffy, mterm);
  pres = kRcp ? add_pressure_recip(pres, flen, q.den, q.rden) : add_pressure(pres, flen, q.den);
  moment += mterm;
}

// attractSpheres (src/particles_kernel.cuh:310-342) for a partner with
// attraction coefficient aB: the force on A and its pressure term.
__device__ __forceinline__ void attract_force(const PairCtx& q, float rx, float ry, float dist, float rB, float aB,
                                              float& afx, float& afy, float& pres) {
  const float aA = q.mA.attract;
  const float adA = (float)((double)(q.rA * 2.0f) * 1.2);
  const float adB = (float)((double)(rB * 2.0f) * 1.2);
  const float nx = rx / dist, ny = ry / dist;
  afx = 0.0f;
  afy = 0.0f;
  if (dist < adA) {
    const float s = aA * (adA - dist);
    afx += s * nx;
    afy += s * ny;
  }
  if (dist < adB) {
    const float s = aB * (adB - dist);
    afx += s * nx;
    afy += s * ny;
  }
  pres = add_pressure(pres, sqrtf(afx * afx + afy * afy), q.den);
}

// Velocity and spin of the sorted particles: vel[j], omega[j], or with a
// gather index vel[gat[j]], omega[gat[j]] (the previous step's arrays,
// before the sort).
// Element i of an array.  kOff32: addressed by a 32-bit byte offset from the
// array's base, which the loads take as a scalar base plus one 32-bit lane
// register (64-bit address arithmetic per load otherwise); only for arrays
// of at most 4 GiB, which the launcher checks.
template <bool kOff32, class T>
__device__ __forceinline__ const T& ld(const T* __restrict__ base, uint32_t i) {
  if constexpr (kOff32)
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T));
  else
    return base[i];
}

struct Dyn {
  const float2* __restrict__ vel;
  const float* __restrict__ omega;
  const uint32_t* __restrict__ gat;
  template <bool kOff32>
  __device__ __forceinline__ uint32_t at(uint32_t j) const { return gat ? ld<kOff32>(gat, j) : j; }
};

// A neighbour candidate as the pair law reads it.  With kDyn its velocity,
// spin and material row are loaded together with its position and radius
// (one dependent load fewer for every contact, three more registers);
// without, they are loaded only for candidates that touch.
struct Cand {
  float2 p;
  float r;
  uint32_t t;
  float2 v;
  float w;
};

// The row's per-cell partial sum: collideCell's sums start at +0 for every
// cell and collideD adds them in cell order.  Candidates that exert no force
// add +0, which leaves a partial sum unchanged (a sum that starts at +0 is
// never -0), so the partial sum is flushed into the total only where a
// candidate that may exert a force lies in a later cell of the row than the
// sum's (cell index (j >= b0) + (j >= b1), the row's cell ends), and at the
// row's end: the same additions in the same order as flushing at every cell
// end.
struct RowSum {
  float cfx = 0.0f, cfy = 0.0f;
  uint32_t cell = 0, b0, b1;
  __device__ __forceinline__ void at(uint32_t j, float& fx, float& fy) {
    const uint32_t k = (uint32_t)(j >= b0) + (uint32_t)(j >= b1);
    if (k != cell) {
      fx += cfx;
      fy += cfy;
      cfx = 0.0f;
      cfy = 0.0f;
      cell = k;
    }
  }
};

template <bool kRadius, bool kOne, bool kDyn, bool kO, bool kRcp = false>
__device__ __forceinline__ void candidate_at(const StepConst& c, const PairCtx& q, uint32_t j, bool self,
                                             const Cand& b, const Dyn& d, const uint32_t* __restrict__ type,
                                             RowSum& rsum, float& fx, float& fy, float& moment, float& pres) {
  const float rx = b.p.x - q.p.x, ry = b.p.y - q.p.y;
  const float d2 = rx * rx + ry * ry;
  const float R = q.rA + b.r;
  const bool maybe = !(d2 > (R * R) * 1.00000095367431640625f) || R < 1.0842022e-19f;
  // the particle itself is skipped (collideCell's j != index test)
  if (self || !(maybe || c.attract)) return;
  rsum.at(j, fx, fy);
  float& cfx = rsum.cfx;
  float& cfy = rsum.cfy;
  const float dist = sqrtf(d2);
  const float rB = b.r;
  uint32_t tB = b.t;
  if (kRadius && !kOne && !kDyn) tB = clamp_idx(ld<kO>(type, j), c.num_types);
  if (dist < R) {
    float ffx, ffy;
    float2 vB;
    float wB;
    if (kDyn) {
      vB = b.v;
      wB = b.w;
    } else {
      const uint32_t jv = d.at<kO>(j);
      vB = ld<kO>(d.vel, jv);
      wB = ld<kO>(d.omega, jv);
    }
    contact_force<kRcp>(c, q, rx, ry, dist, R, rB, tB, vB, wB, ffx, ffy, moment, pres);
    cfx += ffx;
    cfy += ffy;
  }
  if (c.attract) {
    float afx, afy;
    attract_force(q, rx, ry, dist, rB, c.mat[tB].attract, afx, afy, pres);
    cfx += afx;
    cfy += afy;
  }
}

// The pair with the controlled particle (src/particles_kernel.cuh:527-541,
// USE_ATOMIC 0): f = collideSpheres + attractSpheres, summed on its own and
// then added to the particle's force and moment; the reaction on the
// controlled particle, -f and -m / rA * rC, goes to per-particle vectors.
__device__ __forceinline__ void control_pair(const StepConst& c, const PairCtx& q, const dem_controlled& C,
                                             float& fx, float& fy, float& moment, float& pres, float& react_x,
                                             float& react_y, float& react_m) {
  const uint32_t tC = clamp_idx(C.type, c.num_types);
  const float rC = c.mat[tC].radius;
  const float rx = C.pos[0] - q.p.x, ry = C.pos[1] - q.p.y;
  const float dist = sqrtf(rx * rx + ry * ry);
  const float R = q.rA + rC;
  float m = 0.0f, f_x = 0.0f, f_y = 0.0f;
  if (dist < R) {
    float mterm = 0.0f;
    contact_force(c, q, rx, ry, dist, R, rC, tC, make_float2(C.vel[0], C.vel[1]), C.omega, f_x, f_y, mterm, pres);
    m = m + mterm;  // collideSpheres adds into the caller's m (0)
  }
  if (c.attract) {
    float afx, afy;
    attract_force(q, rx, ry, dist, rC, c.mat[tC].attract, afx, afy, pres);
    f_x += afx;
    f_y += afy;
  }
  fx += f_x;
  fy += f_y;
  moment += m;
  react_x = -f_x;
  react_y = -f_y;
  react_m = -m / q.rA * rC;
}

// Candidate j's position and radius (and, without a radius array, its
// material row), plus with kDyn its velocity, spin and material row.
// (kOne: every particle has material row 0, so no type is read.)
template <bool kRadius, bool kOne, bool kDyn, bool kO>
__device__ __forceinline__ void candidate_load(const StepConst& c, uint32_t j, const float2* __restrict__ pos,
                                               const uint32_t* __restrict__ type, const float* __restrict__ radius,
                                               const Dyn& d, Cand& b) {
  b.p = ld<kO>(pos, j);
  if (kRadius) {
    b.r = ld<kO>(radius, j);
    b.t = kDyn && !kOne ? clamp_idx(ld<kO>(type, j), c.num_types) : 0u;
  } else {
    b.t = kOne ? 0u : clamp_idx(ld<kO>(type, j), c.num_types);
    b.r = c.mat[b.t].radius;
  }
  if (kDyn) {
    const uint32_t jv = d.at<kO>(j);
    b.v = ld<kO>(d.vel, jv);
    b.w = ld<kO>(d.omega, jv);
  }
}

template <bool kRadius, bool kOne, bool kDyn, bool kO, bool kRcp = false>
__device__ __forceinline__ void candidate(const StepConst& c, const PairCtx& q, uint32_t j, bool self,
                                          const float2* __restrict__ pos, const Dyn& d,
                                          const uint32_t* __restrict__ type, const float* __restrict__ radius,
                                          RowSum& rsum, float& fx, float& fy, float& moment, float& pres) {
  Cand b;
  candidate_load<kRadius, kOne, kDyn, kO>(c, j, pos, type, radius, d, b);
  candidate_at<kRadius, kOne, kDyn, kO, kRcp>(c, q, j, self, b, d, type, rsum, fx, fy, moment, pres);
}

// Forces on sorted particle i from its 3x3 cell neighbourhood (the body of
// collideD, src/particles_kernel.cuh:459-549): sum of the pair forces fx,
// fy, the moment and the pressure, plus the particle's own radius, mass,
// inertia and material (for integrate).
struct Own {
  float r, mass, inertia;
  uint32_t t;
  Material m;
  Instance in;
  float2 p, v;
  float w;
};

// Candidate modes of the collide kernels: kPlain walks each row's candidates
// one load at a time (each candidate's velocity and spin loaded with its
// position); kPrefetch loads each one iteration ahead; kBatch loads the first
// candidate of all three rows at once before walking them (three loads in
// flight); kList first walks the rows testing only position and radius and
// sets a bit per candidate that may touch, then runs the contact law over the
// set bits of all three rows, so a wave executes the contact law as often as
// its busiest lane has contacts instead of at every candidate where some lane
// has one.  The order of the sums is the same in every mode.  Preloading
// two or three candidates a row needs 73 or 81 VGPRs (6 waves per SIMD
// instead of 8) and was slower.
constexpr int kPlain = 0, kPrefetch = 1, kBatch = 2, kList = 3;
constexpr int kBatchN = 1;

// Material layouts of the kernels: one material row (no type is read), one
// instance with several rows, or an ensemble of instances.
constexpr int kOneMat = 0, kManyMat = 1, kInstances = 2;

// kUni: every lane of the wave has material row t_uni (a wave-uniform value,
// so the material row and the instance geometry are scalar loads and the
// grid arithmetic scalar); the ensemble kernels take this path for every
// wave whose particles all belong to one instance.
template <bool kRadius, int kMat, int kMode, bool kControl, bool kO, bool kUni = false>
__device__ __forceinline__ void collide_particle(const StepConst& c, uint32_t i, const float2* __restrict__ pos,
                                                 const Dyn& d, const uint32_t* __restrict__ type,
                                                 const float* __restrict__ radius,
                                                 const uint32_t* __restrict__ cell_start,
                                                 const uint32_t* __restrict__ cell_end, uint32_t n, Own& own,
                                                 const ControlOut& ctl, float& fx_out, float& fy_out,
                                                 float& moment_out, float& pres_out, uint32_t t_uni = 0) {
  constexpr bool kMulti = kMat == kInstances, kOne = kMat == kOneMat;
  PairCtx q;
  own.t = kOne ? 0u : (kUni ? t_uni : clamp_idx(type[i], c.num_types));
  q.tA = own.t;
  q.mA = c.mat[own.t];
  own.m = q.mA;
  own.in = kMulti ? c.inst[clamp_idx(q.mA.inst, c.num_instances)] : c.inst0;
  const Instance& in = own.in;
  body<kRadius>(q.mA, radius, i, q.rA, own.mass, own.inertia);
  own.r = q.rA;
  q.p = pos[i];
  // The particle's own velocity and spin feed only the contact law (and the
  // fused kernel's integrate): the contact-list walk loads them once it has
  // marked a candidate (or at the end, where a caller uses them), the other
  // modes up front.
  bool have_dyn = false;
  auto load_own_dyn = [&]() {
    const uint32_t iv = d.at<kO>(i);
    q.v = d.vel[iv];
    q.w = d.omega[iv];
    have_dyn = true;
  };
  if (kMode != kList || kControl) load_own_dyn();
  own.p = q.p;
  q.den = pressure_den(q.rA);
  q.rden = 0.0;  // the contact-list walk sets it before its first contact
  q.dt = c.dt;
  const bool clamp = c.flags & DEM_GRID_CLAMP;
  const int G0 = (int)in.grid[0], G1 = (int)in.grid[1];

  int gx, gy;
  grid_pos(in, q.p, gx, gy);
  if (clamp) clamp_cell(in, gx, gy);
  // x columns of a row: clamp -> out-of-range skipped; wrap -> & (G-1)
  const bool vx0 = !clamp || gx > 0, vx2 = !clamp || gx < G0 - 1;
  const uint32_t x0 = ((uint32_t)gx - 1u) & (clamp ? 0xFFFFFFFFu : (uint32_t)G0 - 1u);
  const uint32_t x1 = clamp ? (uint32_t)gx : ((uint32_t)gx & ((uint32_t)G0 - 1u));
  const uint32_t x2 = ((uint32_t)gx + 1u) & (clamp ? 0xFFFFFFFFu : (uint32_t)G0 - 1u);
  // a wrap that moves a column to the other edge breaks key contiguity
  const bool contiguous = clamp || (x0 + 1u == x1 && x1 + 1u == x2);

  float fx = 0.0f, fy = 0.0f, moment = 0.0f, pres = 0.0f;
  if (contiguous) {
    // The usual case: every row's three cells are consecutive keys, so a row
    // is one candidate range [rs, re) with the cell ends b0, b1 inside it.
    // The ranges of all three rows are loaded first (their loads overlap),
    // and each candidate's position and radius are fetched one iteration
    // ahead; the summation order is unchanged.
    uint32_t rs[3], re[3], b0[3], b1[3];
#pragma unroll
    for (int y = 0; y < 3; ++y) {
      const int iy = gy + y - 1;
      uint32_t cy;
      bool ok = true;
      if (clamp) {
        ok = iy >= 0 && iy < G1;
        cy = ok ? (uint32_t)iy : 0u;
      } else {
        cy = (uint32_t)iy & ((uint32_t)G1 - 1u);
      }
      const uint32_t row = in.cell_base + cy * (uint32_t)G0;
      // All six words are loaded unconditionally from cells inside the grid
      // (a column or row outside it reads the centre column or row 0) and
      // then masked: no branch per load.
      const uint32_t a0 = row + (vx0 ? x0 : x1), a1 = row + x1, a2 = row + (vx2 ? x2 : x1);
      const uint32_t s0r = ld<kO>(cell_start, a0), s1r = ld<kO>(cell_start, a1), s2r = ld<kO>(cell_start, a2);
      const uint32_t e0r = ld<kO>(cell_end, a0), e1r = ld<kO>(cell_end, a1), e2r = ld<kO>(cell_end, a2);
      const uint32_t s0 = ok && vx0 ? s0r : kEmpty;
      const uint32_t s1 = ok ? s1r : kEmpty;
      const uint32_t s2 = ok && vx2 ? s2r : kEmpty;
      const bool n0 = s0 != kEmpty, n1 = s1 != kEmpty, n2 = s2 != kEmpty;
      const uint32_t e0 = n0 ? min(e0r, n) : 0u;
      const uint32_t e1 = n1 ? min(e1r, n) : 0u;
      const uint32_t e2 = n2 ? min(e2r, n) : 0u;
      rs[y] = n0 ? s0 : (n1 ? s1 : (n2 ? s2 : 0u));
      re[y] = n2 ? e2 : (n1 ? e1 : e0);
      b0[y] = n0 ? e0 : kNoBoundary;
      b1[y] = n1 ? e1 : kNoBoundary;
    }
    Cand cb[3][kBatchN];
    if (kMode == kBatch) {
      // index 0 stands in for a candidate past the row's end (never used)
#pragma unroll
      for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int k = 0; k < kBatchN; ++k) {
          const uint32_t j = rs[y] + (uint32_t)k;
          candidate_load<kRadius, kOne, false, kO>(c, j < re[y] ? j : 0u, pos, type, radius, d, cb[y][k]);
        }
    }
    if (kMode == kList && !c.attract && re[0] - rs[0] <= 64u && re[1] - rs[1] <= 64u && re[2] - rs[2] <= 64u) {
      // Rows of at most 64 candidates: first one bit per candidate that may
      // touch (the pair law's own prefilter, from position and radius), then
      // the law over the set bits of all three rows in row order, so a wave
      // runs the law as often as its busiest lane has contacts.  Candidates
      // that exert no force add +0, so skipping them leaves every sum
      // unchanged; a row's partial sum is flushed when the walk moves on to
      // a later row.  (With attraction every candidate within reach pulls,
      // so the rows are walked one candidate at a time instead.)
      // The three rows are walked together, the k-th candidate of each
      // loaded at once (three loads in flight; a row already walked
      // re-reads the particle's own slot, which is never marked).
      // (the marks are kept as 32-bit halves: bit k of a row's mask is the
      // same bit for every lane, so setting it is one select and one or)
      uint32_t hlo[3] = {0u, 0u, 0u}, hhi[3] = {0u, 0u, 0u};
      const uint32_t len0 = re[0] - rs[0], len1 = re[1] - rs[1], len2 = re[2] - rs[2];
      const uint32_t L = max(len0, max(len1, len2));
      // positions k of the rows' k-th candidates, in [k0, k1), marking into
      // word w (bit k - 32 w)
      auto mark = [&](uint32_t k0, uint32_t k1, uint32_t (&w)[3], uint32_t wbase) {
        for (uint32_t k = k0; k < k1; ++k) {
          Cand b[3];
          const uint32_t j0 = k < len0 ? rs[0] + k : i, j1 = k < len1 ? rs[1] + k : i, j2 = k < len2 ? rs[2] + k : i;
          candidate_load<kRadius, kOne, false, kO>(c, j0, pos, type, radius, d, b[0]);
          candidate_load<kRadius, kOne, false, kO>(c, j1, pos, type, radius, d, b[1]);
          candidate_load<kRadius, kOne, false, kO>(c, j2, pos, type, radius, d, b[2]);
          const uint32_t jy[3] = {j0, j1, j2};
          const uint32_t bit = 1u << (k - wbase);
#pragma unroll
          for (int y = 0; y < 3; ++y) {
            const float rx = b[y].p.x - q.p.x, ry = b[y].p.y - q.p.y;
            const float d2 = rx * rx + ry * ry;
            const float R = q.rA + b[y].r;
            const bool maybe = !(d2 > (R * R) * 1.00000095367431640625f) || R < 1.0842022e-19f;
            w[y] |= maybe && jy[y] != i ? bit : 0u;
          }
        }
      };
      mark(0u, min(L, 32u), hlo, 0u);
      if (L > 32u) mark(32u, L, hhi, 32u);
      uint64_t h[3];
#pragma unroll
      for (int y = 0; y < 3; ++y) h[y] = ((uint64_t)hhi[y] << 32) | hlo[y];
      RowSum rsum;
      rsum.b0 = b0[0];
      rsum.b1 = b1[0];
      int row = 0;
      // the particle's velocity, spin and the pressure terms' reciprocal,
      // only where a candidate may touch
      if (h[0] | h[1] | h[2]) {
        if (!have_dyn) load_own_dyn();
        q.rden = 1.0 / q.den;
      }
      while (h[0] | h[1] | h[2]) {
        const bool in0 = h[0] != 0, in1 = !in0 && h[1] != 0;
        const int y = in0 ? 0 : (in1 ? 1 : 2);
        const uint64_t hy = in0 ? h[0] : (in1 ? h[1] : h[2]);
        const uint32_t j = (in0 ? rs[0] : (in1 ? rs[1] : rs[2])) + (uint32_t)__builtin_ctzll(hy);
        const uint64_t rest = hy & (hy - 1u);
        h[0] = in0 ? rest : h[0];
        h[1] = in1 ? rest : h[1];
        h[2] = (!in0 && !in1) ? rest : h[2];
        if (y != row) {
          fx += rsum.cfx;
          fy += rsum.cfy;
          rsum = RowSum{};
          rsum.b0 = in0 ? b0[0] : (in1 ? b0[1] : b0[2]);
          rsum.b1 = in0 ? b1[0] : (in1 ? b1[1] : b1[2]);
          row = y;
        }
        candidate<kRadius, kOne, true, kO, true>(c, q, j, false, pos, d, type, radius, rsum, fx, fy, moment, pres);
      }
      fx += rsum.cfx;
      fy += rsum.cfy;
    } else {
    if (!have_dyn) load_own_dyn();
#pragma unroll
    for (int y = 0; y < 3; ++y) {
      RowSum rsum;
      rsum.b0 = b0[y];
      rsum.b1 = b1[y];
      const uint32_t end = re[y];
      if (kMode == kPrefetch) {
        // each candidate's position and radius fetched one iteration ahead
        // (pays off with many candidates per row, costs registers)
        uint32_t j = rs[y];
        Cand nx{};
        if (j < end) candidate_load<kRadius, kOne, false, kO>(c, j, pos, type, radius, d, nx);
        for (; j < end; ++j) {
          const Cand b = nx;
          if (j + 1u < end) candidate_load<kRadius, kOne, false, kO>(c, j + 1u, pos, type, radius, d, nx);
          candidate_at<kRadius, kOne, false, kO>(c, q, j, j == i, b, d, type, rsum, fx, fy, moment, pres);
        }
      } else {
        uint32_t j = rs[y];
        if (kMode == kBatch) {
#pragma unroll
          for (int k = 0; k < kBatchN; ++k, ++j)
            if (j < end)
              candidate_at<kRadius, kOne, false, kO>(c, q, j, j == i, cb[y][k], d, type, rsum, fx, fy, moment, pres);
        }
        for (; j < end; ++j)
          candidate<kRadius, kOne, kMode == kPlain, kO>(c, q, j, j == i, pos, d, type, radius, rsum, fx, fy, moment,
                                                     pres);
      }
      fx += rsum.cfx;
      fy += rsum.cfy;
    }
    }
  } else {
    // wrap mode at the grid's x edge: the three cells one by one, in the
    // reference order
    if (!have_dyn) load_own_dyn();
    for (int y = -1; y <= 1; ++y) {
      const uint32_t cy = ((uint32_t)gy + (uint32_t)y) & ((uint32_t)G1 - 1u);
      const uint32_t row = in.cell_base + cy * (uint32_t)G0;
      const uint32_t xs[3] = {x0, x1, x2};
      for (int sg = 0; sg < 3; ++sg) {
        const uint32_t st = cell_start[row + xs[sg]];
        if (st == kEmpty) continue;
        const uint32_t en = min(cell_end[row + xs[sg]], n);
        RowSum rsum;  // one cell: no boundary inside
        rsum.b0 = kNoBoundary;
        rsum.b1 = kNoBoundary;
        for (uint32_t j = st; j < en; ++j)
          candidate<kRadius, kOne, false, kO>(c, q, j, j == i, pos, d, type, radius, rsum, fx, fy, moment, pres);
        fx += rsum.cfx;
        fy += rsum.cfy;
      }
    }
  }
  if (kControl) {
    float rx, ry, rm;
    control_pair(c, q, *ctl.state, fx, fy, moment, pres, rx, ry, rm);
    ctl.react_x[i] = rx;
    ctl.react_y[i] = ry;
    ctl.react_m[i] = rm;
  }
  if (!have_dyn) load_own_dyn();  // (dead, and dropped, where the caller does not use them)
  own.v = q.v;
  own.w = q.w;
  fx_out = fx;
  fy_out = fy;
  moment_out = moment;
  pres_out = pres;
}

#define DEM_COLLIDE_PARAMS                                                                    \
  StepConst c, const float2* __restrict__ pos, const float2* __restrict__ vel,              \
      float2* __restrict__ acc, const float* __restrict__ omega, float* __restrict__ alpha, \
      const uint32_t* __restrict__ type, const float* __restrict__ radius,                  \
      const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,       \
      float* __restrict__ pressure, ControlOut ctl, const uint32_t* __restrict__ gat,                \
      uint32_t* __restrict__ quiet, uint32_t n

template <bool kRadius, int kMat, int kMode, bool kControl, bool kO>
__device__ __forceinline__ void collide_kernel_body(DEM_COLLIDE_PARAMS) {
  const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Own own;
  float fx, fy, moment, pres;
  const Dyn d{vel, omega, gat};
  bool done = false;
  if constexpr (kMat == kInstances) {
    // waves inside one instance (all but the few that straddle a boundary)
    const uint32_t t = clamp_idx(type[i], c.num_types);
    const uint32_t t0 = __builtin_amdgcn_readfirstlane(t);
    if (__ballot(t != t0) == 0) {
      collide_particle<kRadius, kMat, kMode, kControl, kO, true>(c, i, pos, d, type, radius, cell_start, cell_end, n,
                                                                 own, ctl, fx, fy, moment, pres, t0);
      done = true;
    }
  }
  if (!done)
    collide_particle<kRadius, kMat, kMode, kControl, kO>(c, i, pos, d, type, radius, cell_start, cell_end, n, own,
                                                         ctl, fx, fy, moment, pres);
  // A wave whose discs all feel no force stores zeros without the three
  // divisions: the sums start at +0 and are never -0, so fx, fy and moment
  // are +0 exactly when nothing acted, and +0 over a positive finite mass or
  // inertia is +0.  With a quiet array (the owned simulation's in-place
  // step, whose integrate reads it) the wave records that instead of
  // storing the zeros: quiet[i / 64] = 1 when acc and alpha of the wave's 64
  // slots are +0 and were not written, else 0.
  // (with a radius array the inertia is t / 5, t = ((2 mass) r) r, and t / 5
  // rounds to a positive finite value exactly when 3 * 2^-149 <= t <= FLT_MAX:
  // the check needs no division, which only a wave with a force then does)
  bool inertia_ok;
  if constexpr (kRadius) {
    const float t = ((2.0f * own.mass) * own.r) * own.r;
    inertia_ok = t >= __uint_as_float(3u) && t <= 3.402823466e38f;
  } else {
    inertia_ok = own.inertia > 0.0f && own.inertia <= 3.402823466e38f;
  }
  const bool still = fx == 0.0f && fy == 0.0f && moment == 0.0f && own.mass > 0.0f &&
                     own.mass <= 3.402823466e38f && inertia_ok;
  const bool wave_still = __ballot(!still) == 0;
  if (quiet && __lane_id() == 0) quiet[i / (uint32_t)kWave] = wave_still ? 1u : 0u;
  if (wave_still) {
    if (!quiet) {
      acc[i] = make_float2(0.0f, 0.0f);
      alpha[i] = 0.0f;
    }
  } else {
    acc[i] = make_float2(fx / own.mass, fy / own.mass);
    alpha[i] = moment / own.inertia;
  }
  pressure[i] = pres;
}

template <bool kRadius, int kMat, int kMode, bool kControl, bool kO>
__global__ __launch_bounds__(256) void k_collide(DEM_COLLIDE_PARAMS) {
  collide_kernel_body<kRadius, kMat, kMode, kControl, kO>(c, pos, vel, acc, omega, alpha, type, radius, cell_start,
                                                          cell_end, pressure, ctl, gat, quiet, n);
}

// The one-material batched mode needed 65 VGPRs with 64-bit addresses; held to 64 (8
// waves per SIMD) it does not spill.  The other batched layouts would spill
// 12-19 registers under the same bound, so they keep k_collide.
template <bool kRadius>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_collide_batch(DEM_COLLIDE_PARAMS) {
  collide_kernel_body<kRadius, kOneMat, kBatch, false, true>(c, pos, vel, acc, omega, alpha, type, radius,
                                                             cell_start, cell_end, pressure, ctl, gat, quiet, n);
}
#undef DEM_COLLIDE_PARAMS

// kO (32-bit offsets) for states whose arrays stay under 4 GiB; larger ones
// (or DEM_STEP_WIDE) run the plain mode with 64-bit addresses.
template <bool kRadius, int kMat, int kMode, bool kControl, bool kO>
constexpr auto collide_kernel() {
  if constexpr (!kO)
    return &k_collide<kRadius, kMat, kPlain, kControl, false>;
  else if constexpr (kMat == kOneMat && kMode == kBatch && !kControl)
    return &k_collide_batch<kRadius>;
  else
    return &k_collide<kRadius, kMat, kMode, kControl, true>;
}

// ------------------------------------------------------------- integrate --
__device__ __forceinline__ float wrap_theta(float th) {
  const double twopi = 2 * M_PI;
  if ((double)th > twopi) th = (float)((double)th - twopi * floor((double)th / twopi));
  if (th < 0.0f) th = (float)((double)th - twopi * floor((double)th / twopi));
  return th;
}

// integrateSystemD (src/particles_kernel.cuh:580-627) with collideBoundary
// (:388-456) inlined for the bottom wall, on one particle: p, v, w, th are
// updated in place; pres gains the bottom wall's term when it is hit (the
// return value says whether the bottom-wall law ran).  With pres_src, pres
// is first loaded from it, and only on that branch.
__device__ __forceinline__ bool integrate_particle(const StepConst& c, const Material& m, const Instance& in, float r,
                                                   float mass, float inertia, float2 a, float alpha_i, float2& p,
                                                   float2& v, float& w, float& th, float& pres,
                                                   const float* pres_src) {
  const float dt = c.dt;
  v.x = v.x + (c.gx + a.x) * dt;
  v.y = v.y + (c.gy + a.y) * dt;
  p.x = p.x + v.x * dt;
  p.y = p.y + v.y * dt;
  w = w + alpha_i * dt;
  th = th + w * dt;
  th = wrap_theta(th);

  const float bd = -m.bdamp;
  const float Lx = in.box[0], Ly = in.box[1];
  if (p.x > Lx - r) {
    p.x = Lx - r;
    v.x = v.x * bd;
  }
  if (p.x < r) {
    p.x = r;
    v.x = v.x * bd;
  }
  if (!(c.flags & DEM_OPEN_TOP) && p.y > Ly - r) {
    p.y = Ly - r;
    v.y = v.y * bd;
  }
  if (p.y < r) {
    // collideBoundary(pos, vel, omega, type, (pos.x, 0), wallAlpha, pressure)
    float wallAlpha = 0.0f;
    float pr = pres_src ? *pres_src : pres;
    const float rx = p.x - p.x, ry = 0.0f - p.y;
    const float dist = sqrtf(rx * rx + ry * ry);
    float wfx = 0.0f, wfy = 0.0f;
    if (dist < r) {
      const float nx = rx / dist, ny = ry / dist;
      const float tx = -ny, ty = nx;
      const float rvx = -v.x, rvy = -v.y;
      const float vn = rvx * nx + rvy * ny;
      const float vtx = rvx - vn * nx, vty = rvy - vn * ny;
      const float kt = series(m.kt, c.bkt);
      const float s = -mass;
      wfx = s * c.gx;
      wfy = s * c.gy;
      if (vn < 0.0f) {
        const float s2 = m.cn * vn;
        wfx += s2 * nx;
        wfy += s2 * ny;
      }
      const float flen = sqrtf(wfx * wfx + wfy * wfy);
      pr = add_pressure(pr, flen, pressure_den(r));
      if (m.mu != 0.0f) {
        const float sr = r * w;
        const float cvx = vtx - sr * tx, cvy = vty - sr * ty;
        float ftx = (kt * cvx) * dt, fty = (kt * cvy) * dt;
        const float ftmax = m.mu * flen;
        if (!sqrt_below(ftx * ftx + fty * fty, ftmax)) {
          const float cl = sqrtf(cvx * cvx + cvy * cvy);
          ftx = (ftmax * cvx) / cl;
          fty = (ftmax * cvy) / cl;
        }
        wfx = ftx;
        wfy = fty;
        wallAlpha = wallAlpha + (r * (ftx * tx + fty * ty)) / inertia;
      }
    }
    const float ax = wfx / mass, ay = wfy / mass;
    v.x = v.x + ax * dt;
    v.y = v.y + ay * dt;
    p.x = p.x + (ax * dt) * dt;
    p.y = p.y + (ay * dt) * dt;
    w = w + wallAlpha * dt;
    th = th + (wallAlpha * dt) * dt;
    p.y = r;
    v.y = v.y * bd;
    pres = pr;
    return true;
  }
  return false;
}

// The next step's hash of an integrated particle (NextKeys): the cell key of
// its new position p, and whether it left the cell of its sorted slot (the
// state stays in this order until the next sort).  The key the particle is
// sorted by is the hash of its position before integrate, p_old, so it is
// recomputed here rather than read.
__device__ __forceinline__ void next_key(const StepConst& c, const NextKeys& nk, float2 p, float2 p_old,
                                         const Instance& in, uint32_t i, uint32_t n) {
  const uint32_t k = cell_key_in(c, in, p);
  const bool mover = k != cell_key_in(c, in, p_old);
  if (mover) nk.key[i] = k;  // a stayer's key is its sorted key
  const uint64_t moved = __ballot(mover);
  if (__lane_id() == 0) {  // lane 0 holds the wave's first index, a multiple of 64
    const uint32_t w0 = i / 32u, nwords = (n + 31u) / 32u;
    nk.bits[w0] = (uint32_t)moved;
    if (w0 + 1u < nwords) nk.bits[w0 + 1u] = (uint32_t)(moved >> 32);
    const uint32_t cnt = (uint32_t)__popcll(moved);
    if (cnt) atomicAdd(&nk.tile_count[i / kResortTile], cnt);
  }
}

// integrateSystemD over all particles.  The input and output arrays may be
// the same (in place, as the reference) or different (the owned
// simulation writes the integrated state back to its home buffer); each
// lane reads and writes only its own slot.
template <bool kRadius, int kMat, bool kNext, int kPer>
__global__ __launch_bounds__(256) void k_integrate(
    StepConst c, const float2* pos_in, const float2* vel_in, const float* theta_in, const float* omega_in,
    float2* pos_out, float2* vel_out, float* theta_out, float* omega_out, const float2* __restrict__ acc,
    const float* __restrict__ alpha, const uint32_t* __restrict__ type, const float* __restrict__ radius,
    float* __restrict__ pressure, NextKeys nk, const uint32_t* __restrict__ gat,
    const uint32_t* __restrict__ quiet, uint32_t n) {
  constexpr bool kMulti = kMat == kInstances;
  const bool in_place = !gat && omega_out == omega_in && theta_out == theta_in;
  // kPer particles a lane, as k_hash.  The inputs of all kPer particles are
  // loaded before any output is stored (the in and out arrays may be the
  // same, so the compiler could not move a later particle's loads above an
  // earlier one's stores); each lane reads and writes only its own slots.
  const uint32_t base = blockIdx.x * blockDim.x * kPer + threadIdx.x;
  uint32_t t[kPer];
  float2 p_old[kPer], v[kPer], a[kPer];
  float w[kPer], th[kPer], al[kPer], rad[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const uint32_t i = base + (uint32_t)k * blockDim.x;
    if (i >= n) continue;
    t[k] = kMat == kOneMat ? 0u : clamp_idx(type[i], c.num_types);
    // velocity, spin and angle through the gather index when one is given
    const uint32_t g = gat ? gat[i] : i;
    p_old[k] = pos_in[i];
    v[k] = vel_in[g];
    w[k] = omega_in[g];
    th[k] = theta_in[g];
    // a run of 64 slots that collide marked quiet has acc = alpha = +0 (not
    // stored); the wave's 64 slots are one such run (one scalar load)
    const bool still = quiet && quiet[__builtin_amdgcn_readfirstlane(i) / (uint32_t)kWave] != 0u;
    a[k] = still ? make_float2(0.0f, 0.0f) : acc[i];
    al[k] = still ? 0.0f : alpha[i];
    rad[k] = kRadius ? radius[i] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const uint32_t i = base + (uint32_t)k * blockDim.x;
    if (i >= n) continue;
    auto run = [&](const Material& m, const Instance& in) {
      float r, mass, inertia;
      body<kRadius>(m, rad + k, 0u, r, mass, inertia);
      float2 p = p_old[k], vv = v[k];
      float ww = w[k], tt = th[k];
      float pres = 0.0f;
      if (integrate_particle(c, m, in, r, mass, inertia, a[k], al[k], p, vv, ww, tt, pres, pressure + i))
        pressure[i] = pres;
      pos_out[i] = p;
      vel_out[i] = vv;
      // in place, a spin or angle whose bits did not change is not stored
      // again (a disc at rest: no spin, no moment)
      if (!in_place || __float_as_uint(ww) != __float_as_uint(w[k])) omega_out[i] = ww;
      if (!in_place || __float_as_uint(tt) != __float_as_uint(th[k])) theta_out[i] = tt;
      if (kNext) next_key(c, nk, p, p_old[k], in, i, n);
    };
    if constexpr (kMulti) {
      // a wave inside one instance reads its material row and geometry as
      // scalars
      const uint32_t t0 = __builtin_amdgcn_readfirstlane(t[k]);
      if (__ballot(t[k] != t0) == 0) {
        const Material m = c.mat[t0];
        run(m, c.inst[clamp_idx(m.inst, c.num_instances)]);
        continue;
      }
    }
    const Material m = c.mat[t[k]];
    run(m, kMulti ? c.inst[clamp_idx(m.inst, c.num_instances)] : c.inst0);
  }
}

// collide + integrate + the next step's hash in one pass over the sorted
// particles (the owned simulation's step): forces from the sorted state
// (pos, vel, omega, theta in the *_in arrays), accelerations and pressure
// written as collide writes them, and the integrated state written to the
// *_out arrays (the unsorted buffer of this step, dead after reorder), so no
// lane overwrites a neighbour's input.  The arithmetic is that of k_collide
// followed by k_integrate.
template <bool kRadius, int kMat, int kMode, bool kControl>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_collide_integrate(
    StepConst c, const float2* __restrict__ pos_in, const float2* __restrict__ vel_in,
    const float* __restrict__ omega_in, const float* __restrict__ theta_in, const uint32_t* __restrict__ type,
    const float* __restrict__ radius, const uint32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_end, float2* __restrict__ acc, float* __restrict__ alpha,
    float* __restrict__ pressure, float2* __restrict__ pos_out, float2* __restrict__ vel_out,
    float* __restrict__ omega_out, float* __restrict__ theta_out, NextKeys nk, ControlOut ctl, uint32_t n) {
  const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Own own;
  float fx, fy, moment, pres;
  const Dyn d{vel_in, omega_in, nullptr};
  collide_particle<kRadius, kMat, kMode, kControl, false>(c, i, pos_in, d, type, radius, cell_start, cell_end, n, own,
                                                         ctl, fx, fy, moment, pres);
  const float2 a = make_float2(fx / own.mass, fy / own.mass);
  const float al = moment / own.inertia;
  acc[i] = a;
  alpha[i] = al;
  float2 p = own.p, v = own.v;
  float w = own.w, th = theta_in[i];
  (void)integrate_particle(c, own.m, own.in, own.r, own.mass, own.inertia, a, al, p, v, w, th, pres, nullptr);
  pressure[i] = pres;
  pos_out[i] = p;
  vel_out[i] = v;
  omega_out[i] = w;
  theta_out[i] = th;
  next_key(c, nk, p, own.p, own.in, i, n);
}

inline dim3 grid_for(uint32_t n, int block) { return dim3((n + (uint32_t)block - 1u) / (uint32_t)block); }

}  // namespace

// Particles a lane of the streaming kernels (hash, reorder, integrate): 1, 2
// or 4 (the autotuner's choice); anything else runs as 1.
inline int per_of(int per) { return per == 2 || per == 4 ? per : 1; }
#define DEM_PER(MACRO, PER) \
  do {                      \
    if ((PER) == 4) MACRO(4); \
    else if ((PER) == 2) MACRO(2); \
    else MACRO(1);          \
  } while (0)

hipError_t launch_hash(const StepConst& c, hipStream_t s, int block, const float2* pos,
                       const uint32_t* type, uint32_t* hash, uint32_t* index, uint32_t n, int per) {
  if (n == 0) return hipSuccess;
  per = per_of(per);
  const dim3 grid = grid_for((n + (uint32_t)per - 1u) / (uint32_t)per, block);
#define DEM_HASH(P)                                                                                        \
  do {                                                                                                    \
    if (type && c.num_instances > 1)                                                                      \
      hipLaunchKernelGGL((k_hash<true, P>), grid, dim3(block), 0, s, c, pos, type, hash, index, n);       \
    else                                                                                                  \
      hipLaunchKernelGGL((k_hash<false, P>), grid, dim3(block), 0, s, c, pos, type, hash, index, n);      \
  } while (0)
  DEM_PER(DEM_HASH, per);
#undef DEM_HASH
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, uint32_t* p, uint32_t value, uint32_t count) {
  if (count == 0) return hipSuccess;
  const int vec = (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
  const uint32_t threads = (count + 3u) / 4u;
  hipLaunchKernelGGL(k_fill, grid_for(threads, 256), dim3(256), 0, s, p, value, count, vec);
  return hipGetLastError();
}

hipError_t launch_reorder(hipStream_t s, int block, uint32_t* cell_start, uint32_t* cell_end,
                          float2* s_pos, float2* s_vel, float* s_theta, float* s_omega,
                          uint32_t* s_id, uint32_t* s_loc, uint32_t* s_type, float* s_radius,
                          const uint32_t* hash, const uint32_t* index, const float2* o_pos,
                          const float2* o_vel, const float* o_theta, const float* o_omega,
                          const uint32_t* o_id, const uint32_t* o_type, const float* o_radius,
                          uint32_t n, uint32_t num_cells, bool full, int per) {
  if (n == 0) return hipSuccess;
  per = per_of(per);
  const dim3 grid = grid_for((n + (uint32_t)per - 1u) / (uint32_t)per, block);
#define DEM_REORDER(R, F, P)                                                                                 \
  hipLaunchKernelGGL((k_reorder<R, F, P>), grid, dim3(block), 0, s, cell_start, cell_end, s_pos, s_vel,       \
                     s_theta, s_omega, s_id, s_loc, s_type, s_radius, hash, index, o_pos, o_vel, o_theta,     \
                     o_omega, o_id, o_type, o_radius, n, num_cells)
#define DEM_REORDER_P(P)                                                     \
  do {                                                                       \
    if (s_radius) {                                                          \
      if (full) DEM_REORDER(true, true, P); else DEM_REORDER(true, false, P); \
    } else {                                                                 \
      if (full) DEM_REORDER(false, true, P); else DEM_REORDER(false, false, P); \
    }                                                                        \
  } while (0)
  DEM_PER(DEM_REORDER_P, per);
#undef DEM_REORDER_P
#undef DEM_REORDER
  return hipGetLastError();
}

// Kernel variant dispatch: kRadius x material layout x (candidate mode,
// control), where the controlled particle always uses the plain mode.
inline int mat_kind(const StepConst& c) {
  return c.num_instances > 1 ? kInstances : (c.num_types == 1 ? kOneMat : kManyMat);
}
#define DEM_DISPATCH_RM(MACRO, MODE, X)                                  \
  do {                                                                   \
    const int mk_ = mat_kind(c);                                         \
    if (radius) {                                                        \
      if (mk_ == kInstances) MACRO(true, kInstances, MODE, X);           \
      else if (mk_ == kOneMat) MACRO(true, kOneMat, MODE, X);            \
      else MACRO(true, kManyMat, MODE, X);                               \
    } else {                                                             \
      if (mk_ == kInstances) MACRO(false, kInstances, MODE, X);          \
      else if (mk_ == kOneMat) MACRO(false, kOneMat, MODE, X);           \
      else MACRO(false, kManyMat, MODE, X);                              \
    }                                                                    \
  } while (0)
#define DEM_DISPATCH(MACRO)                                              \
  do {                                                                   \
    if (ctl) DEM_DISPATCH_RM(MACRO, kPlain, true);                       \
    else if (mode == kPrefetch) DEM_DISPATCH_RM(MACRO, kPrefetch, false);\
    else if (mode == kBatch) DEM_DISPATCH_RM(MACRO, kBatch, false);      \
    else if (mode == kList) DEM_DISPATCH_RM(MACRO, kList, false);        \
    else DEM_DISPATCH_RM(MACRO, kPlain, false);                          \
  } while (0)

hipError_t launch_collide(const StepConst& c, hipStream_t s, int block, const float2* pos,
                          const float2* vel, float2* acc, const float* omega, float* alpha,
                          const uint32_t* type, const float* radius, const uint32_t* cell_start,
                          const uint32_t* cell_end, float* pressure, uint32_t n, int mode,
                          const ControlOut* ctl, const uint32_t* gat, uint32_t* quiet) {
  if (n == 0) return hipSuccess;
  if (quiet && block % kWave) return hipErrorInvalidValue;  // the flags cover aligned 64-particle runs
  const ControlOut co = ctl ? *ctl : ControlOut{};
  // 32-bit byte offsets reach every element while n * 8 B and cells * 4 B
  // stay within 4 GiB
  const bool o32 = !(mode & kWideMode) && n <= (1u << 29) && c.cells <= (1u << 30);
  mode &= ~kWideMode;
#define DEM_COLLIDE(R, M, P, X)                                                                             \
  do {                                                                                                      \
    if (o32)                                                                                                \
      hipLaunchKernelGGL((collide_kernel<R, M, P, X, true>()), grid_for(n, block), dim3(block), 0, s, c, pos, \
                         vel, acc, omega, alpha, type, radius, cell_start, cell_end, pressure, co, gat, quiet, n); \
    else                                                                                                    \
      hipLaunchKernelGGL((collide_kernel<R, M, P, X, false>()), grid_for(n, block), dim3(block), 0, s, c,     \
                         pos, vel, acc, omega, alpha, type, radius, cell_start, cell_end, pressure, co, gat, quiet, \
                         n);                                                                                \
  } while (0)
  DEM_DISPATCH(DEM_COLLIDE);
#undef DEM_COLLIDE
  return hipGetLastError();
}

hipError_t launch_collide_integrate(const StepConst& c, hipStream_t s, int block, const float2* pos_in,
                                    const float2* vel_in, const float* omega_in, const float* theta_in,
                                    const uint32_t* type, const float* radius, const uint32_t* cell_start,
                                    const uint32_t* cell_end, float2* acc, float* alpha, float* pressure,
                                    float2* pos_out, float2* vel_out, float* omega_out, float* theta_out,
                                    const NextKeys& next, uint32_t n, int mode, const ControlOut* ctl) {
  if (n == 0) return hipSuccess;
  if (block % kWave) return hipErrorInvalidValue;  // waves must cover aligned 64-particle runs
  const ControlOut co = ctl ? *ctl : ControlOut{};
#define DEM_CI(R, M, P, X)                                                                                      \
  hipLaunchKernelGGL((k_collide_integrate<R, M, P, X>), grid_for(n, block), dim3(block), 0, s, c, pos_in, vel_in, \
                     omega_in, theta_in, type, radius, cell_start, cell_end, acc, alpha, pressure, pos_out,        \
                     vel_out, omega_out, theta_out, next, co, n)
  // the fused kernel has no batched mode: with integrate's registers on top
  // of the batch it spills heavily; it loads one candidate at a time instead
  if (ctl) DEM_DISPATCH_RM(DEM_CI, kPlain, true);
  else if (mode == kPrefetch) DEM_DISPATCH_RM(DEM_CI, kPrefetch, false);
  else if (mode == kList) DEM_DISPATCH_RM(DEM_CI, kList, false);
  else DEM_DISPATCH_RM(DEM_CI, kPlain, false);
#undef DEM_CI
  return hipGetLastError();
}
#undef DEM_DISPATCH
#undef DEM_DISPATCH_RM

// ---- controlled particle: reaction sum and move ------------------------------
// reduceVecD (src/particles_kernel.cuh:552-578) on the three reaction
// vectors: block b sums values [512 b, 512 b + 512) (zeros past n) by the
// same pairwise tree.
__global__ __launch_bounds__(kReactBlock) void k_reduce_reaction(const float* __restrict__ rx,
                                                                const float* __restrict__ ry,
                                                                const float* __restrict__ rm, float* __restrict__ bx,
                                                                float* __restrict__ by, float* __restrict__ bm,
                                                                uint32_t n) {
  __shared__ float cx[kReactBlock], cy[kReactBlock], cm[kReactBlock];
  const uint32_t t = threadIdx.x, i = blockIdx.x * kReactBlock + t;
  cx[t] = i < n ? rx[i] : 0.0f;
  cy[t] = i < n ? ry[i] : 0.0f;
  cm[t] = i < n ? rm[i] : 0.0f;
  __syncthreads();
  for (uint32_t h = kReactBlock / 2; h != 0; h /= 2) {
    if (t < h) {
      cx[t] += cx[t + h];
      cy[t] += cy[t + h];
      cm[t] += cm[t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    bx[blockIdx.x] = cx[0];
    by[blockIdx.x] = cy[0];
    bm[blockIdx.x] = cm[0];
  }
}

// The block sums added in order (src/functions.cu:505-513, there on the
// host), one lane per quantity, then the kinematic move of the controlled
// particle (src/main.cu:510-514) when `move` is set.
__global__ __launch_bounds__(64) void k_control_finish(const float* __restrict__ bx, const float* __restrict__ by,
                                                      const float* __restrict__ bm, uint32_t blocks,
                                                      float* __restrict__ reaction, dem_controlled* state, int move) {
  const uint32_t t = threadIdx.x;
  if (t < 3) {
    const float* b = t == 0 ? bx : (t == 1 ? by : bm);
    float acc = 0.0f;
    for (uint32_t k = 0; k < blocks; ++k) acc += b[k];
    reaction[t] = acc;
  }
  if (t == 0 && move) {
    dem_controlled& C = *state;
    C.pos[0] = (float)((double)C.pos[0] + C.move[0]);
    C.pos[1] = (float)((double)C.pos[1] + C.move[1]);
    if ((double)C.pos[1] < C.wrap_below_y) C.pos[1] = C.reset_y;
    if ((double)C.pos[0] > C.wrap_above_x) C.pos[0] = C.reset_x;
  }
}

hipError_t launch_control_finish(hipStream_t s, const ControlOut& ctl, float* bx, float* by, float* bm, uint32_t n,
                                 float* reaction, dem_controlled* state, bool move) {
  const uint32_t blocks = (n + kReactBlock - 1u) / kReactBlock;
  if (n) hipLaunchKernelGGL(k_reduce_reaction, dim3(blocks), dim3(kReactBlock), 0, s, ctl.react_x, ctl.react_y,
                            ctl.react_m, bx, by, bm, n);
  hipLaunchKernelGGL(k_control_finish, dim3(1), dim3(64), 0, s, bx, by, bm, n ? blocks : 0u, reaction, state,
                     move ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_integrate(const StepConst& c, hipStream_t s, int block, float2* pos, float2* vel,
                            const float2* acc, float* theta, float* omega, const float* alpha,
                            const uint32_t* type, const float* radius, float* pressure, uint32_t n,
                            const NextKeys* next, int per) {
  return launch_integrate_to(c, s, block, pos, vel, theta, omega, pos, vel, theta, omega, acc, alpha, type, radius,
                             pressure, n, next, nullptr, per);
}

hipError_t launch_integrate_to(const StepConst& c, hipStream_t s, int block, const float2* pos_in,
                               const float2* vel_in, const float* theta_in, const float* omega_in, float2* pos_out,
                               float2* vel_out, float* theta_out, float* omega_out, const float2* acc,
                               const float* alpha, const uint32_t* type, const float* radius, float* pressure,
                               uint32_t n, const NextKeys* next, const uint32_t* gat, int per,
                               const uint32_t* quiet) {
  if (n == 0) return hipSuccess;
  if (block % kWave) return hipErrorInvalidValue;  // waves must cover aligned 64-particle runs
  per = per_of(per);
  const dim3 grid = grid_for((n + (uint32_t)per - 1u) / (uint32_t)per, block);
  const NextKeys nk = next ? *next : NextKeys{};
#define DEM_INTEGRATE(R, M, X, P)                                                                         \
  hipLaunchKernelGGL((k_integrate<R, M, X, P>), grid, dim3(block), 0, s, c, pos_in, vel_in, theta_in,    \
                     omega_in, pos_out, vel_out, theta_out, omega_out, acc, alpha, type, radius, pressure, nk, \
                     gat, quiet, n)
#define DEM_INTEGRATE_X(R, M, P)            \
  do {                                      \
    if (next) DEM_INTEGRATE(R, M, true, P); \
    else DEM_INTEGRATE(R, M, false, P);     \
  } while (0)
#define DEM_INTEGRATE_P(P)                                                \
  do {                                                                    \
    const int mk = mat_kind(c);                                           \
    if (radius) {                                                         \
      if (mk == kInstances) DEM_INTEGRATE_X(true, kInstances, P);         \
      else if (mk == kOneMat) DEM_INTEGRATE_X(true, kOneMat, P);          \
      else DEM_INTEGRATE_X(true, kManyMat, P);                            \
    } else {                                                              \
      if (mk == kInstances) DEM_INTEGRATE_X(false, kInstances, P);        \
      else if (mk == kOneMat) DEM_INTEGRATE_X(false, kOneMat, P);         \
      else DEM_INTEGRATE_X(false, kManyMat, P);                           \
    }                                                                     \
  } while (0)
  DEM_PER(DEM_INTEGRATE_P, per);
#undef DEM_INTEGRATE_P
#undef DEM_INTEGRATE_X
#undef DEM_INTEGRATE
  return hipGetLastError();
}

}  // namespace dem
