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
  if (!(g == g)) return 0;
  if (g >= 2147483648.0f) return 2147483647;
  if (g < -2147483648.0f) return -2147483647 - 1;
  return (int)g;
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

__device__ __forceinline__ double pressure_den(float r) {
  return 4 * M_PI * (double)r * (double)r;
}

// ---------------------------------------------------------------- calcHash --
// calcHashD (src/particles_kernel.cuh:133-151).
template <bool kMulti>
__global__ __launch_bounds__(256) void k_hash(StepConst c, const float2* __restrict__ pos,
                                              const uint32_t* __restrict__ type,
                                              uint32_t* __restrict__ hash,
                                              uint32_t* __restrict__ index, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 p = pos[i];
  const Instance in =
      kMulti ? c.inst[clamp_idx(c.mat[clamp_idx(type[i], c.num_types)].inst, c.num_instances)]
             : c.inst0;
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
  hash[i] = in.cell_base + cy * in.grid[0] + cx;
  index[i] = i;
}

// -------------------------------------------------------------- cell fill --
// cudaMemset(cellStart, 0xff) of src/functions.cu:371 as a kernel (a memset
// node in a captured graph is avoided on purpose).
__global__ __launch_bounds__(256) void k_fill4(uint4* __restrict__ p, uint32_t v, uint32_t n4) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) p[i] = make_uint4(v, v, v, v);
}

__global__ __launch_bounds__(256) void k_fill1(uint32_t* __restrict__ p, uint32_t v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// --------------------------------------------------------------- reorder --
// reorderDataAndFindCellStartD (src/particles_kernel.cuh:156-233).  The
// previous key is read from global memory (it is in the same cache line for
// all lanes but lane 0).  Keys >= num_cells, indices >= n and IDs >= n are
// skipped instead of being written out of bounds.
template <bool kRadius>
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
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
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
  if (k >= n) return;
  s_pos[i] = o_pos[k];
  s_vel[i] = o_vel[k];
  s_theta[i] = o_theta[k];
  s_omega[i] = o_omega[k];
  const uint32_t id = o_id[k];
  s_id[i] = id;
  if (id < n) s_loc[id] = i;
  s_type[i] = o_type[k];
  if (kRadius) s_radius[i] = o_radius[k];
}

// --------------------------------------------------------------- collide --
// collideD (src/particles_kernel.cuh:459-549) with collideCell (:345-385),
// collideSpheres (:235-307) and attractSpheres (:310-342) inlined.  The
// neighbour's velocity, spin and material are only loaded for actual
// contacts (the reference loads them for every candidate; the values used
// are the same).  Per-cell partial sums are kept, as collideCell returns one
// float2 per cell that collideD adds up.
template <bool kRadius, bool kMulti>
__global__ __launch_bounds__(256) void k_collide(
    StepConst c, const float2* __restrict__ pos, const float2* __restrict__ vel,
    float2* __restrict__ acc, const float* __restrict__ omega, float* __restrict__ alpha,
    const uint32_t* __restrict__ type, const float* __restrict__ radius,
    const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,
    float* __restrict__ pressure, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;

  const uint32_t tA = clamp_idx(type[i], c.num_types);
  const Material mA = c.mat[tA];
  const Instance in = kMulti ? c.inst[clamp_idx(mA.inst, c.num_instances)] : c.inst0;
  float rA, massA, inertiaA;
  body<kRadius>(mA, radius, i, rA, massA, inertiaA);
  const float2 p = pos[i];
  const float2 v = vel[i];
  const float w = omega[i];
  const double den = pressure_den(rA);
  const float dt = c.dt;
  const bool clamp = c.flags & DEM_GRID_CLAMP;

  int gx, gy;
  grid_pos(in, p, gx, gy);
  if (clamp) clamp_cell(in, gx, gy);

  float fx = 0.0f, fy = 0.0f, moment = 0.0f, pres = 0.0f;
  for (int y = -1; y <= 1; ++y) {
    for (int x = -1; x <= 1; ++x) {
      uint32_t cx, cy;
      if (clamp) {
        const int ix = gx + x, iy = gy + y;
        if (ix < 0 || iy < 0 || ix >= (int)in.grid[0] || iy >= (int)in.grid[1]) continue;
        cx = (uint32_t)ix;
        cy = (uint32_t)iy;
      } else {
        cx = ((uint32_t)gx + (uint32_t)x) & (in.grid[0] - 1u);
        cy = ((uint32_t)gy + (uint32_t)y) & (in.grid[1] - 1u);
      }
      const uint32_t key = in.cell_base + cy * in.grid[0] + cx;
      const uint32_t st = cell_start[key];
      if (st == kEmpty) continue;
      const uint32_t en = min(cell_end[key], n);
      float cfx = 0.0f, cfy = 0.0f;
      for (uint32_t j = st; j < en; ++j) {
        if (j == i) continue;
        const float2 pB = pos[j];
        uint32_t tB = 0;
        float rB;
        if (kRadius) {
          rB = radius[j];
        } else {
          tB = clamp_idx(type[j], c.num_types);
          rB = c.mat[tB].radius;
        }
        const float rx = pB.x - p.x, ry = pB.y - p.y;
        const float dist = sqrtf(rx * rx + ry * ry);
        const float R = rA + rB;
        if (kRadius && (dist < R || c.attract)) tB = clamp_idx(type[j], c.num_types);
        if (dist < R) {
          const float kBn = c.mat[tB].kn, kBt = c.mat[tB].kt, cBn = c.mat[tB].cn, muB = c.mat[tB].mu;
          const float2 vB = vel[j];
          const float wB = omega[j];
          const float nx = rx / dist, ny = ry / dist;
          const float tx = -ny, ty = nx;
          const float rvx = vB.x - v.x, rvy = vB.y - v.y;
          const float vn = rvx * nx + rvy * ny;
          const float vtx = rvx - vn * nx, vty = rvy - vn * ny;
          const float kn = series(mA.kn, kBn);
          const float kt = series(mA.kt, kBt);
          const float cn = series(mA.cn, cBn);
          const float mu = (mA.mu + muB) / 2;
          const float s = -kn * (R - dist);
          float ffx = s * nx, ffy = s * ny;
          if (vn < 0.0f) {
            const float s2 = cn * vn;
            ffx += s2 * nx;
            ffy += s2 * ny;
          }
          const float flen = sqrtf(ffx * ffx + ffy * ffy);
          pres = add_pressure(pres, flen, den);
          if (mu != 0.0f) {
            const float sr = rA * w + rB * wB;
            const float cvx = vtx - sr * tx, cvy = vty - sr * ty;
            float ftx = (kt * cvx) * dt, fty = (kt * cvy) * dt;
            const float ftmax = mu * flen;
            if (!(sqrtf(ftx * ftx + fty * fty) < ftmax)) {
              const float cl = sqrtf(cvx * cvx + cvy * cvy);
              ftx = (ftmax * cvx) / cl;
              fty = (ftmax * cvy) / cl;
            }
            ffx += ftx;
            ffy += fty;
            moment += rA * (ftx * tx + fty * ty);
          }
          cfx += ffx;
          cfy += ffy;
        }
        if (c.attract) {
          const float aA = mA.attract, aB = c.mat[tB].attract;
          const float adA = (float)((double)(rA * 2.0f) * 1.2);
          const float adB = (float)((double)(rB * 2.0f) * 1.2);
          const float nx = rx / dist, ny = ry / dist;
          float afx = 0.0f, afy = 0.0f;
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
          pres = add_pressure(pres, sqrtf(afx * afx + afy * afy), den);
          cfx += afx;
          cfy += afy;
        }
      }
      fx += cfx;
      fy += cfy;
    }
  }
  acc[i] = make_float2(fx / massA, fy / massA);
  alpha[i] = moment / inertiaA;
  pressure[i] = pres;
}

// ------------------------------------------------------------- integrate --
__device__ __forceinline__ float wrap_theta(float th) {
  const double twopi = 2 * M_PI;
  if ((double)th > twopi) th = (float)((double)th - twopi * floor((double)th / twopi));
  if (th < 0.0f) th = (float)((double)th - twopi * floor((double)th / twopi));
  return th;
}

// integrateSystemD (src/particles_kernel.cuh:580-627) with collideBoundary
// (:388-456) inlined for the bottom wall.
template <bool kRadius, bool kMulti>
__global__ __launch_bounds__(256) void k_integrate(
    StepConst c, float2* __restrict__ pos, float2* __restrict__ vel, const float2* __restrict__ acc,
    float* __restrict__ theta, float* __restrict__ omega, const float* __restrict__ alpha,
    const uint32_t* __restrict__ type, const float* __restrict__ radius,
    float* __restrict__ pressure, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = clamp_idx(type[i], c.num_types);
  const Material m = c.mat[t];
  const Instance in = kMulti ? c.inst[clamp_idx(m.inst, c.num_instances)] : c.inst0;
  float r, mass, inertia;
  body<kRadius>(m, radius, i, r, mass, inertia);
  const float dt = c.dt;

  const float2 a = acc[i];
  float2 v = vel[i];
  v.x = v.x + (c.gx + a.x) * dt;
  v.y = v.y + (c.gy + a.y) * dt;
  float2 p = pos[i];
  p.x = p.x + v.x * dt;
  p.y = p.y + v.y * dt;
  float w = omega[i] + alpha[i] * dt;
  float th = theta[i] + w * dt;
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
    float pr = pressure[i];
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
        if (!(sqrtf(ftx * ftx + fty * fty) < ftmax)) {
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
    pressure[i] = pr;
  }
  pos[i] = p;
  vel[i] = v;
  omega[i] = w;
  theta[i] = th;
}

inline dim3 grid_for(uint32_t n, int block) { return dim3((n + (uint32_t)block - 1u) / (uint32_t)block); }

}  // namespace

hipError_t launch_hash(const StepConst& c, hipStream_t s, int block, const float2* pos,
                       const uint32_t* type, uint32_t* hash, uint32_t* index, uint32_t n) {
  if (n == 0) return hipSuccess;
  if (type && c.num_instances > 1)
    hipLaunchKernelGGL(k_hash<true>, grid_for(n, block), dim3(block), 0, s, c, pos, type, hash, index, n);
  else
    hipLaunchKernelGGL(k_hash<false>, grid_for(n, block), dim3(block), 0, s, c, pos, type, hash, index, n);
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, uint32_t* p, uint32_t value, uint32_t count) {
  if (count == 0) return hipSuccess;
  uint32_t head = 0;
  if ((reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
    const uint32_t n4 = count / 4u;
    if (n4) hipLaunchKernelGGL(k_fill4, grid_for(n4, 256), dim3(256), 0, s, reinterpret_cast<uint4*>(p), value, n4);
    head = n4 * 4u;
  }
  if (count > head)
    hipLaunchKernelGGL(k_fill1, grid_for(count - head, 256), dim3(256), 0, s, p + head, value, count - head);
  return hipGetLastError();
}

hipError_t launch_reorder(hipStream_t s, int block, uint32_t* cell_start, uint32_t* cell_end,
                          float2* s_pos, float2* s_vel, float* s_theta, float* s_omega,
                          uint32_t* s_id, uint32_t* s_loc, uint32_t* s_type, float* s_radius,
                          const uint32_t* hash, const uint32_t* index, const float2* o_pos,
                          const float2* o_vel, const float* o_theta, const float* o_omega,
                          const uint32_t* o_id, const uint32_t* o_type, const float* o_radius,
                          uint32_t n, uint32_t num_cells) {
  if (n == 0) return hipSuccess;
  if (s_radius)
    hipLaunchKernelGGL(k_reorder<true>, grid_for(n, block), dim3(block), 0, s, cell_start, cell_end, s_pos,
                       s_vel, s_theta, s_omega, s_id, s_loc, s_type, s_radius, hash, index, o_pos, o_vel,
                       o_theta, o_omega, o_id, o_type, o_radius, n, num_cells);
  else
    hipLaunchKernelGGL(k_reorder<false>, grid_for(n, block), dim3(block), 0, s, cell_start, cell_end, s_pos,
                       s_vel, s_theta, s_omega, s_id, s_loc, s_type, s_radius, hash, index, o_pos, o_vel,
                       o_theta, o_omega, o_id, o_type, o_radius, n, num_cells);
  return hipGetLastError();
}

hipError_t launch_collide(const StepConst& c, hipStream_t s, int block, const float2* pos,
                          const float2* vel, float2* acc, const float* omega, float* alpha,
                          const uint32_t* type, const float* radius, const uint32_t* cell_start,
                          const uint32_t* cell_end, float* pressure, uint32_t n) {
  if (n == 0) return hipSuccess;
  const bool multi = c.num_instances > 1;
#define DEM_COLLIDE(R, M)                                                                        \
  hipLaunchKernelGGL((k_collide<R, M>), grid_for(n, block), dim3(block), 0, s, c, pos, vel, acc, \
                     omega, alpha, type, radius, cell_start, cell_end, pressure, n)
  if (radius) {
    if (multi) DEM_COLLIDE(true, true); else DEM_COLLIDE(true, false);
  } else {
    if (multi) DEM_COLLIDE(false, true); else DEM_COLLIDE(false, false);
  }
#undef DEM_COLLIDE
  return hipGetLastError();
}

hipError_t launch_integrate(const StepConst& c, hipStream_t s, int block, float2* pos, float2* vel,
                            const float2* acc, float* theta, float* omega, const float* alpha,
                            const uint32_t* type, const float* radius, float* pressure, uint32_t n) {
  if (n == 0) return hipSuccess;
  const bool multi = c.num_instances > 1;
#define DEM_INTEGRATE(R, M)                                                                         \
  hipLaunchKernelGGL((k_integrate<R, M>), grid_for(n, block), dim3(block), 0, s, c, pos, vel, acc, \
                     theta, omega, alpha, type, radius, pressure, n)
  if (radius) {
    if (multi) DEM_INTEGRATE(true, true); else DEM_INTEGRATE(true, false);
  } else {
    if (multi) DEM_INTEGRATE(false, true); else DEM_INTEGRATE(false, false);
  }
#undef DEM_INTEGRATE
  return hipGetLastError();
}

}  // namespace dem
