// adjoint.hip -- reverse mode of the DEM step (collide + integrate) on gfx950.
//
// The forward keeps a checkpoint of the state every K steps.  The backward
// walks the checkpoints in reverse; for each segment it re-runs the K
// forward steps, storing the sorted pre-collide state and the permutation of
// each step (the neighbour lists are not stored: cells are rebuilt from the
// stored sorted positions), then per step in reverse:
//   cells   -> forces (the forward collide kernel, bit-identical acc/alpha)
//   k_adj_integrate : reverse of integrateSystemD (src/particles_kernel.cuh:
//                     580-627) incl. the walls and the bottom-wall law
//                     (:388-456), replaying the FP32 forward for its
//                     switch decisions
//   k_adj_collide   : reverse of the pair law (:235-307) over the 3x3
//                     neighbourhood, one lane per particle gathering the
//                     neighbours' force/moment adjoints (no atomics on state)
//   k_adj_permute   : adjoint of the sort permutation (scatter; after an
//                     in-place re-sort of more than kScatterMax particles, in
//                     place at the moved slots instead)
// Switches (contact, approach-only damping, stick/slip, wall hits) are
// frozen at the forward's decisions.  In exact arithmetic the pair law is
// F = s n + f_t t, tau = f_t with s = -kn (R - dist) + [vn<0] cn vn,
// c_t = (vB - vA).t - (rA wA + rB wB), f_t = kt dt c_t (stick) or
// mu |s| sgn(c_t) (slip); force on B = -F, moments rA f_t and rB f_t.  The
// adjoint state and the vector-Jacobian products are FP64; parameter
// gradients (kn, kt, cn, mu per material, through the series combination)
// are accumulated in FP64.  Attraction is not differentiated (rejected).
#include "capi_internal.h"

#include <math.h>

#include <algorithm>
#include <new>
#include <vector>

#pragma clang fp contract(off)

namespace dem {
namespace {

constexpr int kParams = 4;  // kn, kt, cn, mu
// Up to this many particles the adjoint of an in-place step's permutation is
// one full scatter (a launch costs more than copying the whole adjoint
// state); beyond it, only the listed slots are saved and put back.
constexpr uint32_t kScatterMax = 262144;

__device__ __forceinline__ uint32_t clampi(uint32_t t, uint32_t count) { return t < count ? t : count - 1u; }

__device__ __forceinline__ float series_f(float a, float b) {
  if (a == 0.0f || b == 0.0f) return 0.0f;
  return (a * b) / (a + b);
}

__device__ __forceinline__ void series_grad(double a, double b, double& da, double& db) {
  if (a == 0.0 || b == 0.0) {
    da = db = 0.0;
    return;
  }
  const double inv = 1.0 / ((a + b) * (a + b));
  da = b * b * inv;
  db = a * a * inv;
}

__device__ __forceinline__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// Same conversion as the forward kernels: NaN -> 0, saturation, truncation.
__device__ __forceinline__ int to_int_sat(float g) {
  int r = (int)fminf(fmaxf(g, -2147483648.0f), 2147483520.0f);
  r = (g >= 2147483648.0f) ? 2147483647 : r;
  return (g == g) ? r : 0;
}

template <bool kRadius>
__device__ __forceinline__ void body_of(const Material& m, const float* __restrict__ radius, uint32_t i, float& r,
                                        float& mass, float& inertia) {
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

// Adds a lane's gradient contributions for material t: one atomic per
// parameter per wave when the whole wave shares the material, else per lane.
__device__ __forceinline__ void flush_grads(double g[kParams], uint32_t t, bool active, double* __restrict__ grad) {
  const uint32_t t0 = __shfl(t, 0);
  const bool uniform = __all(!active || t == t0);
  if (uniform) {
#pragma unroll
    for (int p = 0; p < kParams; ++p) {
      double v = active ? g[p] : 0.0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      g[p] = v;
    }
    if (__lane_id() == 0) {
#pragma unroll
      for (int p = 0; p < kParams; ++p)
        if (g[p] != 0.0) atomicAdd(&grad[kParams * t0 + p], g[p]);
    }
  } else if (active) {
#pragma unroll
    for (int p = 0; p < kParams; ++p)
      if (g[p] != 0.0) atomicAdd(&grad[kParams * t + p], g[p]);
  }
}

struct AdjIn {
  const float2* pos;
  const float2* vel;
  const float* omega;
  const float* radius;
  const uint32_t* type;
};

// ---------------------------------------------------------------------------
__global__ void k_adj_init(const uint32_t* __restrict__ id, const float* __restrict__ gpos,
                           const float* __restrict__ gvel, double2* __restrict__ lx, double2* __restrict__ lv,
                           double* __restrict__ lw, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = min(id[i], n - 1u);
  lx[i] = gpos ? make_double2(gpos[2 * d], gpos[2 * d + 1]) : make_double2(0.0, 0.0);
  lv[i] = gvel ? make_double2(gvel[2 * d], gvel[2 * d + 1]) : make_double2(0.0, 0.0);
  lw[i] = 0.0;
}

__global__ void k_adj_out(const uint32_t* __restrict__ id, const double2* __restrict__ lx,
                          const double2* __restrict__ lv, float* __restrict__ gpos0, float* __restrict__ gvel0,
                          uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = id[i];
  if (d >= n) return;
  if (gpos0) {
    gpos0[2 * d] = (float)lx[i].x;
    gpos0[2 * d + 1] = (float)lx[i].y;
  }
  if (gvel0) {
    gvel0[2 * d] = (float)lv[i].x;
    gvel0[2 * d + 1] = (float)lv[i].y;
  }
}

// sorted-state snapshot of one step: pos, vel, omega, radius, type, index
// One step's slot: the sorted pre-integrate state, the step's sorted keys
// and permutation, and collide's accelerations.
// (quiet: the in-place step's runs of 64 slots whose acc and alpha collide
// left unwritten as +0, or null)
__global__ void k_store_slot(const float2* __restrict__ pos, const float2* __restrict__ vel,
                             const float* __restrict__ omega, const float* __restrict__ radius,
                             const uint32_t* __restrict__ type, const uint32_t* __restrict__ index,
                             const uint32_t* __restrict__ key, const float2* __restrict__ acc,
                             const float* __restrict__ alpha, const uint32_t* __restrict__ quiet,
                             float2* __restrict__ spos, float2* __restrict__ svel,
                             float* __restrict__ somega, float* __restrict__ sradius, uint32_t* __restrict__ stype,
                             uint32_t* __restrict__ sindex, uint32_t* __restrict__ skey, float2* __restrict__ sacc,
                             float* __restrict__ salpha, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  spos[i] = pos[i];
  svel[i] = vel[i];
  somega[i] = omega[i];
  if (radius) sradius[i] = radius[i];
  stype[i] = type[i];
  sindex[i] = index ? index[i] : i;  // without an index: the identity (an in-place step's, before its listed slots)
  skey[i] = key[i];
  const bool still = quiet && quiet[i / 64u] != 0u;
  sacc[i] = still ? make_float2(0.0f, 0.0f) : acc[i];
  salpha[i] = still ? 0.0f : alpha[i];
}

// cellStart/cellEnd of sorted keys (cellStart pre-filled with EMPTY)
__global__ void k_cell_bounds(const uint32_t* __restrict__ key, uint32_t* __restrict__ cs, uint32_t* __restrict__ ce,
                              uint32_t n, uint32_t num_cells) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t h = key[i];
  if (i == 0 || h != key[i - 1]) {
    if (h < num_cells) cs[h] = i;
    if (i > 0 && key[i - 1] < num_cells) ce[key[i - 1]] = i;
  }
  if (i == n - 1 && h < num_cells) ce[h] = n;
}

__global__ void k_adj_permute(const uint32_t* __restrict__ index, const double2* __restrict__ lxp,
                              const double2* __restrict__ lvp, const double* __restrict__ lwp,
                              double2* __restrict__ lx, double2* __restrict__ lv, double* __restrict__ lw,
                              uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t src = index[i];
  if (src >= n) return;
  lx[src] = lxp[i];
  lv[src] = lvp[i];
  lw[src] = lwp[i];
}

// The same for a step whose index is the identity except at a few slots (an
// in-place re-sort), in place on the pre-step adjoint: the listed slots form
// a permutation among themselves, so their values are first copied aside
// (into the post-step arrays, no longer needed), then written to their source
// slots; the caller then swaps the roles of the two arrays.
__global__ void k_adj_perm_save(const uint32_t* __restrict__ index, const double2* __restrict__ lxp,
                                const double2* __restrict__ lvp, const double* __restrict__ lwp,
                                double2* __restrict__ tx, double2* __restrict__ tv, double* __restrict__ tw,
                                uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || index[i] == i) return;
  tx[i] = lxp[i];
  tv[i] = lvp[i];
  tw[i] = lwp[i];
}

__global__ void k_adj_perm_put(const uint32_t* __restrict__ index, const double2* __restrict__ tx,
                               const double2* __restrict__ tv, const double* __restrict__ tw,
                               double2* __restrict__ lxp, double2* __restrict__ lvp, double* __restrict__ lwp,
                               uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t src = index[i];
  if (src == i || src >= n) return;
  lxp[src] = tx[i];
  lvp[src] = tv[i];
  lwp[src] = tw[i];
}

// ---------------------------------------------------------------------------
// Reverse of integrateSystemD for one particle.
template <bool kRadius, bool kMulti>
__global__ __launch_bounds__(256) void k_adj_integrate(StepConst c, AdjIn S, const float2* __restrict__ acc,
                                                       const float* __restrict__ alpha,
                                                       const double2* __restrict__ lx, const double2* __restrict__ lv,
                                                       const double* __restrict__ lw, double2* __restrict__ lxp,
                                                       double2* __restrict__ lvp, double* __restrict__ lwp,
                                                       double2* __restrict__ lF, double* __restrict__ lM,
                                                       double* __restrict__ grad, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < n;
  const uint32_t ii = active ? i : 0u;
  const uint32_t t = clampi(S.type[ii], c.num_types);
  double g[kParams] = {0.0, 0.0, 0.0, 0.0};
  if (active) {
    const Material m = c.mat[t];
    const Instance in = kMulti ? c.inst[clampi(m.inst, c.num_instances)] : c.inst0;
    float r, mass, inertia;
    body_of<kRadius>(m, S.radius, i, r, mass, inertia);
    const float dt = c.dt;
    // FP32 forward replay (same operations as k_integrate) for the switches
    const float2 a = acc[i];
    float2 v1 = S.vel[i];
    v1.x = v1.x + (c.gx + a.x) * dt;
    v1.y = v1.y + (c.gy + a.y) * dt;
    float2 p1 = S.pos[i];
    p1.x = p1.x + v1.x * dt;
    p1.y = p1.y + v1.y * dt;
    float w1 = S.omega[i] + alpha[i] * dt;
    const float bd = -m.bdamp;
    bool fr = false, fl = false, ft = false, fb = false;
    if (p1.x > in.box[0] - r) { p1.x = in.box[0] - r; v1.x = v1.x * bd; fr = true; }
    if (p1.x < r) { p1.x = r; v1.x = v1.x * bd; fl = true; }
    if (!(c.flags & DEM_OPEN_TOP) && p1.y > in.box[1] - r) { p1.y = in.box[1] - r; v1.y = v1.y * bd; ft = true; }
    bool contact = false, damp = false, fric = false, stick = false;
    float sy = 0.0f, cx = 0.0f, P = 0.0f, ktw = 0.0f, f0x = 0.0f, f0y = 0.0f;
    if (p1.y < r) {
      fb = true;
      const float rx = p1.x - p1.x, ry = 0.0f - p1.y;
      const float dist = sqrtf(rx * rx + ry * ry);
      if (dist < r) {
        contact = true;
        const float nx = rx / dist, ny = ry / dist;
        const float tx = -ny, ty = nx;
        sy = tx;  // t = (sy, 0)
        const float rvx = -v1.x, rvy = -v1.y;
        const float vn = rvx * nx + rvy * ny;
        const float vtx = rvx - vn * nx, vty = rvy - vn * ny;
        ktw = series_f(m.kt, c.bkt);
        const float s = -mass;
        f0x = s * c.gx;
        f0y = s * c.gy;
        if (vn < 0.0f) {
          damp = true;
          const float s2 = m.cn * vn;
          f0x += s2 * nx;
          f0y += s2 * ny;
        }
        P = sqrtf(f0x * f0x + f0y * f0y);
        if (m.mu != 0.0f) {
          fric = true;
          const float sr = r * w1;
          const float cvx = vtx - sr * tx, cvy = vty - sr * ty;
          const float ftx = (ktw * cvx) * dt, fty = (ktw * cvy) * dt;
          stick = sqrtf(ftx * ftx + fty * fty) < m.mu * P;
          cx = cvx;
        }
      }
    }
    // reverse, in double
    const double ddt = dt, dbd = bd, dm = mass, dI = inertia, dr = r;
    double2 lp1 = lx[i], lv1 = lv[i];
    double lw1 = lw[i];
    if (fb) {
      lv1.y *= dbd;
      lp1.y = 0.0;
      const double lFwx = (lv1.x * ddt + lp1.x * (ddt * ddt)) / dm;
      const double lFwy = (lv1.y * ddt + lp1.y * (ddt * ddt)) / dm;
      const double law = lw1 * ddt;
      if (contact) {
        if (fric) {
          const double lFt = lFwx + law * dr * (double)sy / dI;
          if (stick) {
            const double lc = lFt * (double)ktw * ddt;
            double da, db;
            series_grad(m.kt, c.bkt, da, db);
            g[1] += lFt * ddt * (double)cx * da;
            lv1.x += -lc;
            lw1 += -(double)sy * dr * lc;
          } else {
            const double lP = lFt * (double)m.mu * sgn(cx);
            g[3] += lFt * (double)P * sgn(cx);
            if (damp && P > 0.0f) {
              const double lf0y = lP * (double)f0y / (double)P;
              lv1.y += lf0y * (-(double)m.cn);
              g[2] += lf0y * (-(double)v1.y);
            }
          }
        } else if (damp) {
          lv1.y += lFwy * (-(double)m.cn);
          g[2] += lFwy * (-(double)v1.y);
        }
      }
    }
    if (ft) { lv1.y *= dbd; lp1.y = 0.0; }
    if (fl) { lv1.x *= dbd; lp1.x = 0.0; }
    if (fr) { lv1.x *= dbd; lp1.x = 0.0; }
    lv1.x += lp1.x * ddt;
    lv1.y += lp1.y * ddt;
    lxp[i] = lp1;
    lvp[i] = lv1;
    lwp[i] = lw1;
    lF[i] = make_double2(lv1.x * (ddt / dm), lv1.y * (ddt / dm));
    lM[i] = lw1 * ddt / dI;
  }
  flush_grads(g, t, active, grad);
}

// ---------------------------------------------------------------------------
// Reverse of the pair law for one contact of particle A with candidate j:
// accumulates the adjoints of A's position, velocity and spin and A's share
// of the parameter gradients.  The contact, approach and stick decisions are
// the forward's, replayed in FP32 with its operations (the exact distance
// prefilter of the forward collide skips the candidates that cannot touch);
// the derivatives are in double.
struct AdjA {
  float2 pA, vA;
  float wA, rA;
  uint32_t tA;
  Material A;
  double2 lFA;
  double lMA;
  double2 dx_acc, drv_acc;
  double dw_acc;
};

template <bool kRadius>
__device__ __forceinline__ void adj_candidate(const StepConst& c, const AdjIn& S, const double2* __restrict__ lF,
                                              const double* __restrict__ lM, double* __restrict__ grad, uint32_t i,
                                              uint32_t j, AdjA& q, double g[kParams]) {
  if (j == i) return;
  uint32_t tB = 0;
  float rB;
  if (kRadius) {
    rB = S.radius[j];
  } else {
    tB = clampi(S.type[j], c.num_types);
    rB = c.mat[tB].radius;
  }
  const float2 pB = S.pos[j];
  // FP32 forward replay of collideSpheres' decisions
  const float rx = pB.x - q.pA.x, ry = pB.y - q.pA.y;
  const float d2 = rx * rx + ry * ry;
  const float R = q.rA + rB;
  // d2 > fl(R*R)*(1+2^-20) implies sqrtf(d2) >= R (the forward's prefilter)
  if (d2 > (R * R) * 1.00000095367431640625f && !(R < 1.0842022e-19f)) return;
  const float dist = sqrtf(d2);
  if (!(dist < R)) return;
  if (kRadius) tB = clampi(S.type[j], c.num_types);
  const float2 vB = S.vel[j];
  const float wB = S.omega[j];
  const float nx = rx / dist, ny = ry / dist;
  const float tx = -ny, ty = nx;
  const float rvx = vB.x - q.vA.x, rvy = vB.y - q.vA.y;
  const float vn = rvx * nx + rvy * ny;
  const float vtx = rvx - vn * nx, vty = rvy - vn * ny;
  const float4 pr = c.pair[q.tA * c.num_types + tB];  // kn, kt, cn in series and the mean mu
  const float kn = pr.x, kt = pr.y, cn = pr.z, mu = pr.w;
  const float dt = c.dt;
  const bool damp = vn < 0.0f;
  bool stick = false;
  const bool fric = mu != 0.0f;
  if (fric) {
    const float s = -kn * (R - dist);
    float ffx = s * nx, ffy = s * ny;
    if (damp) {
      const float s2 = cn * vn;
      ffx += s2 * nx;
      ffy += s2 * ny;
    }
    const float flen = sqrtf(ffx * ffx + ffy * ffy);
    const float sr = q.rA * q.wA + rB * wB;
    const float cvx = vtx - sr * tx, cvy = vty - sr * ty;
    const float ftx = (kt * cvx) * dt, fty = (kt * cvy) * dt;
    stick = sqrtf(ftx * ftx + fty * fty) < mu * flen;
  }
  // closed-form derivatives in double
  const double dn_x = nx, dn_y = ny, dt_x = tx, dt_y = ty, ddist = dist;
  const double inv_d = 1.0 / ddist;
  const double dvn = (double)rvx * dn_x + (double)rvy * dn_y;
  const double dvt = (double)rvx * dt_x + (double)rvy * dt_y;
  const double dkn = kn, dkt = kt, dcn = cn, dmu = mu, ddtm = dt;
  const double delta = (double)R - ddist;
  const double s = -dkn * delta + (damp ? dcn * dvn : 0.0);
  const double P = fabs(s);
  double ct = 0.0, ft = 0.0;
  if (fric) {
    ct = dvt - ((double)q.rA * q.wA + (double)rB * wB);
    ft = stick ? dkt * ddtm * ct : dmu * P * sgn(ct);
  }
  const double2 lFB = lF[j];
  const double gFx = q.lFA.x - lFB.x, gFy = q.lFA.y - lFB.y;
  const double gtau = (double)q.rA * q.lMA + (double)rB * lM[j];
  const double a = gFx * dn_x + gFy * dn_y;
  const double gFt = gFx * dt_x + gFy * dt_y;
  const double bb = gFt + gtau;
  const double cdamp = damp ? dcn : 0.0;
  const double dsx = dn_x * dkn + dt_x * (cdamp * dvt * inv_d);
  const double dsy = dn_y * dkn + dt_y * (cdamp * dvt * inv_d);
  const double dsrx = dn_x * cdamp, dsry = dn_y * cdamp;
  double dftx = 0.0, dfty = 0.0, dftrx = 0.0, dftry = 0.0, dftw = 0.0, slip_k = 0.0;
  if (fric) {
    if (stick) {
      dftx = dt_x * (-dkt * ddtm * dvn * inv_d);
      dfty = dt_y * (-dkt * ddtm * dvn * inv_d);
      dftrx = dt_x * (dkt * ddtm);
      dftry = dt_y * (dkt * ddtm);
      dftw = -dkt * ddtm;
    } else {
      slip_k = dmu * sgn(s) * sgn(ct);
      dftx = dsx * slip_k;
      dfty = dsy * slip_k;
      dftrx = dsrx * slip_k;
      dftry = dsry * slip_k;
    }
  }
  q.dx_acc.x += dsx * a + dt_x * (s * gFt * inv_d) + dftx * bb - dt_x * (ft * a * inv_d);
  q.dx_acc.y += dsy * a + dt_y * (s * gFt * inv_d) + dfty * bb - dt_y * (ft * a * inv_d);
  q.drv_acc.x += dsrx * a + dftrx * bb;
  q.drv_acc.y += dsry * a + dftry * bb;
  q.dw_acc += dftw * bb;
  if (i < j) {
    const double gkn = a * (-delta) + bb * slip_k * (-delta);
    const double gcn = damp ? (a * dvn + bb * slip_k * dvn) : 0.0;
    const double gkt = (fric && stick) ? bb * ddtm * ct : 0.0;
    const double gmu = (fric && !stick) ? bb * P * sgn(ct) : 0.0;
    double ka, kb, ta, tb, ca, cb;
    if (c.dpair) {
      const double* d = c.dpair + 6 * ((size_t)q.tA * c.num_types + tB);
      ka = d[0];
      kb = d[1];
      ta = d[2];
      tb = d[3];
      ca = d[4];
      cb = d[5];
    } else {
      const Material& B = c.mat[tB];
      series_grad(q.A.kn, B.kn, ka, kb);
      series_grad(q.A.kt, B.kt, ta, tb);
      series_grad(q.A.cn, B.cn, ca, cb);
    }
    g[0] += gkn * ka;
    g[1] += gkt * ta;
    g[2] += gcn * ca;
    g[3] += gmu * 0.5;
    if (tB == q.tA) {
      g[0] += gkn * kb;
      g[1] += gkt * tb;
      g[2] += gcn * cb;
      g[3] += gmu * 0.5;
    } else {
      if (gkn * kb != 0.0) atomicAdd(&grad[kParams * tB + 0], gkn * kb);
      if (gkt * tb != 0.0) atomicAdd(&grad[kParams * tB + 1], gkt * tb);
      if (gcn * cb != 0.0) atomicAdd(&grad[kParams * tB + 2], gcn * cb);
      if (gmu != 0.0) atomicAdd(&grad[kParams * tB + 3], gmu * 0.5);
    }
  }
}

// Reverse of the pair law for particle i over its 3x3 neighbourhood, in the
// forward's candidate order: cells by row (y = -1, 0, 1) and column; the
// three cells of a row are consecutive keys, walked as one range of sorted
// candidates (cell by cell where a wrapping grid breaks that at an edge).
template <bool kRadius, bool kMulti>
__global__ __launch_bounds__(256) void k_adj_collide(StepConst c, AdjIn S, const uint32_t* __restrict__ cell_start,
                                                     const uint32_t* __restrict__ cell_end,
                                                     const double2* __restrict__ lF, const double* __restrict__ lM,
                                                     double2* __restrict__ lxp, double2* __restrict__ lvp,
                                                     double* __restrict__ lwp, double* __restrict__ grad, uint32_t n) {
  const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  const bool active = i < n;
  const uint32_t ii = active ? i : 0u;
  const uint32_t tA = clampi(S.type[ii], c.num_types);
  double g[kParams] = {0.0, 0.0, 0.0, 0.0};
  if (active) {
    AdjA q;
    q.tA = tA;
    q.A = c.mat[tA];
    const Instance in = kMulti ? c.inst[clampi(q.A.inst, c.num_instances)] : c.inst0;
    float massA, inertiaA;
    body_of<kRadius>(q.A, S.radius, i, q.rA, massA, inertiaA);
    q.pA = S.pos[i];
    q.vA = S.vel[i];
    q.wA = S.omega[i];
    const bool clamp = c.flags & DEM_GRID_CLAMP;
    int gx = to_int_sat(floorf(q.pA.x * (float)in.grid[0] / in.extent[0]));
    int gy = to_int_sat(floorf(q.pA.y * (float)in.grid[1] / in.extent[1]));
    if (clamp) {
      gx = min(max(gx, 0), (int)in.grid[0] - 1);
      gy = min(max(gy, 0), (int)in.grid[1] - 1);
    }
    q.lFA = lF[i];
    q.lMA = lM[i];
    q.dx_acc = make_double2(0.0, 0.0);
    q.drv_acc = make_double2(0.0, 0.0);
    q.dw_acc = 0.0;
    const uint32_t G0 = in.grid[0], G1 = in.grid[1];
    const bool vx0 = !clamp || gx > 0, vx2 = !clamp || gx < (int)G0 - 1;
    const uint32_t x0 = ((uint32_t)gx - 1u) & (clamp ? 0xFFFFFFFFu : G0 - 1u);
    const uint32_t x1 = clamp ? (uint32_t)gx : ((uint32_t)gx & (G0 - 1u));
    const uint32_t x2 = ((uint32_t)gx + 1u) & (clamp ? 0xFFFFFFFFu : G0 - 1u);
    const bool contiguous = clamp || (x0 + 1u == x1 && x1 + 1u == x2);
    // Contiguous rows (the usual case): all three rows' candidate ranges are
    // loaded at once; rows of at most 64 candidates are walked together
    // marking the candidates the forward's prefilter lets through (three
    // loads in flight), then the pair adjoint runs over the marks in row
    // order.  Candidates out of contact contribute nothing, so the sums are
    // the same, in the same order, as walking every candidate.
    bool walked = false;
    if (contiguous) {
      uint32_t rs[3], re[3];
#pragma unroll
      for (int y = 0; y < 3; ++y) {
        const int iy = gy + y - 1;
        const bool ok = !clamp || (iy >= 0 && iy < (int)G1);
        const uint32_t cy = clamp ? (ok ? (uint32_t)iy : 0u) : ((uint32_t)iy & (G1 - 1u));
        const uint32_t row = in.cell_base + cy * G0;
        const uint32_t a0 = row + (vx0 ? x0 : x1), a1 = row + x1, a2 = row + (vx2 ? x2 : x1);
        const uint32_t s0r = cell_start[a0], s1r = cell_start[a1], s2r = cell_start[a2];
        const uint32_t e0r = cell_end[a0], e1r = cell_end[a1], e2r = cell_end[a2];
        const uint32_t s0 = ok && vx0 ? s0r : kEmpty, s1 = ok ? s1r : kEmpty, s2 = ok && vx2 ? s2r : kEmpty;
        const bool n0 = s0 != kEmpty, n1 = s1 != kEmpty, n2 = s2 != kEmpty;
        rs[y] = n0 ? s0 : (n1 ? s1 : (n2 ? s2 : 0u));
        re[y] = (n0 || n1 || n2) ? min(n2 ? e2r : (n1 ? e1r : e0r), n) : 0u;
        if (re[y] < rs[y]) re[y] = rs[y];
      }
      const uint32_t len0 = re[0] - rs[0], len1 = re[1] - rs[1], len2 = re[2] - rs[2];
      if (len0 <= 64u && len1 <= 64u && len2 <= 64u) {
        walked = true;
        uint64_t h[3] = {0u, 0u, 0u};
        const uint32_t L = max(len0, max(len1, len2));
        for (uint32_t k = 0; k < L; ++k) {
          const uint32_t jy[3] = {k < len0 ? rs[0] + k : i, k < len1 ? rs[1] + k : i, k < len2 ? rs[2] + k : i};
          float2 pb[3];
          float rb[3];
#pragma unroll
          for (int y = 0; y < 3; ++y) {
            pb[y] = S.pos[jy[y]];
            rb[y] = kRadius ? S.radius[jy[y]] : c.mat[clampi(S.type[jy[y]], c.num_types)].radius;
          }
#pragma unroll
          for (int y = 0; y < 3; ++y) {
            const float rx = pb[y].x - q.pA.x, ry = pb[y].y - q.pA.y;
            const float d2 = rx * rx + ry * ry;
            const float R = q.rA + rb[y];
            const bool maybe = !(d2 > (R * R) * 1.00000095367431640625f) || R < 1.0842022e-19f;
            h[y] |= (uint64_t)(maybe && jy[y] != i) << k;
          }
        }
#pragma unroll
        for (int y = 0; y < 3; ++y)
          for (uint64_t hy = h[y]; hy; hy &= hy - 1u)
            adj_candidate<kRadius>(c, S, lF, lM, grad, i, rs[y] + (uint32_t)__builtin_ctzll(hy), q, g);
      }
    }
    for (int y = -1; y <= 1 && !walked; ++y) {
      const int iy = gy + y;
      uint32_t cy;
      if (clamp) {
        if (iy < 0 || iy >= (int)G1) continue;
        cy = (uint32_t)iy;
      } else {
        cy = (uint32_t)iy & (G1 - 1u);
      }
      const uint32_t row = in.cell_base + cy * G0;
      if (contiguous) {
        // the row's candidates: from the first non-empty cell's start to the
        // last non-empty cell's end
        const uint32_t a0 = row + (vx0 ? x0 : x1), a1 = row + x1, a2 = row + (vx2 ? x2 : x1);
        const uint32_t s0 = vx0 ? cell_start[a0] : kEmpty, s1 = cell_start[a1], s2 = vx2 ? cell_start[a2] : kEmpty;
        const bool n0 = s0 != kEmpty, n1 = s1 != kEmpty, n2 = s2 != kEmpty;
        if (!(n0 || n1 || n2)) continue;
        const uint32_t rs = n0 ? s0 : (n1 ? s1 : s2);
        const uint32_t re = min(n2 ? cell_end[a2] : (n1 ? cell_end[a1] : cell_end[a0]), n);
        for (uint32_t j = rs; j < re; ++j) adj_candidate<kRadius>(c, S, lF, lM, grad, i, j, q, g);
      } else {
        const uint32_t xs[3] = {x0, x1, x2};
        for (int k = 0; k < 3; ++k) {
          const uint32_t key = row + xs[k];
          const uint32_t st = cell_start[key];
          if (st == kEmpty) continue;
          const uint32_t en = min(cell_end[key], n);
          for (uint32_t j = st; j < en; ++j) adj_candidate<kRadius>(c, S, lF, lM, grad, i, j, q, g);
        }
      }
    }
    double2 lxi = lxp[i], lvi = lvp[i];
    lxi.x -= q.dx_acc.x;
    lxi.y -= q.dx_acc.y;
    lvi.x -= q.drv_acc.x;
    lvi.y -= q.drv_acc.y;
    lxp[i] = lxi;
    lvp[i] = lvi;
    lwp[i] = lwp[i] + (double)q.rA * q.dw_acc;
  }
  flush_grads(g, tA, active, grad);
}

inline dim3 grid_of(uint32_t n) { return dim3((n + 255u) / 256u); }

}  // namespace
}  // namespace dem

using namespace dem;

// One device copy of the current state buffer of a sim.
struct AdjCheckpoint {
  float2 *pos = nullptr, *vel = nullptr;
  float *theta = nullptr, *omega = nullptr, *radius = nullptr;
  uint32_t *id = nullptr, *type = nullptr, *loc = nullptr;
  int cur = 0;
};

struct dem_adjoint {
  dem_sim* sim = nullptr;
  uint32_t K = 100;
  uint32_t steps = 0;
  std::vector<AdjCheckpoint> ckpt;
  AdjCheckpoint final_state;
  // segment slots: K x n
  float2 *spos = nullptr, *svel = nullptr, *sacc = nullptr;
  float *somega = nullptr, *sradius = nullptr, *salpha = nullptr;
  uint32_t *stype = nullptr, *sindex = nullptr, *skey = nullptr;
  size_t slot_off = 0;  // the slot the step hook fills
  // per slot of the segment: the step was re-sorted in place (its index is
  // the identity except at a few slots)
  std::vector<char> slot_inplace;
  // the slots hold the last segment as dem_adj_forward recorded it (until a
  // backward pass recomputes other segments into them)
  bool last_recorded = false;
  // dem_adj_forward calls started so far, and the context's parameter
  // generation the last complete forward ran under (0: none complete)
  uint64_t generation = 0;
  uint64_t ctx_gen = 0;
  // adjoint state (FP64)
  double2 *lx = nullptr, *lv = nullptr, *lxp = nullptr, *lvp = nullptr, *lF = nullptr;
  double *lw = nullptr, *lwp = nullptr, *lM = nullptr, *grad = nullptr;
  std::vector<void*> allocs;
};

namespace {

template <class T>
int adj_alloc(dem_adjoint* a, T** p, size_t count) {
  hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1));
  if (e != hipSuccess) return hip_fail(a->sim->ctx, e, "hipMalloc(adjoint)");
  a->allocs.push_back((void*)*p);
  return DEM_OK;
}

int alloc_checkpoint(dem_adjoint* a, AdjCheckpoint& c) {
  const uint32_t n = a->sim->n;
  int r = DEM_OK;
  if (!r) r = adj_alloc(a, &c.pos, n);
  if (!r) r = adj_alloc(a, &c.vel, n);
  if (!r) r = adj_alloc(a, &c.theta, n);
  if (!r) r = adj_alloc(a, &c.omega, n);
  if (!r) r = adj_alloc(a, &c.id, n);
  if (!r) r = adj_alloc(a, &c.type, n);
  if (!r) r = adj_alloc(a, &c.loc, n);
  if (!r && a->sim->rad) r = adj_alloc(a, &c.radius, n);
  return r;
}

hipError_t copy_state(dem_sim* m, AdjCheckpoint& c, bool to_checkpoint, hipStream_t s) {
  const size_t n = m->n;
  const int b = to_checkpoint ? m->cur : c.cur;
  struct Item { void* ck; void* st; size_t bytes; } items[] = {
      {c.pos, m->pos[b], 8 * n},   {c.vel, m->vel[b], 8 * n}, {c.theta, m->theta[b], 4 * n},
      {c.omega, m->omega[b], 4 * n}, {c.id, m->id[b], 4 * n},     {c.type, m->type[b], 4 * n},
      {c.loc, m->loc[b], 4 * n},   {c.radius, m->rad ? m->radius[b] : nullptr, 4 * n}};
  for (const Item& it : items) {
    if (!it.ck || !it.st) continue;
    hipError_t e = to_checkpoint ? hipMemcpyAsync(it.ck, it.st, it.bytes, hipMemcpyDeviceToDevice, s)
                                 : hipMemcpyAsync(it.st, it.ck, it.bytes, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
  }
  if (to_checkpoint) c.cur = m->cur;
  else {
    m->cur = c.cur;
    m->cell_start = m->cs[m->cur];
    m->perm_list = false;
  }
  return hipSuccess;
}

hipError_t store_slot_hook(void* user, dem_sim* m, hipStream_t s, int dyn, int stat, bool list) {
  auto* a = static_cast<dem_adjoint*>(user);
  const size_t off = a->slot_off;
  const uint32_t n = m->n;
  a->slot_inplace[off / n] = list ? 1 : 0;
  hipLaunchKernelGGL(k_store_slot, grid_of(n), dim3(256), 0, s, m->pos[dyn], m->vel[dyn], m->omega[dyn],
                     m->rad ? m->radius[stat] : nullptr, m->type[stat], list ? nullptr : m->sorted_idx,
                     m->sorted_key, m->acc, m->alpha, m->acc_lazy ? m->quiet : nullptr,
                     a->spos + off, a->svel + off, a->somega + off, a->sradius ? a->sradius + off : nullptr,
                     a->stype + off, a->sindex + off, a->skey + off, a->sacc + off, a->salpha + off, n);
  hipError_t e = hipGetLastError();
  // an in-place step: its index is the identity (just stored) except at the
  // listed slots
  if (e == hipSuccess && list) e = launch_shift_index(s, m->shift, a->sindex + off, n, true);
  return e;
}

}  // namespace

extern "C" {

int dem_adj_create(dem_sim* sim, uint32_t checkpoint_every, dem_adjoint** out) {
  if (!sim || !out) return DEM_ERR_INVALID;
  *out = nullptr;
  dem_context* ctx = sim->ctx;
  if (checkpoint_every == 0) return fail(ctx, DEM_ERR_INVALID, "checkpoint_every must be >= 1");
  if (ctx->c.attract) return fail(ctx, DEM_ERR_INVALID, "the adjoint does not cover attraction");
  if (ctx->h_control.enabled) return fail(ctx, DEM_ERR_INVALID, "the adjoint does not cover the controlled particle");
  if (int r = ensure_device(ctx)) return r;
  auto* a = new (std::nothrow) dem_adjoint();
  if (!a) return fail(ctx, DEM_ERR_ALLOC, "host allocation failed");
  a->sim = sim;
  a->K = checkpoint_every;
  const size_t n = sim->n, slots = (size_t)checkpoint_every * n;
  int r = DEM_OK;
  if (!r) r = adj_alloc(a, &a->spos, slots);
  if (!r) r = adj_alloc(a, &a->svel, slots);
  if (!r) r = adj_alloc(a, &a->somega, slots);
  if (!r && sim->rad) r = adj_alloc(a, &a->sradius, slots);
  if (!r) r = adj_alloc(a, &a->stype, slots);
  if (!r) r = adj_alloc(a, &a->sindex, slots);
  if (!r) r = adj_alloc(a, &a->skey, slots);
  if (!r) r = adj_alloc(a, &a->sacc, slots);
  if (!r) r = adj_alloc(a, &a->salpha, slots);
  if (!r) r = adj_alloc(a, &a->lx, n);
  if (!r) r = adj_alloc(a, &a->lv, n);
  if (!r) r = adj_alloc(a, &a->lw, n);
  if (!r) r = adj_alloc(a, &a->lxp, n);
  if (!r) r = adj_alloc(a, &a->lvp, n);
  if (!r) r = adj_alloc(a, &a->lwp, n);
  if (!r) r = adj_alloc(a, &a->lF, n);
  if (!r) r = adj_alloc(a, &a->lM, n);
  if (!r) r = adj_alloc(a, &a->grad, (size_t)kParams * DEM_MAX_TYPES);
  if (!r) r = alloc_checkpoint(a, a->final_state);
  if (r) {
    dem_adj_destroy(a);
    return r;
  }
  *out = a;
  return DEM_OK;
}

int dem_adj_destroy(dem_adjoint* a) {
  if (!a) return DEM_ERR_INVALID;
  (void)hipSetDevice(a->sim->ctx->device);
  (void)hipStreamSynchronize(a->sim->ctx->stream);
  for (void* p : a->allocs) (void)hipFree(p);
  delete a;
  return DEM_OK;
}

int dem_adj_generation(dem_adjoint* a, uint64_t* generation) {
  if (!a || !generation) return DEM_ERR_INVALID;
  *generation = a->generation;
  return DEM_OK;
}

int dem_adj_forward(dem_adjoint* a, uint32_t steps) {
  if (!a) return DEM_ERR_INVALID;
  ++a->generation;
  a->steps = 0;
  a->ctx_gen = 0;
  dem_sim* m = a->sim;
  dem_context* ctx = m->ctx;
  if (int r = sim_check(m)) return r;
  if (ctx->c.attract) return fail(ctx, DEM_ERR_INVALID, "the adjoint does not cover attraction");
  if (ctx->h_control.enabled) return fail(ctx, DEM_ERR_INVALID, "the adjoint does not cover the controlled particle");
  const uint32_t nck = steps / a->K + 1;
  while (a->ckpt.size() < nck) {
    a->ckpt.emplace_back();
    if (int r = alloc_checkpoint(a, a->ckpt.back())) {
      a->ckpt.pop_back();
      return r;
    }
  }
  hipStream_t s = ctx->stream;
  // the last segment's steps are recorded into the slots as they run, so
  // that the backward pass, which starts with that segment, need not
  // recompute it
  const uint32_t last0 = steps ? ((steps - 1) / a->K) * a->K : 0;
  a->last_recorded = false;
  a->slot_inplace.assign(a->K, 0);
  StepHook hook;
  hook.fn = store_slot_hook;
  hook.user = a;
  for (uint32_t k = 0; k < steps; ++k) {
    if (k % a->K == 0) DEM_HIP(ctx, copy_state(m, a->ckpt[k / a->K], true, s), "checkpoint");
    const bool record = k >= last0;
    if (record) a->slot_off = (size_t)(k - last0) * m->n;
    DEM_HIP(ctx, sim_step_on(m, s, nullptr, sim_use_resort(m), record ? &hook : nullptr), "forward step");
    m->resort_ok = true;
  }
  DEM_HIP(ctx, copy_state(m, a->final_state, true, s), "final state");
  a->steps = steps;
  a->last_recorded = steps > 0;
  a->ctx_gen = ctx->generation;
  return DEM_OK;
}

int dem_adj_backward(dem_adjoint* a, const float* grad_pos, const float* grad_vel, float* grad_pos0,
                     float* grad_vel0, double* grad_params) {
  if (!a) return DEM_ERR_INVALID;
  dem_sim* m = a->sim;
  dem_context* ctx = m->ctx;
  if (int r = sim_check(m)) return r;
  if (a->steps == 0) return fail(ctx, DEM_ERR_STATE, "dem_adj_forward has not run");
  if (a->ctx_gen != ctx->generation)
    return fail(ctx, DEM_ERR_STATE, "the context's parameters changed after dem_adj_forward; run it again");
  hipStream_t s = ctx->stream;
  const uint32_t n = m->n;
  const StepConst& c = ctx->c;
  const bool mul = multi(ctx);
  // adjoint of the final state, from gradients given by particle id
  hipLaunchKernelGGL(k_adj_init, grid_of(n), dim3(256), 0, s, a->final_state.id, grad_pos, grad_vel, a->lx, a->lv,
                     a->lw, n);
  DEM_HIP(ctx, hipMemsetAsync(a->grad, 0, sizeof(double) * kParams * c.num_types, s), "grad reset");
  const uint32_t nseg = (a->steps + a->K - 1) / a->K;
  for (int seg = (int)nseg - 1; seg >= 0; --seg) {
    const uint32_t k0 = (uint32_t)seg * a->K, len = std::min(a->K, a->steps - k0);
    // re-run the segment with the owned step (re-sorted after its first
    // step), storing each step's sorted pre-integrate state, keys,
    // permutation and accelerations
    if (seg == (int)nseg - 1 && a->last_recorded) {
      // recorded by the forward pass
    } else {
      DEM_HIP(ctx, copy_state(m, a->ckpt[seg], false, s), "restore checkpoint");
      m->resort_ok = false;
      StepHook hook;
      hook.fn = store_slot_hook;
      hook.user = a;
      a->slot_inplace.assign(a->K, 0);
      for (uint32_t k = 0; k < len; ++k) {
        a->slot_off = (size_t)k * n;
        DEM_HIP(ctx, sim_step_on(m, s, nullptr, sim_use_resort(m), &hook), "segment step");
        m->resort_ok = true;
      }
    }
    a->last_recorded = false;  // the slots are about to be reused
    // reverse through the segment: cell ranges from the stored keys, forces
    // from the stored accelerations
    for (int k = (int)len - 1; k >= 0; --k) {
      const size_t off = (size_t)k * n;
      AdjIn S{a->spos + off, a->svel + off, a->somega + off, a->sradius ? a->sradius + off : nullptr, a->stype + off};
      DEM_HIP(ctx, launch_fill(s, m->cell_start, kEmpty, m->cells), "adj fill");
      hipLaunchKernelGGL(k_cell_bounds, grid_of(n), dim3(256), 0, s, a->skey + off, m->cell_start, m->cell_end, n,
                         m->cells);
#define DEM_ADJ(K, ...)                                                                                 \
  do {                                                                                                  \
    if (m->rad) {                                                                                       \
      if (mul) hipLaunchKernelGGL((K<true, true>), grid_of(n), dim3(256), 0, s, __VA_ARGS__);             \
      else hipLaunchKernelGGL((K<true, false>), grid_of(n), dim3(256), 0, s, __VA_ARGS__);               \
    } else {                                                                                            \
      if (mul) hipLaunchKernelGGL((K<false, true>), grid_of(n), dim3(256), 0, s, __VA_ARGS__);            \
      else hipLaunchKernelGGL((K<false, false>), grid_of(n), dim3(256), 0, s, __VA_ARGS__);              \
    }                                                                                                   \
  } while (0)
      DEM_ADJ(k_adj_integrate, c, S, (const float2*)(a->sacc + off), (const float*)(a->salpha + off),
              (const double2*)a->lx,
              (const double2*)a->lv, (const double*)a->lw, a->lxp, a->lvp, a->lwp, a->lF, a->lM, a->grad, n);
      DEM_ADJ(k_adj_collide, c, S, (const uint32_t*)m->cell_start, (const uint32_t*)m->cell_end,
              (const double2*)a->lF, (const double*)a->lM, a->lxp, a->lvp, a->lwp, a->grad, n);
#undef DEM_ADJ
      if (a->slot_inplace[k] && n > kScatterMax) {
        hipLaunchKernelGGL(k_adj_perm_save, grid_of(n), dim3(256), 0, s, a->sindex + off, a->lxp, a->lvp, a->lwp,
                           a->lx, a->lv, a->lw, n);
        hipLaunchKernelGGL(k_adj_perm_put, grid_of(n), dim3(256), 0, s, a->sindex + off, a->lx, a->lv, a->lw,
                           a->lxp, a->lvp, a->lwp, n);
        std::swap(a->lx, a->lxp);
        std::swap(a->lv, a->lvp);
        std::swap(a->lw, a->lwp);
      } else {
        // the full scatter (for an in-place step too, whose stored index is
        // the identity outside its listed slots: one launch instead of two,
        // which is faster wherever launches cost more than the whole copy)
        hipLaunchKernelGGL(k_adj_permute, grid_of(n), dim3(256), 0, s, a->sindex + off, a->lxp, a->lvp, a->lwp,
                           a->lx, a->lv, a->lw, n);
      }
      DEM_HIP(ctx, hipGetLastError(), "adjoint kernels");
    }
  }
  // lx/lv now follow the order of the first checkpoint
  hipLaunchKernelGGL(k_adj_out, grid_of(n), dim3(256), 0, s, a->ckpt[0].id, a->lx, a->lv, grad_pos0, grad_vel0, n);
  if (grad_params)
    DEM_HIP(ctx, hipMemcpyAsync(grad_params, a->grad, sizeof(double) * kParams * c.num_types, hipMemcpyDeviceToDevice, s),
            "grad copy");
  DEM_HIP(ctx, copy_state(m, a->final_state, false, s), "restore final state");
  m->resort_ok = false;  // the recomputation rewrote the cell arrays
  DEM_HIP(ctx, hipStreamSynchronize(s), "adjoint sync");
  return DEM_OK;
}

}  // extern "C"
