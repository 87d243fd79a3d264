// adjoint_oracle.hpp -- FP64 host adjoint of DemSo's collide + integrate.
//
// TEST INFRASTRUCTURE ONLY (the reference has no adjoint; this is the host
// FP64 restatement the HIP adjoint is checked against, itself checked against
// central finite differences of the FP64 forward in tests/test_adjoint_cpu.py).
//
// Forward: the oracle's step (dem_oracle.hpp) in double, keeping for every
// step the sorted pre-collide state and the sort permutation.  Backward, per
// step in reverse: recompute cells and forces, reverse integrate
// (src/particles_kernel.cuh:580-627, walls and the bottom-wall law :388-456
// included), reverse the pair law (:235-307) over the 3x3 neighbourhood, and
// scatter through the permutation.  Switches (contact, approach-only damping,
// stick/slip, wall hits) are frozen at the forward's decisions.  In exact
// arithmetic the pair law is F = s n + f_t t, tau = f_t with
//   s   = -kn (R - dist) + [vn < 0] cn vn,     vn  = (vB - vA) . n
//   c_t = (vB - vA) . t - (rA wA + rB wB)
//   f_t = kt dt c_t (stick)  |  mu |s| sgn(c_t) (slip)  |  0 (mu = 0)
// (cv = vt - w t is parallel to t), force on B = -F, moments rA f_t and rB f_t;
// those closed forms are what is differentiated.  Gradients: initial
// positions/velocities, and per material (kn, kt, cn, mu) through the series
// combination of src/particles_kernel.cuh:269-275.
#pragma once
#include "dem_oracle.hpp"

namespace orc {

enum { P_KN = 0, P_KT = 1, P_CN = 2, P_MU = 3, NPARAM = 4 };

struct SeriesGrad {  // d series(a,b)/da, d/db
  double da, db;
};
inline SeriesGrad series_grad(double a, double b) {
  if (a == 0.0 || b == 0.0) return {0.0, 0.0};
  const double s = (a + b) * (a + b);
  return {b * b / s, a * a / s};
}

inline double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }

struct AdjSnapshot {
  std::vector<V2<double>> pos, vel;
  std::vector<double> omega, radius;
  std::vector<uint32_t> type, index, id;
};

class AdjointOracle {
 public:
  System sys;
  std::vector<Material> mat;
  std::vector<Instance> inst;
  uint32_t n = 0, num_cells = 0;
  bool has_radius = false;
  std::vector<AdjSnapshot> snaps;
  Sim<double> sim;
  std::vector<uint32_t> id0;  // ids of the initial state (input order)

  void init(const System& s, uint32_t nt, uint32_t ni, uint32_t n_, uint32_t cells, bool rad) {
    mat.assign(s.mat, s.mat + nt);
    inst.assign(s.inst, s.inst + ni);
    sys = s;
    sys.mat = mat.data();
    sys.inst = inst.data();
    n = n_;
    num_cells = cells;
    has_radius = rad;
    sim.sys = sys;
    sim.resize(n, cells, rad);
  }

  // Forward `steps` steps from the sim's current buffer, keeping snapshots.
  void forward(uint32_t steps) {
    snaps.clear();
    id0 = sim.id[sim.cur];
    for (uint32_t k = 0; k < steps; ++k) {
      State<double> o = sim.st(sim.cur), s = sim.st(sim.cur ^ 1);
      calc_hash<double>(sim.sys, o.pos, o.type, sim.hash.data(), sim.index.data(), n);
      sort_particles(sim.hash.data(), sim.index.data(), n);
      reorder_and_find_cell_start<double>(sim.cell_start.data(), sim.cell_end.data(), s, sim.hash.data(),
                                          sim.index.data(), o, n, num_cells);
      AdjSnapshot sn;
      sn.pos.assign(s.pos, s.pos + n);
      sn.vel.assign(s.vel, s.vel + n);
      sn.omega.assign(s.omega, s.omega + n);
      sn.radius = has_radius ? std::vector<double>(s.radius, s.radius + n) : std::vector<double>();
      sn.type.assign(s.type, s.type + n);
      sn.id.assign(s.id, s.id + n);
      sn.index = sim.index;
      snaps.push_back(std::move(sn));
      collide<double>(sim.sys, s.pos, s.vel, sim.acc.data(), s.omega, sim.alpha.data(), s.type, s.radius,
                      sim.cell_start.data(), sim.cell_end.data(), n, sim.pressure.data());
      integrate<double>(sim.sys, s.pos, s.vel, sim.acc.data(), s.theta, s.omega, sim.alpha.data(), s.type,
                        s.radius, sim.pressure.data(), n);
      sim.cur ^= 1;
    }
  }

  // Backward from dL/dpos, dL/dvel of the final state given BY PARTICLE ID.
  // Returns dL/dpos0, dL/dvel0 by particle id and per-material gradients.
  void backward(const double* gpos_by_id, const double* gvel_by_id, double* gpos0_by_id,
                double* gvel0_by_id, double* gparams) const {
    const uint32_t K = (uint32_t)snaps.size();
    std::vector<V2<double>> lx(n), lv(n);
    std::vector<double> lw(n, 0.0);
    std::fill(gparams, gparams + NPARAM * mat.size(), 0.0);
    // final state order = sorted order of the last step
    const std::vector<uint32_t>& idT = K ? snaps[K - 1].id : id0;
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t d = idT[i];
      lx[i] = V2<double>{gpos_by_id ? gpos_by_id[2 * d] : 0.0, gpos_by_id ? gpos_by_id[2 * d + 1] : 0.0};
      lv[i] = V2<double>{gvel_by_id ? gvel_by_id[2 * d] : 0.0, gvel_by_id ? gvel_by_id[2 * d + 1] : 0.0};
    }
    std::vector<uint32_t> cs(num_cells), ce(num_cells), hash(n), idx(n);
    std::vector<V2<double>> acc(n), lF(n), lxp(n), lvp(n);
    std::vector<double> alpha(n), pres(n), lM(n), lwp(n);
    for (int k = (int)K - 1; k >= 0; --k) {
      const AdjSnapshot& S = snaps[k];
      const double* rad = has_radius ? S.radius.data() : nullptr;
      // cells of the sorted state (its keys are sorted)
      calc_hash<double>(sys, S.pos.data(), S.type.data(), hash.data(), idx.data(), n);
      std::fill(cs.begin(), cs.end(), EMPTY);
      for (uint32_t i = 0; i < n; ++i) {
        if (i == 0 || hash[i] != hash[i - 1]) {
          cs[hash[i]] = i;
          if (i > 0) ce[hash[i - 1]] = i;
        }
        if (i == n - 1) ce[hash[i]] = n;
      }
      collide<double>(sys, S.pos.data(), S.vel.data(), acc.data(), S.omega.data(), alpha.data(), S.type.data(),
                      rad, cs.data(), ce.data(), n, pres.data());
      // reverse integrate
      for (uint32_t i = 0; i < n; ++i)
        integrate_adjoint(i, S, rad, acc[i], alpha[i], lx[i], lv[i], lw[i], lxp[i], lvp[i], lwp[i], lF[i], lM[i],
                          gparams);
      // reverse collide
      for (uint32_t i = 0; i < n; ++i) collide_adjoint(i, S, rad, cs, ce, lF, lM, lxp[i], lvp[i], lwp[i], gparams);
      // reverse the permutation: sorted slot i came from input slot index[i]
      for (uint32_t i = 0; i < n; ++i) {
        const uint32_t src = S.index[i];
        lx[src] = lxp[i];
        lv[src] = lvp[i];
        lw[src] = lwp[i];
      }
    }
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t d = id0[i];
      if (gpos0_by_id) { gpos0_by_id[2 * d] = lx[i].x; gpos0_by_id[2 * d + 1] = lx[i].y; }
      if (gvel0_by_id) { gvel0_by_id[2 * d] = lv[i].x; gvel0_by_id[2 * d + 1] = lv[i].y; }
    }
  }

 private:
  Body<double> body_of(uint32_t i, const AdjSnapshot& S, const double* rad) const {
    return body<double>(sys, rad, i, S.type[i]);
  }

  void add_param(double* g, uint32_t type, int p, double v) const {
    const uint32_t t = type < mat.size() ? type : (uint32_t)mat.size() - 1;
    g[NPARAM * t + p] += v;
  }

  // Reverse of integrateSystemD for particle i of snapshot S (sorted order).
  void integrate_adjoint(uint32_t i, const AdjSnapshot& S, const double* rad, V2<double> a, double alpha,
                         V2<double> lpo, V2<double> lvo, double lwo, V2<double>& lp, V2<double>& lv, double& lw,
                         V2<double>& lF, double& lM, double* g) const {
    const uint32_t t = S.type[i];
    const Material& M = sys.material(t);
    const Instance& in = sys.instance_of(t);
    const Body<double> b = body_of(i, S, rad);
    const double dt = sys.dt, gx = sys.g[0], gy = sys.g[1];
    const double bd = -double(M.bdamp);
    // forward replay with switch decisions
    V2<double> v1 = S.vel[i] + (mk(gx, gy) + a) * dt;
    V2<double> p1 = S.pos[i] + v1 * dt;
    double w1 = S.omega[i] + alpha * dt;
    const double Lx = in.box[0], Ly = in.box[1];
    bool fr = false, fl = false, ft = false, fb = false;
    if (p1.x > Lx - b.r) { p1.x = Lx - b.r; v1.x *= bd; fr = true; }
    if (p1.x < b.r) { p1.x = b.r; v1.x *= bd; fl = true; }
    if (!(sys.flags & OPEN_TOP) && p1.y > Ly - b.r) { p1.y = Ly - b.r; v1.y *= bd; ft = true; }
    // bottom-wall law intermediates
    bool contact = false, damp = false, fric = false, stick = false;
    double sy = 0, cx = 0, P = 0, ktw = 0;
    V2<double> force0{0, 0};
    if (p1.y < b.r) {
      fb = true;
      const double dist = std::fabs(p1.y);
      if (dist < b.r) {
        contact = true;
        sy = p1.y > 0 ? 1.0 : -1.0;  // n = (0, -sy), t = (sy, 0)
        const double vn = sy * v1.y;
        force0 = mk(-b.m * gx, -b.m * gy);
        if (vn < 0) { damp = true; force0.y += -double(M.cn) * v1.y; }
        P = std::sqrt(force0.x * force0.x + force0.y * force0.y);
        if (M.mu != 0.0f) {
          fric = true;
          ktw = series<double>(M.kt, sys.bkt);
          cx = -v1.x - sy * b.r * w1;
          stick = std::fabs(ktw * cx * dt) < double(M.mu) * P;
        }
      }
    }
    // reverse
    V2<double> lp1 = lpo, lv1 = lvo;
    double lw1 = lwo;
    if (fb) {
      // (c) p.y = r, v.y *= bd
      lv1.y *= bd;
      lp1.y = 0.0;
      // (b) v += F/m dt, p += F/m dt^2, w += aw dt
      V2<double> lFw = (lv1 * dt + lp1 * (dt * dt)) / b.m;
      const double law = lw1 * dt;
      if (contact) {
        if (fric) {
          const double lFt = lFw.x + law * b.r * sy / b.inertia;
          if (stick) {
            const double lc = lFt * ktw * dt;
            const SeriesGrad sg = series_grad(M.kt, sys.bkt);
            add_param(g, t, P_KT, lFt * dt * cx * sg.da);
            lv1.x += -lc;
            lw1 += -sy * b.r * lc;
          } else {
            const double lP = lFt * double(M.mu) * sgn(cx);
            add_param(g, t, P_MU, lFt * P * sgn(cx));
            if (damp && P > 0) {
              const double lf0y = lP * force0.y / P;
              lv1.y += lf0y * (-double(M.cn));
              add_param(g, t, P_CN, lf0y * (-v1.y));
            }
          }
        } else if (damp) {
          lv1.y += lFw.y * (-double(M.cn));
          add_param(g, t, P_CN, lFw.y * (-v1.y));
        }
      }
    }
    if (ft) { lv1.y *= bd; lp1.y = 0.0; }
    if (fl) { lv1.x *= bd; lp1.x = 0.0; }
    if (fr) { lv1.x *= bd; lp1.x = 0.0; }
    // stage 1: v1 = v + (g+a)dt; p1 = p + v1 dt; w1 = w + alpha dt
    lv1 = lv1 + lp1 * dt;
    lp = lp1;
    lv = lv1;
    lw = lw1;
    lF = lv1 * (dt / b.m);
    lM = lw1 * dt / b.inertia;
  }

  // Reverse of the pair law for particle i (sorted order) over its neighbours.
  void collide_adjoint(uint32_t i, const AdjSnapshot& S, const double* rad, const std::vector<uint32_t>& cs,
                       const std::vector<uint32_t>& ce, const std::vector<V2<double>>& lF,
                       const std::vector<double>& lM, V2<double>& lx, V2<double>& lv, double& lw,
                       double* g) const {
    const bool clamp = sys.flags & GRID_CLAMP;
    const uint32_t tA = S.type[i];
    const Instance& in = sys.instance_of(tA);
    const Body<double> bA = body_of(i, S, rad);
    const Material& A = sys.material(tA);
    int gx, gy;
    calc_grid_pos(in, S.pos[i], gx, gy);
    if (clamp) own_cell(in, true, gx, gy, gx, gy);
    const double dt = sys.dt;
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x) {
        int cx = gx + x, cy = gy + y;
        if (clamp) {
          if (cx < 0 || cy < 0 || cx >= int(in.grid[0]) || cy >= int(in.grid[1])) continue;
        } else {
          own_cell(in, false, int(uint32_t(gx) + uint32_t(x)), int(uint32_t(gy) + uint32_t(y)), cx, cy);
        }
        const uint32_t h = cell_key(in, cx, cy);
        if (cs[h] == EMPTY) continue;
        for (uint32_t j = cs[h]; j < std::min(ce[h], n); ++j) {
          if (j == i) continue;
          const uint32_t tB = S.type[j];
          const Material& B = sys.material(tB);
          const double rB = body_of(j, S, rad).r;
          const V2<double> dx = S.pos[j] - S.pos[i];
          const double dist = std::sqrt(dot(dx, dx));
          const double R = bA.r + rB;
          if (!(dist < R)) continue;
          const V2<double> nn = dx / dist;
          const V2<double> tt = mk(-nn.y, nn.x);
          const V2<double> rv = S.vel[j] - S.vel[i];
          const double vn = dot(rv, nn), vt = dot(rv, tt);
          const double kn = series<double>(A.kn, B.kn), kt = series<double>(A.kt, B.kt);
          const double cn = series<double>(A.cn, B.cn), mu = (double(A.mu) + double(B.mu)) / 2.0;
          const bool damp = vn < 0;
          const double delta = R - dist;
          const double s = -kn * delta + (damp ? cn * vn : 0.0);
          const double P = std::fabs(s);
          bool fric = mu != 0.0, stick = false;
          double ct = 0.0, ft = 0.0;
          if (fric) {
            ct = vt - (bA.r * S.omega[i] + rB * S.omega[j]);
            // the forward's stick test, evaluated as collide_spheres does
            V2<double> force = (-kn * delta) * nn;
            if (damp) force += (cn * vn) * nn;
            const V2<double> relVel_t = rv - vn * nn;
            const V2<double> cv = relVel_t - (bA.r * S.omega[i] + rB * S.omega[j]) * tt;
            const V2<double> Ftrial = (kt * cv) * dt;
            stick = length(Ftrial) < mu * length(force);
            ft = stick ? kt * dt * ct : mu * P * sgn(ct);
          }
          const V2<double> gF = lF[i] - lF[j];
          const double gtau = bA.r * lM[i] + rB * lM[j];
          const double a = dot(gF, nn), gFt = dot(gF, tt), bb = gFt + gtau;
          // ds/ddx, ds/drv
          const V2<double> ds_dx = nn * kn + tt * ((damp ? cn : 0.0) * vt / dist);
          const V2<double> ds_drv = nn * (damp ? cn : 0.0);
          V2<double> dft_dx{0, 0}, dft_drv{0, 0};
          double dft_dw = 0.0, slip_k = 0.0;
          if (fric) {
            if (stick) {
              dft_dx = tt * (-kt * dt * vn / dist);
              dft_drv = tt * (kt * dt);
              dft_dw = -kt * dt;
            } else {
              slip_k = mu * sgn(s) * sgn(ct);
              dft_dx = ds_dx * slip_k;
              dft_drv = ds_drv * slip_k;
            }
          }
          const V2<double> dPhi_dx = ds_dx * a + tt * (s * gFt / dist) + dft_dx * bb - tt * (ft * a / dist);
          const V2<double> dPhi_drv = ds_drv * a + dft_drv * bb;
          const double dPhi_dw = dft_dw * bb;
          lx = lx - dPhi_dx;
          lv = lv - dPhi_drv;
          lw += bA.r * dPhi_dw;
          if (i < j) {
            const double dkn = a * (-delta) + bb * slip_k * (-delta);
            const double dcn = damp ? (a * vn + bb * slip_k * vn) : 0.0;
            const double dkt = (fric && stick) ? bb * dt * ct : 0.0;
            const double dmu = (fric && !stick) ? bb * P * sgn(ct) : 0.0;
            const SeriesGrad gk = series_grad(A.kn, B.kn), gt = series_grad(A.kt, B.kt), gc = series_grad(A.cn, B.cn);
            add_param(g, tA, P_KN, dkn * gk.da);
            add_param(g, tB, P_KN, dkn * gk.db);
            add_param(g, tA, P_KT, dkt * gt.da);
            add_param(g, tB, P_KT, dkt * gt.db);
            add_param(g, tA, P_CN, dcn * gc.da);
            add_param(g, tB, P_CN, dcn * gc.db);
            add_param(g, tA, P_MU, dmu * 0.5);
            add_param(g, tB, P_MU, dmu * 0.5);
          }
        }
      }
  }
};

}  // namespace orc
