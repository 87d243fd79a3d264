// oracle_capi.cpp -- extern "C" entry points over dem_oracle.hpp for ctypes.
//
// TEST INFRASTRUCTURE ONLY (see dem_oracle.hpp).  Loaded by the tests, the
// smoke check and the benchmark's CPU-baseline leg; never by the product.
#include "adjoint_oracle.hpp"
#include "render_oracle.hpp"

#include <chrono>
#include <new>
#ifdef _OPENMP
#include <omp.h>
#endif

using orc::Instance;
using orc::Material;
using orc::System;
using orc::V2;

namespace {

template <class T> struct OwnedSim {
  orc::Sim<T> sim;
  std::vector<Material> mat;
  std::vector<Instance> inst;
};

template <class T>
OwnedSim<T>* sim_create(const System* s, uint32_t num_types, uint32_t num_instances, uint32_t n,
                        uint32_t num_cells, int has_radius) {
  auto* o = new (std::nothrow) OwnedSim<T>();
  if (!o) return nullptr;
  o->mat.assign(s->mat, s->mat + num_types);
  o->inst.assign(s->inst, s->inst + num_instances);
  o->sim.sys = *s;
  o->sim.sys.mat = o->mat.data();
  o->sim.sys.inst = o->inst.data();
  o->sim.resize(n, num_cells, has_radius != 0);
  return o;
}

template <class T, class S>
void copy_in(std::vector<T>& dst, const S* src, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] = T(src[i]);
}

template <class T, class S>
void copy_in2(std::vector<V2<T>>& dst, const S* src, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] = V2<T>{T(src[2 * i]), T(src[2 * i + 1])};
}

template <class T, class D>
void copy_out(D* dst, const std::vector<T>& src) {
  if (!dst) return;
  for (size_t i = 0; i < src.size(); ++i) dst[i] = D(src[i]);
}

template <class T, class D>
void copy_out2(D* dst, const std::vector<V2<T>>& src) {
  if (!dst) return;
  for (size_t i = 0; i < src.size(); ++i) { dst[2 * i] = D(src[i].x); dst[2 * i + 1] = D(src[i].y); }
}

// State arrays come in/out as the same element type S (float or double).
template <class T, class S>
int sim_set_state(OwnedSim<T>* o, const S* pos, const S* vel, const S* theta, const S* omega,
                  const uint32_t* id, const uint32_t* type, const uint32_t* loc, const S* radius) {
  auto& m = o->sim;
  const int c = m.cur;
  const size_t n = m.n;
  copy_in2(m.pos[c], pos, n);
  copy_in2(m.vel[c], vel, n);
  copy_in(m.theta[c], theta, n);
  copy_in(m.omega[c], omega, n);
  copy_in(m.id[c], id, n);
  copy_in(m.type[c], type, n);
  copy_in(m.loc[c], loc, n);
  if (m.has_radius) {
    if (!radius) return -1;
    copy_in(m.radius[c], radius, n);
  }
  return 0;
}

template <class T, class S>
int sim_get_state(OwnedSim<T>* o, S* pos, S* vel, S* theta, S* omega, uint32_t* id, uint32_t* type,
                  uint32_t* loc, S* radius, S* acc, S* alpha, S* pressure, uint32_t* hash,
                  uint32_t* index, uint32_t* cell_start, uint32_t* cell_end) {
  auto& m = o->sim;
  const int c = m.cur;
  copy_out2(pos, m.pos[c]);
  copy_out2(vel, m.vel[c]);
  copy_out(theta, m.theta[c]);
  copy_out(omega, m.omega[c]);
  copy_out(id, m.id[c]);
  copy_out(type, m.type[c]);
  copy_out(loc, m.loc[c]);
  if (m.has_radius) copy_out(radius, m.radius[c]);
  copy_out2(acc, m.acc);
  copy_out(alpha, m.alpha);
  copy_out(pressure, m.pressure);
  copy_out(hash, m.hash);
  copy_out(index, m.index);
  copy_out(cell_start, m.cell_start);
  copy_out(cell_end, m.cell_end);
  return 0;
}

}  // namespace

extern "C" {

int orc_abi_version(void) { return 1; }

int orc_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}

int orc_sizeof_system(void) { return int(sizeof(System)); }
int orc_sizeof_material(void) { return int(sizeof(Material)); }
int orc_sizeof_instance(void) { return int(sizeof(Instance)); }

// ---- single stages, FP32 ---------------------------------------------------

int orc_calc_hash(const System* s, const float* pos, const uint32_t* type, uint32_t* hash,
                  uint32_t* index, uint32_t n) {
  orc::calc_hash<float>(*s, reinterpret_cast<const V2<float>*>(pos), type, hash, index, n);
  return 0;
}

int orc_sort_particles(uint32_t* hash, uint32_t* index, uint32_t n) {
  orc::sort_particles(hash, index, n);
  return 0;
}

int orc_reorder(uint32_t* cell_start, uint32_t* cell_end, float* s_pos, float* s_vel, float* s_theta,
                float* s_omega, uint32_t* s_id, uint32_t* s_loc, uint32_t* s_type, float* s_radius,
                const uint32_t* hash, const uint32_t* index, const float* o_pos, const float* o_vel,
                const float* o_theta, const float* o_omega, const uint32_t* o_id,
                const uint32_t* o_type, const float* o_radius, uint32_t n, uint32_t num_cells) {
  orc::State<float> srt{reinterpret_cast<V2<float>*>(s_pos), reinterpret_cast<V2<float>*>(s_vel),
                        s_theta, s_omega, s_id, s_type, s_loc, s_radius};
  orc::State<float> old{reinterpret_cast<V2<float>*>(const_cast<float*>(o_pos)),
                        reinterpret_cast<V2<float>*>(const_cast<float*>(o_vel)),
                        const_cast<float*>(o_theta), const_cast<float*>(o_omega),
                        const_cast<uint32_t*>(o_id), const_cast<uint32_t*>(o_type), nullptr,
                        const_cast<float*>(o_radius)};
  if ((s_radius == nullptr) != (o_radius == nullptr)) return -1;
  orc::reorder_and_find_cell_start<float>(cell_start, cell_end, srt, hash, index, old, n, num_cells);
  return 0;
}

int orc_collide(const System* s, const float* pos, const float* vel, float* acc, const float* omega,
                float* alpha, const uint32_t* type, const float* radius, const uint32_t* cell_start,
                const uint32_t* cell_end, uint32_t n, float* pressure) {
  orc::collide<float>(*s, reinterpret_cast<const V2<float>*>(pos),
                      reinterpret_cast<const V2<float>*>(vel), reinterpret_cast<V2<float>*>(acc),
                      omega, alpha, type, radius, cell_start, cell_end, n, pressure);
  return 0;
}

int orc_integrate(const System* s, float* pos, float* vel, const float* acc, float* theta,
                  float* omega, const float* alpha, const uint32_t* type, const float* radius,
                  float* pressure, uint32_t n) {
  orc::integrate<float>(*s, reinterpret_cast<V2<float>*>(pos), reinterpret_cast<V2<float>*>(vel),
                        reinterpret_cast<const V2<float>*>(acc), theta, omega, alpha, type, radius,
                        pressure, n);
  return 0;
}

// ---- owned simulations (FP32 and FP64 state) -------------------------------

void* orc_sim_create_f32(const System* s, uint32_t nt, uint32_t ni, uint32_t n, uint32_t cells, int rad) {
  return sim_create<float>(s, nt, ni, n, cells, rad);
}
void* orc_sim_create_f64(const System* s, uint32_t nt, uint32_t ni, uint32_t n, uint32_t cells, int rad) {
  return sim_create<double>(s, nt, ni, n, cells, rad);
}
void orc_sim_destroy_f32(void* p) { delete static_cast<OwnedSim<float>*>(p); }
void orc_sim_destroy_f64(void* p) { delete static_cast<OwnedSim<double>*>(p); }

// State is passed in as float arrays for both precisions (the FP64 run starts
// from the same FP32 initial state) and read back as float or double.
int orc_sim_set_state_f32(void* p, const float* pos, const float* vel, const float* theta,
                          const float* omega, const uint32_t* id, const uint32_t* type,
                          const uint32_t* loc, const float* radius) {
  return sim_set_state(static_cast<OwnedSim<float>*>(p), pos, vel, theta, omega, id, type, loc, radius);
}
int orc_sim_set_state_f64(void* p, const float* pos, const float* vel, const float* theta,
                          const float* omega, const uint32_t* id, const uint32_t* type,
                          const uint32_t* loc, const float* radius) {
  return sim_set_state(static_cast<OwnedSim<double>*>(p), pos, vel, theta, omega, id, type, loc, radius);
}

double orc_sim_step_f32(void* p, uint32_t steps) {
  auto* o = static_cast<OwnedSim<float>*>(p);
  auto t0 = std::chrono::steady_clock::now();
  for (uint32_t k = 0; k < steps; ++k) o->sim.step();
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
double orc_sim_step_f64(void* p, uint32_t steps) {
  auto* o = static_cast<OwnedSim<double>*>(p);
  auto t0 = std::chrono::steady_clock::now();
  for (uint32_t k = 0; k < steps; ++k) o->sim.step();
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int orc_sim_get_state_f32(void* p, float* pos, float* vel, float* theta, float* omega, uint32_t* id,
                          uint32_t* type, uint32_t* loc, float* radius, float* acc, float* alpha,
                          float* pressure, uint32_t* hash, uint32_t* index, uint32_t* cell_start,
                          uint32_t* cell_end) {
  return sim_get_state(static_cast<OwnedSim<float>*>(p), pos, vel, theta, omega, id, type, loc, radius,
                       acc, alpha, pressure, hash, index, cell_start, cell_end);
}
int orc_sim_get_state_f64(void* p, double* pos, double* vel, double* theta, double* omega,
                          uint32_t* id, uint32_t* type, uint32_t* loc, double* radius, double* acc,
                          double* alpha, double* pressure, uint32_t* hash, uint32_t* index,
                          uint32_t* cell_start, uint32_t* cell_end) {
  return sim_get_state(static_cast<OwnedSim<double>*>(p), pos, vel, theta, omega, id, type, loc,
                       radius, acc, alpha, pressure, hash, index, cell_start, cell_end);
}

// ---- controlled particle --------------------------------------------------------

int orc_sim_set_control_f32(void* p, const orc::Control* c) {
  static_cast<OwnedSim<float>*>(p)->sim.ctl = *c;
  return 0;
}
int orc_sim_set_control_f64(void* p, const orc::Control* c) {
  static_cast<OwnedSim<double>*>(p)->sim.ctl = *c;
  return 0;
}
// Current controlled-particle state and the last step's reaction
// (force x, y and moment on the controlled particle).
int orc_sim_get_control_f32(void* p, orc::Control* c, double* reaction) {
  auto& m = static_cast<OwnedSim<float>*>(p)->sim;
  if (c) *c = m.ctl;
  for (int k = 0; k < 3 && reaction; ++k) reaction[k] = m.reaction[k];
  return 0;
}
int orc_sim_get_control_f64(void* p, orc::Control* c, double* reaction) {
  auto& m = static_cast<OwnedSim<double>*>(p)->sim;
  if (c) *c = m.ctl;
  for (int k = 0; k < 3 && reaction; ++k) reaction[k] = m.reaction[k];
  return 0;
}

// collide with the controlled particle (FP32), the per-particle reaction
// vectors and their reduction (force x, y, moment) into reaction[3].
int orc_collide_controlled(const System* s, const float* pos, const float* vel, float* acc, const float* omega,
                           float* alpha, const uint32_t* type, const float* radius, const uint32_t* cell_start,
                           const uint32_t* cell_end, uint32_t n, float* pressure, const orc::Control* c,
                           float* react_x, float* react_y, float* react_m, float* reaction) {
  orc::collide<float>(*s, reinterpret_cast<const V2<float>*>(pos), reinterpret_cast<const V2<float>*>(vel),
                      reinterpret_cast<V2<float>*>(acc), omega, alpha, type, radius, cell_start, cell_end, n,
                      pressure, c, react_x, react_y, react_m);
  reaction[0] = orc::reduce_reaction(react_x, n);
  reaction[1] = orc::reduce_reaction(react_y, n);
  reaction[2] = orc::reduce_reaction(react_m, n);
  return 0;
}

// ---- FP64 adjoint -------------------------------------------------------------

void* orc_adj_create(const System* s, uint32_t nt, uint32_t ni, uint32_t n, uint32_t cells, int rad) {
  auto* a = new (std::nothrow) orc::AdjointOracle();
  if (!a) return nullptr;
  a->init(*s, nt, ni, n, cells, rad != 0);
  return a;
}
void orc_adj_destroy(void* p) { delete static_cast<orc::AdjointOracle*>(p); }

int orc_adj_set_state(void* p, const float* pos, const float* vel, const float* theta, const float* omega,
                      const uint32_t* id, const uint32_t* type, const uint32_t* loc, const float* radius) {
  auto* a = static_cast<orc::AdjointOracle*>(p);
  auto& m = a->sim;
  const int c = m.cur;
  for (uint32_t i = 0; i < m.n; ++i) {
    m.pos[c][i] = V2<double>{pos[2 * i], pos[2 * i + 1]};
    m.vel[c][i] = V2<double>{vel[2 * i], vel[2 * i + 1]};
    m.theta[c][i] = theta[i];
    m.omega[c][i] = omega[i];
    m.id[c][i] = id[i];
    m.type[c][i] = type[i];
    m.loc[c][i] = loc[i];
    if (m.has_radius) m.radius[c][i] = radius[i];
  }
  return 0;
}

int orc_adj_forward(void* p, uint32_t steps) {
  static_cast<orc::AdjointOracle*>(p)->forward(steps);
  return 0;
}

// Final positions/velocities by particle id.
int orc_adj_final_state(void* p, double* pos_by_id, double* vel_by_id) {
  auto* a = static_cast<orc::AdjointOracle*>(p);
  auto& m = a->sim;
  const int c = m.cur;
  for (uint32_t i = 0; i < m.n; ++i) {
    const uint32_t d = m.id[c][i];
    if (d >= m.n) return -1;
    pos_by_id[2 * d] = m.pos[c][i].x;
    pos_by_id[2 * d + 1] = m.pos[c][i].y;
    if (vel_by_id) {
      vel_by_id[2 * d] = m.vel[c][i].x;
      vel_by_id[2 * d + 1] = m.vel[c][i].y;
    }
  }
  return 0;
}

int orc_adj_backward(void* p, const double* gpos, const double* gvel, double* gpos0, double* gvel0,
                     double* gparams) {
  static_cast<orc::AdjointOracle*>(p)->backward(gpos, gvel, gpos0, gvel0, gparams);
  return 0;
}

// images: plotParticles and updatePressureScale (render_oracle.hpp)
void orc_plot_particles(uint8_t* img, const float* pos, const float* theta, const uint32_t* type, const float* radius,
                        uint32_t n, const orc::RenderType* types, uint32_t num_types, const orc::RenderParams* rp,
                        const orc::Control* ctl, const float* pressure) {
  orc::plot(img, pos, theta, type, radius, n, types, num_types, *rp, ctl, pressure);
}

float orc_pressure_scale(const float* p, uint32_t n) { return orc::pressure_scale(p, n); }

}  // extern "C"
