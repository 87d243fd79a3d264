// prepare.cpp -- PrepareSim (src/main.cu:54-369) without the renderer.
#include "prepare.hpp"

#include <cmath>
#include <string>

#include "scene_rng.h"

namespace dsml {

void reference_grid(float lx, float ly, float max_radius, uint32_t* gx, uint32_t* gy) {
  if (!(max_radius > 0.0f)) throw DsmlError("no particle type has a positive radius");
  auto axis = [&](float length) -> uint32_t {
    const float cells = length / (4.0f * max_radius);
    if (!(cells >= 1.0f) || cells >= 2147483648.0f)
      throw DsmlError("box side " + std::to_string(length) + " is not between 4 max radii and 2^31 cells");
    const uint32_t grid = (uint32_t)cells;
    int temp = (int)std::log2((double)grid);
    const float cell = lx / (float)(1u << temp);  // the x side for both axes, as src/main.cu:231
    if (cell / 2.0f <= 1.2f * max_radius) temp -= 1;
    else if (cell / 2.0f >= 3.0f * max_radius) temp += 1;
    if (temp < 0 || temp > 30) throw DsmlError("grid sizing out of range");
    return 1u << temp;
  };
  *gx = axis(lx);
  *gy = axis(ly);
}

PreparedScene prepare_sim(const DEMSimulation& sim, uint64_t seed) {
  PreparedScene out;
  const DEMParticles& P = sim.particles;
  const auto& T = sim.properties.particleTypes;
  if (T.size() > DEM_MAX_TYPES) throw DsmlError("too many particle types");

  // particle count (src/main.cu:91-139)
  uint64_t n = P.pos.size();
  for (const UVec2& r : P.num) n += (uint64_t)r.x * r.y;
  for (unsigned t : P.t_num) n += (uint64_t)t * (t + 1) / 2;
  if (n == 0) throw DsmlError("the scene has no particles");
  if (n > 0xFFFFFFFFull) throw DsmlError("too many particles");

  dem_system_props& s = out.sys;
  s.num_particles = (uint32_t)n;
  s.cube_dimension[0] = sim.environment.dimension.x;
  s.cube_dimension[1] = sim.environment.dimension.y;
  s.time_step = sim.parameters.timeStep;
  s.gravity[0] = sim.environment.gravity.x;
  s.gravity[1] = sim.environment.gravity.y;
  s.boundary_normal_stiffness = sim.environment.boundaryNormalStiffness;
  s.boundary_shear_stiffness = sim.environment.boundaryShearStiffness;
  s.flags = 0;  // the reference's wrapping grid and closed box
  // int conversion of simDuration/timeStep (src/main.cu:158)
  if (sim.parameters.simDuration != -1.0f)
    out.duration_steps = (long long)(sim.parameters.simDuration / s.time_step);

  // followed particles: single-particle follow= first, then <follow> IDs
  // (IDs that do not exist are set aside: the reference reads past its
  // arrays for them, src/functions.cu:630)
  std::vector<int> follow = P.followedParticles;
  follow.insert(follow.end(), sim.parameters.followedParticles.begin(), sim.parameters.followedParticles.end());
  for (int f : follow) {
    if (f < 0 || (uint64_t)f >= n) out.missing_followed.push_back(f);
    else out.followed.push_back(f);
  }

  // material rows (src/main.cu:180-197); inertia 2 m r^2 / 5 in float
  out.types.resize(T.size());
  for (size_t i = 0; i < T.size(); ++i) {
    dem_particle_props& p = out.types[i];
    p = dem_particle_props{};
    p.mass = T[i].mass;
    p.radius = T[i].radius;
    p.normal_stiffness = T[i].normalStiffness;
    p.shear_stiffness = T[i].shearStiffness;
    p.normal_damping = T[i].normalDamping;
    p.boundary_damping = T[i].boundaryDamping;
    p.friction_coefficient = T[i].frictionCoefficient;
    p.color_r = T[i].color[0];
    p.color_g = T[i].color[1];
    p.color_b = T[i].color[2];
    p.inertia = 2.0f * p.mass * p.radius * p.radius / 5.0f;
    p.attract_coefficient = T[i].attractCoefficient;
  }

  // largest radius, skipping the controlled particle's type (src/main.cu:199-207)
  float max_radius = 0.0f;
  for (size_t i = 0; i < T.size(); ++i) {
    if ((int)i == P.controlledType) continue;
    if (max_radius < out.types[i].radius) max_radius = out.types[i].radius;
  }
  out.max_radius = max_radius;
  out.controlled_type = P.controlledType;
  {
    dem_render_params& rp = out.render;
    rp.image_dim[1] = (int32_t)sim.parameters.imageDIMy;
    rp.image_dim[0] = (int32_t)(s.cube_dimension[0] / s.cube_dimension[1] * (float)rp.image_dim[1]);
    if (P.controlledType >= 0) {
      const float plot_radius = out.types[P.controlledType].radius;
      rp.control_dim[0] = (int32_t)(std::ceil((float)rp.image_dim[0] / s.cube_dimension[0] * plot_radius) * 2.0f);
      if (rp.control_dim[0] < 2) rp.control_dim[0] = 2;
      rp.control_dim[1] = (int32_t)(std::ceil((float)rp.image_dim[1] / s.cube_dimension[1] * plot_radius) * 2.0f);
      if (rp.control_dim[1] < 2) rp.control_dim[1] = 2;
    }
    rp.bg_color = 0;
    rp.view_rotations = 1;
    rp.color_by_pressure = 0;
    rp.max_pressure = 1.0f;
  }
  reference_grid(s.cube_dimension[0], s.cube_dimension[1], max_radius, &s.grid_size[0], &s.grid_size[1]);
  s.num_cells = s.grid_size[0] * s.grid_size[1];

  out.pos.assign(2 * n, 0.0f);
  out.vel.assign(2 * n, 0.0f);
  out.theta.assign(n, 0.0f);
  out.omega.assign(n, 0.0f);
  out.id.resize(n);
  out.loc.resize(n);
  out.type.resize(n);
  for (uint64_t i = 0; i < n; ++i) out.id[i] = out.loc[i] = (uint32_t)i;

  uint64_t at = 0;
  out.seed = seed;
  out.num_singles = P.pos.size();
  // single particles (createSingleParticles, src/functions.cu:283-311)
  for (size_t k = 0; k < P.pos.size(); ++k, ++at) {
    out.pos[2 * at] = P.pos[k].x;
    out.pos[2 * at + 1] = P.pos[k].y;
    out.vel[2 * at] = P.vel[k].x;
    out.vel[2 * at + 1] = P.vel[k].y;
    out.theta[at] = P.theta[k];
    out.omega[at] = P.omega[k];
    out.type[at] = (uint32_t)P.type[k];
  }
  // rectangles (createRectanglesD, src/particles_kernel.cuh:43-74)
  for (size_t r = 0; r < P.start.size(); ++r) {
    const Vec2 st = P.start[r];
    const Vec2 side_len{P.end[r].x - st.x, P.end[r].y - st.y};
    const UVec2 side = P.num[r];
    const auto& tv = P.types[r];
    const float dx = side_len.x / (float)(side.x - 1);
    const float dy = side_len.y / (float)(side.y - 1);
    RectBlock blk;
    blk.start[0] = st.x;
    blk.start[1] = st.y;
    blk.side_length[0] = side_len.x;
    blk.side_length[1] = side_len.y;
    blk.side[0] = side.x;
    blk.side[1] = side.y;
    for (int t : tv) blk.types.push_back((uint32_t)t);
    blk.first = at;
    blk.stream = 1000 + 2 * r;
    out.rects.push_back(blk);
    for (uint32_t y = 0; y < side.y; ++y)
      for (uint32_t x = 0; x < side.x; ++x) {
        const uint64_t p = (uint64_t)x + (uint64_t)y * side.x;
        const float nx = (float)dem_rng::jitter_normal(seed, 1000 + 2 * r, p);
        const float ny = (float)dem_rng::jitter_normal(seed, 1001 + 2 * r, p);
        const uint64_t i = at + p;
        out.pos[2 * i] = st.x + dx * ((float)x + (nx - 0.5f) / 100.0f);
        out.pos[2 * i + 1] = st.y + dy * ((float)y + (ny - 0.5f) / 100.0f);
        out.type[i] = (uint32_t)tv[(x + y) % tv.size()];
      }
    at += (uint64_t)side.x * side.y;
  }
  // triangles (createTrianglesD, src/particles_kernel.cuh:76-111; spacing
  // and row height from src/main.cu:305-307)
  for (size_t t = 0; t < P.t_pos.size(); ++t) {
    const uint32_t N = P.t_num[t];
    const uint64_t count = (uint64_t)N * (N + 1) / 2;
    const float space = P.width[t] / (float)(N - 1);
    const float height = (float)((double)space * std::sqrt(3.0) / 2.0);
    const auto& tv = P.t_types[t];
    const Vec2 st = P.t_pos[t];
    TriBlock blk;
    blk.start[0] = st.x;
    blk.start[1] = st.y;
    blk.N = N;
    blk.space = space;
    blk.height = height;
    for (int k : tv) blk.types.push_back((uint32_t)k);
    blk.first = at;
    blk.count = count;
    out.tris.push_back(blk);
    for (uint64_t index = 0; index < count; ++index) {
      const uint32_t radicand = N * N + N - 2u * (uint32_t)(index + 1) + 1u;
      const int L1 = (int)((1.0f + 2.0f * std::sqrt((float)radicand)) / 2.0f);
      const int L = (int)N - L1;
      const int sidx = (int)(((uint32_t)L1 + N + N * N - (uint32_t)(L1 * L1)) / 2u);
      const int C = (int)index - sidx + L1;
      const int A = (int)N - 1;
      const uint64_t i = at + index;
      out.pos[2 * i] = st.x + (float)(L + 2 * C - A) * space / 2.0f;
      out.pos[2 * i + 1] = st.y + (float)L * height;
      out.type[i] = (uint32_t)tv[(uint32_t)(C + L) % tv.size()];
    }
    at += count;
  }
  return out;
}

}  // namespace dsml
