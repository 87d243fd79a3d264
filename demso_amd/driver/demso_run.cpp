// demso_run.cpp -- headless DemSo driver over the HIP step.
//
//   demso_run [options] scene.dsml
//
// Loads a .dsml scene with DemSo's parser rules, prepares the particles as
// PrepareSim does (src/main.cu:54-369) and runs the step loop of SimLooping
// (src/main.cu:371-536) without the OpenGL window: every step writes the
// followed particles' line to the output file in writeOutputFile's format
// (src/functions.cu:608-661), and binary state dumps replace the rendered
// frames.
//
// Two loops:
//   --loop stages  (default) the reference's loop itself: ping-pong buffers
//                  owned by this driver and the five launch wrappers with
//                  their reference signatures (include/demso_reference_shim.hpp)
//   --loop graph   the library's owned simulation, steps replayed from a
//                  captured hipGraph (dem_sim_step_graph)
// Both produce the same numbers; the rows of followed particles are gathered
// on the device every step and copied back in batches, instead of twelve
// 4-byte copies per particle per step.
//
// Options:
//   --steps N          steps to run (default: simduration / timestep)
//   --output FILE      output file of followed particles (default output.csv; "" = none)
//   --dump-every K     write a state dump every K steps (0 = none)
//   --dump-final       write a state dump after the last step
//   --dump-prefix P    dump file prefix (default "state_"); files P<step>.bin
//   --seed S           seed of the rectangles' jitter (default 0)
//   --controlled       the controlled-particle build (USE_BIG_PARTICLE): the
//                      <controlled> tag is required, its type is left out of
//                      the grid sizing, and the particle starts at (5, 9.5)
//                      and moves down 0.005 per step (src/main.cu:141-149,
//                      510-514); its final position and the last step's
//                      reaction on it are printed at the end
//   --device D         HIP device (default 0)
//   --print-config     print the loaded scene and exit (no GPU)
//   --dump-initial F   write the prepared initial state to F and exit (no GPU)
//   --image-every K    write the frame of plotParticles every K steps as a
//                      binary PPM (P6, top row first) named <prefix><step>.ppm
//   --image-prefix P   image file prefix (default "frame_")
//   --color-by-pressure  colour by pressure / the frame's largest pressure
//                      (updatePressureScale before each frame)
//   --no-rotations     no rotation lines;  --bg-color B  background byte
//
// Dump format (little endian): "DEMSODMP", u32 version 1, u32 n, u64 step,
// f64 time, then arrays in particle-ID order: pos f32[n][2], vel f32[n][2],
// theta f32[n], omega f32[n], type u32[n], acc f32[n][2], alpha f32[n],
// pressure f32[n].  The initial dump has zero acc/alpha/pressure.
#include <hip/hip_runtime_api.h>

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "demso.h"
#include "demso_reference_shim.hpp"
#include "dsml.hpp"
#include "prepare.hpp"

namespace {

// The reference's parameter structs (src/main.cuh:39-54, 108-122), as the
// shim's init() reads them.
struct Float2 {
  float x, y;
};
struct UInt2 {
  unsigned x, y;
};
struct SystemProperties {
  unsigned numParticles;
  Float2 cubeDimension;
  UInt2 gridSize;
  unsigned numCells;
  float timeStep;
  Float2 gravity;
  float boundaryNormalStiffness;
  float boundaryShearStiffness;
};
struct ParticleProperties {
  float radius, mass, inertia, normalStiffness, shearStiffness, normalDamping, boundaryDamping,
      frictionCoefficient, attractCoefficient, colorR, colorG, colorB;
};

struct Options {
  std::string scene;
  long long steps = -1;
  std::string output = "output.csv";
  long long dump_every = 0;
  bool dump_final = false;
  std::string dump_prefix = "state_";
  uint64_t seed = 0;
  bool controlled = false;
  int device = 0;
  bool print_config = false;
  std::string dump_initial;
  bool graph = false;
  long long image_every = 0;
  std::string image_prefix = "frame_";
  bool color_by_pressure = false;
  bool rotations = true;
  int bg_color = 0;
};

[[noreturn]] void usage(const char* msg) {
  std::fprintf(stderr, "demso_run: %s\nusage: demso_run [--steps N] [--output FILE] [--dump-every K] [--dump-final] "
                       "[--dump-prefix P] [--seed S] [--controlled] [--loop stages|graph] [--device D] "
                       "[--print-config] [--dump-initial FILE] [--image-every K] [--image-prefix P] [--color-by-pressure] "
                       "[--no-rotations] [--bg-color B] scene.dsml\n",
               msg);
  std::exit(2);
}

Options parse_args(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto value = [&]() -> std::string {
      if (i + 1 >= argc) usage(("missing value after " + a).c_str());
      return argv[++i];
    };
    auto integer = [&]() -> long long {
      const std::string v = value();
      char* end = nullptr;
      const long long x = std::strtoll(v.c_str(), &end, 10);
      if (v.empty() || *end || x < 0) usage(("bad number for " + a + ": " + v).c_str());
      return x;
    };
    if (a == "--steps") o.steps = integer();
    else if (a == "--output") o.output = value();
    else if (a == "--dump-every") o.dump_every = integer();
    else if (a == "--dump-final") o.dump_final = true;
    else if (a == "--dump-prefix") o.dump_prefix = value();
    else if (a == "--seed") o.seed = (uint64_t)integer();
    else if (a == "--controlled") o.controlled = true;
    else if (a == "--device") o.device = (int)integer();
    else if (a == "--print-config") o.print_config = true;
    else if (a == "--dump-initial") o.dump_initial = value();
    else if (a == "--image-every") o.image_every = integer();
    else if (a == "--image-prefix") o.image_prefix = value();
    else if (a == "--color-by-pressure") o.color_by_pressure = true;
    else if (a == "--no-rotations") o.rotations = false;
    else if (a == "--bg-color") o.bg_color = (int)integer();
    else if (a == "--loop") {
      const std::string v = value();
      if (v == "graph") o.graph = true;
      else if (v != "stages") usage("--loop takes stages or graph");
    } else if (!a.empty() && a[0] == '-') usage(("unknown option " + a).c_str());
    else if (o.scene.empty()) o.scene = a;
    else usage("more than one scene file");
  }
  if (o.scene.empty()) usage("no scene file");
  return o;
}

// host arrays in sorted order -> dump in ID order
struct HostState {
  std::vector<float> pos, vel, theta, omega, acc, alpha, pressure;
  std::vector<uint32_t> id, type;
  explicit HostState(size_t n)
      : pos(2 * n), vel(2 * n), theta(n), omega(n), acc(2 * n), alpha(n), pressure(n), id(n), type(n) {}
};

bool write_dump(const std::string& path, const HostState& h, uint64_t step, double time) {
  const size_t n = h.id.size();
  HostState o(n);
  for (size_t i = 0; i < n; ++i) {
    const uint32_t d = h.id[i];
    if (d >= n) {
      std::fprintf(stderr, "demso_run: particle ID %u out of range in the state\n", d);
      return false;
    }
    o.pos[2 * d] = h.pos[2 * i];
    o.pos[2 * d + 1] = h.pos[2 * i + 1];
    o.vel[2 * d] = h.vel[2 * i];
    o.vel[2 * d + 1] = h.vel[2 * i + 1];
    o.theta[d] = h.theta[i];
    o.omega[d] = h.omega[i];
    o.type[d] = h.type[i];
    o.acc[2 * d] = h.acc[2 * i];
    o.acc[2 * d + 1] = h.acc[2 * i + 1];
    o.alpha[d] = h.alpha[i];
    o.pressure[d] = h.pressure[i];
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "demso_run: cannot write %s\n", path.c_str());
    return false;
  }
  const uint32_t version = 1, n32 = (uint32_t)n;
  bool ok = std::fwrite("DEMSODMP", 1, 8, f) == 8;
  ok = ok && std::fwrite(&version, 4, 1, f) == 1 && std::fwrite(&n32, 4, 1, f) == 1;
  ok = ok && std::fwrite(&step, 8, 1, f) == 1 && std::fwrite(&time, 8, 1, f) == 1;
  ok = ok && std::fwrite(o.pos.data(), 4, 2 * n, f) == 2 * n && std::fwrite(o.vel.data(), 4, 2 * n, f) == 2 * n;
  ok = ok && std::fwrite(o.theta.data(), 4, n, f) == n && std::fwrite(o.omega.data(), 4, n, f) == n;
  ok = ok && std::fwrite(o.type.data(), 4, n, f) == n && std::fwrite(o.acc.data(), 4, 2 * n, f) == 2 * n;
  ok = ok && std::fwrite(o.alpha.data(), 4, n, f) == n && std::fwrite(o.pressure.data(), 4, n, f) == n;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) std::fprintf(stderr, "demso_run: short write to %s\n", path.c_str());
  return ok;
}

float as_float(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// writeOutputFile's line (src/functions.cu:622-660) from gathered rows
void write_line(FILE* out, float elapsed, const uint32_t* rows, size_t count) {
  std::fprintf(out, "%f,", elapsed);
  for (size_t f = 0; f < count; ++f) {
    const uint32_t* r = rows + DEM_ROW_WORDS * f;
    std::fprintf(out, "%u,%u,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f", r[0], r[1], as_float(r[2]), as_float(r[3]),
                 as_float(r[4]), as_float(r[5]), as_float(r[6]), as_float(r[7]), as_float(r[8]), as_float(r[9]),
                 as_float(r[10]), as_float(r[11]));
    if (f + 1 != count) std::fprintf(out, ",");
  }
  std::fprintf(out, "\n");
}

// Final state of the controlled particle; %a keeps every bit.
void print_controlled(float x, float y, float fx, float fy, float m) {
  std::printf("controlled particle: pos %a %a reaction %a %a %a\n", (double)x, (double)y, (double)fx, (double)fy,
              (double)m);
}

#define HIP_OK(call)                                                                                  \
  do {                                                                                                \
    const hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                           \
      std::fprintf(stderr, "demso_run: %s failed: %s\n", #call, hipGetErrorString(e_));               \
      return 1;                                                                                       \
    }                                                                                                 \
  } while (0)

#define DEM_OK_OR_FAIL(call, ctx)                                                                     \
  do {                                                                                                \
    const int r_ = (call);                                                                            \
    if (r_ != DEM_OK) {                                                                               \
      std::fprintf(stderr, "demso_run: %s failed (%d): %s\n", #call, r_, dem_last_error(ctx));        \
      return 1;                                                                                       \
    }                                                                                                 \
  } while (0)

// Device buffers of the reference's ParticlesValues (src/main.cuh:58-104)
// used by the step loop, allocated as allocateVectors does
// (src/functions.cu:68-149).
struct Buffers {
  std::vector<void*> all;
  float *pos[2], *vel[2], *theta[2], *omega[2], *acc, *alpha, *pressure;
  uint32_t *id[2], *loc[2], *type[2], *hash, *index, *cellStart, *cellEnd;
  ~Buffers() {
    for (void* p : all) (void)hipFree(p);
  }
  template <class T>
  hipError_t get(T** p, size_t count) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, count * sizeof(T) + 4);
    if (e == hipSuccess) all.push_back(q);
    *p = (T*)q;
    return e;
  }
};

// PrepareSim's particle creation on the device (src/main.cu:249-324): the
// single particles, then each rectangle and triangle with the library's
// generators, into device arrays of n particles; vel, theta and omega of the
// generated blocks start at zero.
int create_on_device(dem_context* ctx, const dsml::PreparedScene& ps, float* pos, float* vel, float* theta,
                     float* omega, uint32_t* id, uint32_t* loc, uint32_t* type) {
  const size_t n = ps.sys.num_particles;
  HIP_OK(hipMemset(vel, 0, 8 * n));
  HIP_OK(hipMemset(theta, 0, 4 * n));
  HIP_OK(hipMemset(omega, 0, 4 * n));
  const uint32_t ns = (uint32_t)ps.num_singles;
  DEM_OK_OR_FAIL(dem_create_single_particles(ctx, pos, vel, theta, omega, id, loc, type, ps.pos.data(), ps.vel.data(),
                                             ps.theta.data(), ps.omega.data(), ps.type.data(), ns, 0),
                 ctx);
  for (const dsml::RectBlock& r : ps.rects) {
    const size_t f = r.first;
    DEM_OK_OR_FAIL(dem_create_rectangles_ex(ctx, pos + 2 * f, id + f, loc + f, type + f, r.start, r.side_length,
                                            r.side, (uint32_t)f, (uint32_t)r.types.size(), r.types.data(), ps.seed,
                                            r.stream),
                   ctx);
  }
  for (const dsml::TriBlock& t : ps.tris) {
    const size_t f = t.first;
    DEM_OK_OR_FAIL(dem_create_triangles(ctx, pos + 2 * f, id + f, loc + f, type + f, t.start, t.N,
                                        (uint32_t)t.types.size(), t.types.data(), t.space, t.height, (uint32_t)f,
                                        (uint32_t)t.count),
                   ctx);
  }
  return 0;
}

// One frame: the image of plotParticles (device RGBA, row 0 at the bottom)
// as a P6 PPM with the top row first.
struct Frames {
  uint8_t* d_img = nullptr;
  std::vector<uint8_t> h_img, rgb;
  ~Frames() {
    if (d_img) (void)hipFree(d_img);
  }
  bool write(const std::string& path, const dem_render_params& rp) {
    const size_t w = (size_t)rp.image_dim[0], h = (size_t)rp.image_dim[1];
    h_img.resize(4 * w * h);
    if (hipMemcpy(h_img.data(), d_img, h_img.size(), hipMemcpyDeviceToHost) != hipSuccess) return false;
    rgb.resize(3 * w * h);
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x)
        for (int c = 0; c < 3; ++c) rgb[3 * ((h - 1 - y) * w + x) + c] = h_img[4 * (y * w + x) + c];
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
      std::fprintf(stderr, "demso_run: cannot write %s\n", path.c_str());
      return false;
    }
    std::fprintf(f, "P6\n%zu %zu\n255\n", w, h);
    const bool ok = std::fwrite(rgb.data(), 1, rgb.size(), f) == rgb.size();
    return (std::fclose(f) == 0) && ok;
  }
};

struct Run {
  const Options& opt;
  const dsml::PreparedScene& ps;
  uint32_t n;
  FILE* out = nullptr;
  uint32_t* d_follow = nullptr;
  uint32_t* d_rows = nullptr;
  std::vector<uint32_t> h_rows;
  size_t batch = 0, pending = 0;
  long long first_pending_step = 0;

  Run(const Options& o, const dsml::PreparedScene& p, uint32_t count) : opt(o), ps(p), n(count) {}

  int flush_rows(hipStream_t s) {
    if (!out || pending == 0) return 0;
    const size_t count = ps.followed.size();
    HIP_OK(hipMemcpyAsync(h_rows.data(), d_rows, pending * count * DEM_ROW_WORDS * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (size_t b = 0; b < pending; ++b) {
      // elapsed time as the reference computes it: timeStep * tempo in float
      const float elapsed = ps.sys.time_step * (float)(int)(first_pending_step + (long long)b);
      write_line(out, elapsed, h_rows.data() + b * count * DEM_ROW_WORDS, count);
    }
    pending = 0;
    return 0;
  }
};

int run_sim(const Options& opt, const dsml::PreparedScene& ps, long long steps) {
  const uint32_t n = ps.sys.num_particles;
  HIP_OK(hipSetDevice(opt.device));
  Run run(opt, ps, n);
  const size_t count = ps.followed.size();
  if (!opt.output.empty()) {
    run.out = std::fopen(opt.output.c_str(), "w");
    if (!run.out) {
      std::fprintf(stderr, "demso_run: cannot open %s\n", opt.output.c_str());
      return 1;
    }
  }
  struct Closer {
    Run& r;
    ~Closer() {
      if (r.out) std::fclose(r.out);
      if (r.d_follow) (void)hipFree(r.d_follow);
      if (r.d_rows) (void)hipFree(r.d_rows);
    }
  } closer{run};
  if (run.out && count) {
    run.batch = 256;
    std::vector<uint32_t> f(ps.followed.begin(), ps.followed.end());
    HIP_OK(hipMalloc((void**)&run.d_follow, count * 4));
    HIP_OK(hipMemcpy(run.d_follow, f.data(), count * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMalloc((void**)&run.d_rows, run.batch * count * DEM_ROW_WORDS * 4));
    run.h_rows.resize(run.batch * count * DEM_ROW_WORDS);
  }

  HostState host(n);
  Frames frames;
  dem_render_params rp = ps.render;
  rp.view_rotations = opt.rotations ? 1u : 0u;
  rp.color_by_pressure = opt.color_by_pressure ? 1u : 0u;
  rp.bg_color = opt.bg_color;
  if (opt.image_every > 0)
    HIP_OK(hipMalloc((void**)&frames.d_img, 4 * (size_t)rp.image_dim[0] * (size_t)rp.image_dim[1]));
  auto want_image = [&](long long step) { return opt.image_every > 0 && step % opt.image_every == 0; };
  auto image_name = [&](long long step) { return opt.image_prefix + std::to_string(step) + ".ppm"; };
  auto dump_name = [&](long long step) { return opt.dump_prefix + std::to_string(step) + ".bin"; };
  auto want_dump = [&](long long step) {
    return (opt.dump_every > 0 && step % opt.dump_every == 0) || (opt.dump_final && step == steps);
  };

  if (!opt.graph) {
    // ---- the reference's loop through the shim --------------------------
    SystemProperties sis{};
    sis.numParticles = n;
    sis.cubeDimension = {ps.sys.cube_dimension[0], ps.sys.cube_dimension[1]};
    sis.gridSize = {ps.sys.grid_size[0], ps.sys.grid_size[1]};
    sis.numCells = ps.sys.num_cells;
    sis.timeStep = ps.sys.time_step;
    sis.gravity = {ps.sys.gravity[0], ps.sys.gravity[1]};
    sis.boundaryNormalStiffness = ps.sys.boundary_normal_stiffness;
    sis.boundaryShearStiffness = ps.sys.boundary_shear_stiffness;
    std::vector<ParticleProperties> part(ps.types.size());
    for (size_t t = 0; t < ps.types.size(); ++t) {
      const dem_particle_props& p = ps.types[t];
      part[t] = ParticleProperties{p.radius,          p.mass,           p.inertia,
                                   p.normal_stiffness, p.shear_stiffness, p.normal_damping,
                                   p.boundary_damping, p.friction_coefficient, p.attract_coefficient,
                                   p.color_r,          p.color_g,        p.color_b};
    }
    if (demso_shim::init(sis, part.data(), (unsigned)part.size(), opt.device, nullptr, ps.sys.flags) != DEM_OK)
      return 1;
    struct ShimDown {
      ~ShimDown() { demso_shim::shutdown(); }
    } shim_down;
    dem_context* ctx = demso_shim::context();
    Buffers b;
    for (int k = 0; k < 2; ++k) {
      HIP_OK(b.get(&b.pos[k], 2 * (size_t)n));
      HIP_OK(b.get(&b.vel[k], 2 * (size_t)n));
      HIP_OK(b.get(&b.theta[k], n));
      HIP_OK(b.get(&b.omega[k], n));
      HIP_OK(b.get(&b.id[k], n));
      HIP_OK(b.get(&b.loc[k], n));
      HIP_OK(b.get(&b.type[k], n));
    }
    HIP_OK(b.get(&b.acc, 2 * (size_t)n));
    HIP_OK(b.get(&b.alpha, n));
    HIP_OK(b.get(&b.pressure, n));
    HIP_OK(b.get(&b.hash, n));
    HIP_OK(b.get(&b.index, n));
    HIP_OK(b.get(&b.cellStart, ps.sys.num_cells));
    HIP_OK(b.get(&b.cellEnd, ps.sys.num_cells));
    // the controlled particle's reaction vectors and block sums
    // (src/functions.cu:104-115)
    const bool control = ps.controlled_type >= 0;
    const uint32_t react_blocks = (n + 511u) / 512u;
    float *cfx = nullptr, *cfy = nullptr, *cm = nullptr, *hx = nullptr, *hy = nullptr, *hm = nullptr;
    if (control) {
      HIP_OK(b.get(&cfx, n));
      HIP_OK(b.get(&cfy, n));
      HIP_OK(b.get(&cm, n));
      HIP_OK(b.get(&hx, react_blocks));
      HIP_OK(b.get(&hy, react_blocks));
      HIP_OK(b.get(&hm, react_blocks));
    }
    Float2 ctrl_pos{5.0f, 9.5f}, ctrl_vel{0.0f, 0.0f}, ctrl_f{0.0f, 0.0f};
    float ctrl_m = 0.0f;
    HIP_OK(hipMemset(b.cellEnd, 0, ps.sys.num_cells * 4));  // zeroed once (src/functions.cu:140)
    HIP_OK(hipMemset(b.acc, 0, 8 * (size_t)n));
    HIP_OK(hipMemset(b.alpha, 0, 4 * (size_t)n));
    HIP_OK(hipMemset(b.pressure, 0, 4 * (size_t)n));
    if (create_on_device(ctx, ps, b.pos[0], b.vel[0], b.theta[0], b.omega[0], b.id[0], b.loc[0], b.type[0]))
      return 1;
    hipStream_t s = nullptr;  // the shim's context launches on the null stream
    int cur = 0;
    if (opt.dump_final && steps == 0) {
      // the created state itself (acc, alpha, pressure are zero)
      HIP_OK(hipDeviceSynchronize());
      HIP_OK(hipMemcpy(host.pos.data(), b.pos[0], 8 * (size_t)n, hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(host.vel.data(), b.vel[0], 8 * (size_t)n, hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(host.theta.data(), b.theta[0], 4 * (size_t)n, hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(host.omega.data(), b.omega[0], 4 * (size_t)n, hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(host.id.data(), b.id[0], 4 * (size_t)n, hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(host.type.data(), b.type[0], 4 * (size_t)n, hipMemcpyDeviceToHost));
      if (!write_dump(dump_name(0), host, 0, 0.0)) return 1;
    }
    for (long long tempo = 1; tempo <= steps; ++tempo) {
      const int o = cur, t = cur ^ 1;
      calcHash(b.pos[o], b.index, b.hash, n);
      sortParticles(b.hash, b.index, n);
      reorderDataAndFindCellStart(b.cellStart, b.cellEnd, b.pos[t], b.vel[t], b.theta[t], b.omega[t], b.id[t],
                                  b.loc[t], b.type[t], b.hash, b.index, b.pos[o], b.vel[o], b.theta[o], b.omega[o],
                                  b.id[o], b.type[o], n, ps.sys.num_cells);
      if (control)
        collide(b.pos[t], b.vel[t], b.acc, b.omega[t], b.alpha, b.type[t], b.cellStart, b.cellEnd, n,
                ps.sys.num_cells, ctrl_pos, ctrl_vel, 0.0f, 0.0f, (unsigned)ps.controlled_type, cfx, hx, cfy, hy, cm,
                hm, &ctrl_f, &ctrl_m, b.pressure);
      else
        collide(b.pos[t], b.vel[t], b.acc, b.omega[t], b.alpha, b.type[t], b.cellStart, b.cellEnd, n,
                ps.sys.num_cells, b.pressure);
      integrateSystem(b.pos[t], b.vel[t], b.acc, b.theta[t], b.omega[t], b.alpha, b.type[t], b.pressure, n);
      if (demso_shim::last_status() != DEM_OK) return 1;
      if (control) {
        // src/main.cu:510-514
        ctrl_pos.y += -.005;
        if (ctrl_pos.y < -1) ctrl_pos.y = 10.5;
        if (ctrl_pos.x > sis.cubeDimension.x + 0.5) ctrl_pos.x = -0.5;
      }
      if (want_image(tempo)) {
        // plotParticles after the step (src/main.cu:538-553)
        if (rp.color_by_pressure) DEM_OK_OR_FAIL(dem_update_pressure_scale(ctx, b.pressure, n, &rp.max_pressure), ctx);
        dem_controlled c{};
        c.pos[0] = ctrl_pos.x;
        c.pos[1] = ctrl_pos.y;
        c.type = control ? (uint32_t)ps.controlled_type : 0u;
        c.enabled = control ? 1u : 0u;
        DEM_OK_OR_FAIL(dem_plot_particles(ctx, frames.d_img, b.pos[t], b.theta[t], b.type[t], nullptr, n, &rp,
                                          control ? &c : nullptr, b.pressure),
                       ctx);
        if (!frames.write(image_name(tempo), rp)) return 1;
      }
      cur = t;
      if (run.d_rows) {
        if (run.pending == 0) run.first_pending_step = tempo;
        DEM_OK_OR_FAIL(dem_gather_rows(ctx, run.d_follow, (uint32_t)count, b.loc[t], b.pos[t], b.vel[t], b.acc,
                                       b.theta[t], b.omega[t], b.alpha, b.id[t], b.type[t], b.pressure, n,
                                       run.d_rows + run.pending * count * DEM_ROW_WORDS),
                       ctx);
        if (++run.pending == run.batch && run.flush_rows(s)) return 1;
      }
      if (want_dump(tempo)) {
        if (run.flush_rows(s)) return 1;
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipMemcpy(host.pos.data(), b.pos[t], 8 * (size_t)n, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(host.vel.data(), b.vel[t], 8 * (size_t)n, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(host.theta.data(), b.theta[t], 4 * (size_t)n, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(host.omega.data(), b.omega[t], 4 * (size_t)n, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(host.id.data(), b.id[t], 4 * (size_t)n, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(host.type.data(), b.type[t], 4 * (size_t)n, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(host.acc.data(), b.acc, 8 * (size_t)n, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(host.alpha.data(), b.alpha, 4 * (size_t)n, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(host.pressure.data(), b.pressure, 4 * (size_t)n, hipMemcpyDeviceToHost));
        if (!write_dump(dump_name(tempo), host, (uint64_t)tempo, (double)(ps.sys.time_step * (float)tempo)))
          return 1;
      }
    }
    if (run.flush_rows(s)) return 1;
    HIP_OK(hipDeviceSynchronize());
    if (control) print_controlled(ctrl_pos.x, ctrl_pos.y, ctrl_f.x, ctrl_f.y, ctrl_m);
    return 0;
  }

  // ---- the owned simulation with graph replays ----------------------------
  dem_context* ctx = nullptr;
  if (dem_create(&ctx, opt.device, nullptr) != DEM_OK) {
    std::fprintf(stderr, "demso_run: dem_create failed: %s\n", dem_last_error(nullptr));
    return 1;
  }
  struct CtxDown {
    dem_context* c;
    dem_sim* sim = nullptr;
    ~CtxDown() {
      if (sim) dem_sim_destroy(sim);
      dem_destroy(c);
    }
  } down{ctx};
  DEM_OK_OR_FAIL(dem_set_params(ctx, &ps.sys, ps.types.data(), (uint32_t)ps.types.size()), ctx);
  DEM_OK_OR_FAIL(dem_sim_create(ctx, n, ps.sys.num_cells, 0, &down.sim), ctx);
  dem_sim* sim = down.sim;
  {
    // create on the device, then hand the arrays to the sim (device to device)
    Buffers g;
    float *pos = nullptr, *vel = nullptr, *theta = nullptr, *omega = nullptr;
    uint32_t *id = nullptr, *loc = nullptr, *type = nullptr;
    HIP_OK(g.get(&pos, 2 * (size_t)n));
    HIP_OK(g.get(&vel, 2 * (size_t)n));
    HIP_OK(g.get(&theta, n));
    HIP_OK(g.get(&omega, n));
    HIP_OK(g.get(&id, n));
    HIP_OK(g.get(&loc, n));
    HIP_OK(g.get(&type, n));
    if (create_on_device(ctx, ps, pos, vel, theta, omega, id, loc, type)) return 1;
    DEM_OK_OR_FAIL(dem_sim_upload(sim, pos, vel, theta, omega, id, type, loc, nullptr), ctx);
  }
  if (ps.controlled_type >= 0) {
    dem_controlled c{};
    c.pos[0] = 5.0f;
    c.pos[1] = 9.5f;
    c.type = (uint32_t)ps.controlled_type;
    c.enabled = 1;
    c.move[1] = -.005;
    c.wrap_below_y = -1.0;
    c.wrap_above_x = (double)ps.sys.cube_dimension[0] + 0.5;
    c.reset_y = 10.5f;
    c.reset_x = -0.5f;
    DEM_OK_OR_FAIL(dem_set_controlled(ctx, &c), ctx);
  }
  hipStream_t s = nullptr;
  long long done = 0;
  if (opt.dump_final && steps == 0) {
    DEM_OK_OR_FAIL(dem_sim_download(sim, host.pos.data(), host.vel.data(), host.theta.data(), host.omega.data(),
                                    host.id.data(), host.type.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, nullptr),
                   ctx);
    if (!write_dump(dump_name(0), host, 0, 0.0)) return 1;
  }
  while (done < steps) {
    // run up to the next step that needs output: every step with followed
    // particles, else the next dump
    long long next = steps;
    if (run.d_rows) next = done + 1;
    else if (opt.dump_every > 0) next = std::min(steps, (done / opt.dump_every + 1) * opt.dump_every);
    if (opt.image_every > 0) next = std::min(next, (done / opt.image_every + 1) * opt.image_every);
    DEM_OK_OR_FAIL(dem_sim_step_graph(sim, (uint32_t)(next - done)), ctx);
    done = next;
    if (run.d_rows) {
      if (run.pending == 0) run.first_pending_step = done;
      DEM_OK_OR_FAIL(dem_sim_gather_rows(sim, run.d_follow, (uint32_t)count,
                                         run.d_rows + run.pending * count * DEM_ROW_WORDS),
                     ctx);
      if (++run.pending == run.batch && run.flush_rows(s)) return 1;
    }
    if (want_image(done)) {
      if (rp.color_by_pressure) DEM_OK_OR_FAIL(dem_sim_pressure_scale(sim, &rp.max_pressure), ctx);
      DEM_OK_OR_FAIL(dem_sim_plot(sim, frames.d_img, &rp), ctx);
      if (!frames.write(image_name(done), rp)) return 1;
    }
    if (want_dump(done)) {
      if (run.flush_rows(s)) return 1;
      DEM_OK_OR_FAIL(dem_sim_download(sim, host.pos.data(), host.vel.data(), host.theta.data(), host.omega.data(),
                                      host.id.data(), host.type.data(), nullptr, nullptr, host.acc.data(),
                                      host.alpha.data(), host.pressure.data(), nullptr, nullptr, nullptr, nullptr),
                     ctx);
      if (!write_dump(dump_name(done), host, (uint64_t)done, (double)(ps.sys.time_step * (float)done))) return 1;
    }
  }
  if (run.flush_rows(s)) return 1;
  DEM_OK_OR_FAIL(dem_synchronize(ctx), ctx);
  if (ps.controlled_type >= 0) {
    dem_controlled c{};
    float r[3];
    DEM_OK_OR_FAIL(dem_get_controlled(ctx, &c, r), ctx);
    print_controlled(c.pos[0], c.pos[1], r[0], r[1], r[2]);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const Options opt = parse_args(argc, argv);
  dsml::DEMSimulation sim;
  dsml::LoadOptions lo;
  lo.controlled_particle_build = opt.controlled;
  lo.notes = &std::cout;
  try {
    sim.loadFromFile(opt.scene, lo);
  } catch (const dsml::DsmlError& e) {
    std::fprintf(stderr, "Parsing Error: %s\n", e.what());
    return 1;
  }
  if (opt.print_config) {
    sim.printConfiguration(std::cout);
    return 0;
  }
  dsml::PreparedScene ps;
  try {
    ps = dsml::prepare_sim(sim, opt.seed);
  } catch (const dsml::DsmlError& e) {
    std::fprintf(stderr, "Scene Error: %s\n", e.what());
    return 1;
  }
  std::printf("\nNumero de Particulas = %u\ngrid %u x %u\n\n", ps.sys.num_particles, ps.sys.grid_size[0],
              ps.sys.grid_size[1]);
  for (int f : ps.missing_followed)
    std::fprintf(stderr, "warning: followed particle %d does not exist (%u particles); it is not written\n", f,
                 ps.sys.num_particles);
  if (!opt.dump_initial.empty()) {
    HostState h(ps.sys.num_particles);
    h.pos = ps.pos;
    h.vel = ps.vel;
    h.theta = ps.theta;
    h.omega = ps.omega;
    h.id = ps.id;
    h.type = ps.type;
    return write_dump(opt.dump_initial, h, 0, 0.0) ? 0 : 1;
  }
  long long steps = opt.steps >= 0 ? opt.steps : ps.duration_steps;
  if (steps < 0) {
    std::fprintf(stderr, "demso_run: the scene has no <simduration>; pass --steps\n");
    return 2;
  }
  if (steps > 0x7FFFFFFFll) {
    std::fprintf(stderr, "demso_run: too many steps\n");
    return 2;
  }
  return run_sim(opt, ps, steps);
}
