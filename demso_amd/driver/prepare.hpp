// prepare.hpp -- from a loaded .dsml scene to step parameters and the
// initial particle state: DemSo's PrepareSim (src/main.cu:54-369).
#pragma once

#include <cstdint>
#include <vector>

#include "demso.h"
#include "dsml.hpp"

namespace dsml {

// A rectangle block as createRectangles takes it (src/functions.cu:198-245):
// its particles are first, first + 1, ...; draws from streams stream and
// stream + 1.
struct RectBlock {
  float start[2], side_length[2];
  uint32_t side[2];
  std::vector<uint32_t> types;
  uint64_t first = 0, stream = 0;
};
// A triangle block as createTriangles takes it (src/functions.cu:247-281).
struct TriBlock {
  float start[2];
  uint32_t N = 0;
  float space = 0.0f, height = 0.0f;
  std::vector<uint32_t> types;
  uint64_t first = 0, count = 0;
};

struct PreparedScene {
  dem_system_props sys{};                  // SystemProperties (src/main.cuh:108-122)
  std::vector<dem_particle_props> types;   // ParticleProperties rows (src/main.cuh:39-54)
  // initial state, particle i has ID i: single particles, then rectangles,
  // then triangles (src/main.cu:249-324)
  std::vector<float> pos, vel, theta, omega;  // pos/vel interleaved x,y
  std::vector<uint32_t> id, loc, type;
  std::vector<int> followed;  // IDs written to the output file (src/main.cu:160-165)
  std::vector<int> missing_followed;  // followed IDs beyond the particle count (not written)
  long long duration_steps = -1;  // simDuration / timeStep, -1 when absent (src/main.cu:158)
  float max_radius = 0.0f;
  // material row of the controlled particle, -1 when the scene was not
  // loaded as the controlled-particle build (src/main.cu:141-149)
  int controlled_type = -1;
  // the blocks again, for generation on the device: particles [0,
  // num_singles) are the single particles (host arrays above), then the
  // rectangles and triangles
  uint64_t num_singles = 0, seed = 0;
  // RenderParameters (src/main.cu:172-177, 209-215): image size and the
  // controlled particle's pixel square
  dem_render_params render{};
  std::vector<RectBlock> rects;
  std::vector<TriBlock> tris;
};


// Power-of-two grid of PrepareSim (src/main.cu:218-236), including its use of
// the box width for the y cell size (src/main.cu:231).  Throws DsmlError
// where the reference's arithmetic would be undefined (no positive radius,
// a box narrower than four radii).
void reference_grid(float lx, float ly, float max_radius, uint32_t* gx, uint32_t* gy);

// Rectangle jitter: the reference draws curand_normal with a time(NULL) seed
// (src/main.cu:291, src/particles_kernel.cuh:65-69), which no run can
// reproduce; here the two draws of particle p of rectangle r are
// dem_rng::jitter_normal(seed, 1000 + 2r, p) and (seed, 1001 + 2r, p)
// (csrc/scene_rng.h: the same bits as the device generator
// dem_create_rectangles_ex and demso_amd/scene.py), in the same formula.
PreparedScene prepare_sim(const DEMSimulation& sim, uint64_t seed);

}  // namespace dsml
