// prepare.hpp -- from a loaded .dsml scene to step parameters and the
// initial particle state: DemSo's PrepareSim (src/main.cu:54-369).
#pragma once

#include <cstdint>
#include <vector>

#include "demso.h"
#include "dsml.hpp"

namespace dsml {

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
};

// SplitMix64 counter generator (Steele, Lea & Flood 2014): value i of stream
// s under seed k is mix(k*G + s*H + (i+1)*G); the same generator as the
// Python scene module, so both produce identical draws.
double uniform01(uint64_t seed, uint64_t stream, uint64_t i);
// Box-Muller normal from streams (s, s + 1000003).
double normal01(uint64_t seed, uint64_t stream, uint64_t i);

// Power-of-two grid of PrepareSim (src/main.cu:218-236), including its use of
// the box width for the y cell size (src/main.cu:231).  Throws DsmlError
// where the reference's arithmetic would be undefined (no positive radius,
// a box narrower than four radii).
void reference_grid(float lx, float ly, float max_radius, uint32_t* gx, uint32_t* gy);

// Rectangle jitter: the reference draws curand_normal with a time(NULL) seed
// (src/main.cu:291, src/particles_kernel.cuh:65-69), which no run can
// reproduce; here the two draws of particle p of rectangle r are
// normal01(seed, 1000 + 2r, p) and normal01(seed, 1001 + 2r, p), in the same
// formula.
PreparedScene prepare_sim(const DEMSimulation& sim, uint64_t seed);

}  // namespace dsml
