// demso_reference_shim.hpp -- DemSo's launch-wrapper signatures over the C ABI.
//
// Header-only C++ shim for the DemSo code base: the functions below have the
// exact signatures of src/functions.cuh:50-72 (calcHash, sortParticles,
// reorderDataAndFindCellStart, integrateSystem, collide in both builds:
// without the controlled particle and with it, USE_BIG_PARTICLE 1 /
// USE_ATOMIC 0) and set_gravity (src/functions.cuh:83), and
// forward to include/demso.h.  A maintainer compiles this in place of the
// kernel launches of src/functions.cu.  The reference's hidden __constant__
// state becomes one process-wide dem_context, created by demso_shim::init()
// from the reference's own SystemProperties / ParticleProperties structs
// (src/main.cuh:39-54, 108-122); the struct types are template parameters so
// this header does not depend on the reference headers.
//
// The reference's wrappers return void; here a failure is printed once to
// stderr and kept in demso_shim::last_status() (never exit()).
#pragma once

#include <cstdio>

#include "demso.h"

#ifndef DEMSO_SHIM_MAX_TYPES
#define DEMSO_SHIM_MAX_TYPES 10 /* MAX_PARTICLES_TYPES, src/main.cuh:33 */
#endif

namespace demso_shim {

inline dem_context*& context() {
  static dem_context* ctx = nullptr;
  return ctx;
}

inline int& last_status() {
  static int status = DEM_OK;
  return status;
}

inline void report(int status, const char* what) {
  last_status() = status;
  if (status != DEM_OK) std::fprintf(stderr, "demso: %s failed (%d): %s\n", what, status, dem_last_error(context()));
}

// Equivalent of the constant-memory upload in allocateVectors
// (src/functions.cu:145-147).  SP/PP are the reference's SystemProperties and
// ParticleProperties; flags may add DEM_GRID_CLAMP / DEM_OPEN_TOP.
template <class SP, class PP>
int init(const SP& sis, const PP* part, unsigned num_types, int device = 0, void* stream = nullptr,
         unsigned flags = 0) {
  if (!context()) {
    const int rc = dem_create(&context(), device, stream);
    if (rc != DEM_OK) {
      report(rc, "dem_create");
      return rc;
    }
  }
  dem_system_props s = {};
  s.num_particles = sis.numParticles;
  s.cube_dimension[0] = sis.cubeDimension.x;
  s.cube_dimension[1] = sis.cubeDimension.y;
  s.grid_size[0] = sis.gridSize.x;
  s.grid_size[1] = sis.gridSize.y;
  s.num_cells = sis.numCells;
  s.time_step = sis.timeStep;
  s.gravity[0] = sis.gravity.x;
  s.gravity[1] = sis.gravity.y;
  s.boundary_normal_stiffness = sis.boundaryNormalStiffness;
  s.boundary_shear_stiffness = sis.boundaryShearStiffness;
  s.flags = flags;
  dem_particle_props p[DEMSO_SHIM_MAX_TYPES];
  if (num_types > DEMSO_SHIM_MAX_TYPES) return DEM_ERR_INVALID;
  for (unsigned t = 0; t < num_types; ++t) {
    p[t] = dem_particle_props{};
    p[t].radius = part[t].radius;
    p[t].mass = part[t].mass;
    p[t].inertia = part[t].inertia;
    p[t].normal_stiffness = part[t].normalStiffness;
    p[t].shear_stiffness = part[t].shearStiffness;
    p[t].normal_damping = part[t].normalDamping;
    p[t].boundary_damping = part[t].boundaryDamping;
    p[t].friction_coefficient = part[t].frictionCoefficient;
    p[t].attract_coefficient = part[t].attractCoefficient;
    p[t].color_r = part[t].colorR;
    p[t].color_g = part[t].colorG;
    p[t].color_b = part[t].colorB;
  }
  int rc = dem_set_params(context(), &s, p, num_types);
  if (rc == DEM_OK) rc = dem_reserve(context(), sis.numParticles);
  report(rc, "dem_set_params");
  return rc;
}

inline void shutdown() {
  if (context()) dem_destroy(context());
  context() = nullptr;
}

}  // namespace demso_shim

// ---- src/functions.cuh:50-72 ------------------------------------------------

inline void calcHash(float* pos, unsigned* gridParticleIndex, unsigned* gridParticleHash, unsigned numParticles) {
  demso_shim::report(dem_calc_hash(demso_shim::context(), pos, gridParticleIndex, gridParticleHash, numParticles),
                     "calcHash");
}

inline void sortParticles(unsigned* dGridParticleHash, unsigned* dGridParticleIndex, unsigned numParticles) {
  demso_shim::report(dem_sort_particles(demso_shim::context(), dGridParticleHash, dGridParticleIndex, numParticles),
                     "sortParticles");
}

inline void reorderDataAndFindCellStart(unsigned* cellStart, unsigned* cellEnd, float* sortedPos, float* sortedVel,
                                        float* sortedTheta, float* sortedOmega, unsigned* sortedID,
                                        unsigned* sortedLoc, unsigned* sortedType, unsigned* gridParticleHash,
                                        unsigned* gridParticleIndex, float* oldPos, float* oldVel, float* oldTheta,
                                        float* oldOmega, unsigned* oldID, unsigned* oldType, unsigned numParticles,
                                        unsigned numCells) {
  demso_shim::report(dem_reorder_data_and_find_cell_start(
                         demso_shim::context(), cellStart, cellEnd, sortedPos, sortedVel, sortedTheta, sortedOmega,
                         sortedID, sortedLoc, sortedType, gridParticleHash, gridParticleIndex, oldPos, oldVel,
                         oldTheta, oldOmega, oldID, oldType, numParticles, numCells),
                     "reorderDataAndFindCellStart");
}

// collide as compiled with USE_BIG_PARTICLE 0 (src/functions.cuh:59-69).
inline void collide(float* oldPos, float* oldVel, float* newAcc, float* oldOmega, float* newAlpha, unsigned* oldType,
                    unsigned* cellStart, unsigned* cellEnd, unsigned numParticles, unsigned numCells,
                    float* pressure) {
  demso_shim::report(dem_collide(demso_shim::context(), oldPos, oldVel, newAcc, oldOmega, newAlpha, oldType, cellStart,
                                 cellEnd, numParticles, numCells, pressure),
                     "collide");
}

// collide as compiled with USE_BIG_PARTICLE 1, USE_ATOMIC 0
// (src/functions.cu:398-521): the pair with the controlled particle, its
// reaction per particle in controlForceVecX/Y and controlMomentVec, their
// block sums in hCFVx/hCFVy/hCMV (device, ceil(n/512) each) and the totals in
// *ctrlF, *ctrlM (host).  F2 is the reference's float2.
template <class F2>
inline void collide(float* oldPos, float* oldVel, float* newAcc, float* oldOmega, float* newAlpha, unsigned* oldType,
                    unsigned* cellStart, unsigned* cellEnd, unsigned numParticles, unsigned numCells, F2 controlPos,
                    F2 controlVel, float controlTheta, float controlOmega, unsigned controlType,
                    float* controlForceVecX, float* hCFVx, float* controlForceVecY, float* hCFVy,
                    float* controlMomentVec, float* hCMV, F2* ctrlF, float* ctrlM, float* pressure) {
  const float state[6] = {controlPos.x, controlPos.y, controlVel.x, controlVel.y, controlTheta, controlOmega};
  float f[2] = {0.0f, 0.0f};
  demso_shim::report(dem_collide_controlled(demso_shim::context(), oldPos, oldVel, newAcc, oldOmega, newAlpha, oldType,
                                            cellStart, cellEnd, numParticles, numCells, state, controlType,
                                            controlForceVecX, hCFVx, controlForceVecY, hCFVy, controlMomentVec, hCMV,
                                            f, ctrlM, pressure),
                     "collide");
  ctrlF->x = f[0];
  ctrlF->y = f[1];
}

inline void integrateSystem(float* pos, float* vel, float* acc, float* theta, float* omega, float* alpha,
                            unsigned* type, float* pressure, unsigned numParticles) {
  demso_shim::report(dem_integrate_system(demso_shim::context(), pos, vel, acc, theta, omega, alpha, type, pressure,
                                          numParticles),
                     "integrateSystem");
}

// set_gravity (src/functions.cu:663-668): SP is the reference's SystemProperties.
template <class SP, class V2>
inline void set_gravity(SP* sisProps, V2 gravity) {
  sisProps->gravity = gravity;
  demso_shim::report(dem_set_gravity(demso_shim::context(), gravity.x, gravity.y), "set_gravity");
}
