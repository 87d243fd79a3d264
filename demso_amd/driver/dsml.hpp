// dsml.hpp -- the .dsml scene model and reader.
//
// The classes mirror DemSo's scene classes (src/datatypes.hpp:29-104) field
// for field, and DEMSimulation::loadFromFile reads a .dsml file with the
// rules of DEMParser (src/parser.cpp:49-377): the same tags, attributes,
// defaults and error messages.  Differences, all on inputs where the
// reference has undefined behaviour: a missing section, a <stiffness>
// without dir= in <environment>, or a malformed colour raise DsmlError
// instead of dereferencing null or overflowing a buffer, and fields the
// reference leaves uninitialised (mass, radius, stiffnesses, dampings of a
// type; the environment's stiffnesses) start at 0.
#pragma once

#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace dsml {

struct DsmlError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Vec2 {
  float x = 0.0f, y = 0.0f;
};

struct UVec2 {
  unsigned x = 0, y = 0;
};

// src/datatypes.hpp:29-35
struct DEMParameters {
  float timeStep = -1.0f;     // required
  float simDuration = -1.0f;  // -1 = run until stopped
  std::vector<int> followedParticles;
  unsigned imageDIMy = 800;
};

// src/datatypes.hpp:37-43
struct DEMEnvironment {
  Vec2 dimension;
  Vec2 gravity;
  float boundaryNormalStiffness = 0.0f;
  float boundaryShearStiffness = 0.0f;
};

// src/datatypes.hpp:45-59
struct DEMParticleType {
  std::string id;
  std::string name;
  float mass = 0.0f;
  float radius = 0.0f;
  float normalStiffness = 0.0f;
  float shearStiffness = 0.0f;
  float normalDamping = 0.0f;
  float boundaryDamping = 0.0f;
  float frictionCoefficient = 0.0f;
  float attractCoefficient = 0.0f;
  float color[3] = {255.0f, 255.0f, 255.0f};
};

// src/datatypes.hpp:61-66
struct DEMProperties {
  std::vector<DEMParticleType> particleTypes;
  void addParticleType(const DEMParticleType& t) { particleTypes.push_back(t); }
  int particleTypeIndexById(const std::string& id) const;  // DsmlError if absent
};

// src/datatypes.hpp:69-94
struct DEMParticles {
  // rectangles
  std::vector<Vec2> start;
  std::vector<Vec2> end;
  std::vector<UVec2> num;
  std::vector<std::vector<int>> types;
  // triangles
  std::vector<Vec2> t_pos;
  std::vector<float> width;
  std::vector<unsigned> t_num;
  std::vector<std::vector<int>> t_types;
  // single particles
  std::vector<Vec2> pos;
  std::vector<Vec2> vel;
  std::vector<float> theta;
  std::vector<float> omega;
  std::vector<int> type;
  std::vector<int> followedParticles;
  // <controlled> (read only when loading as the controlled-particle build)
  int controlledType = -1;
};

struct LoadOptions {
  // Read <particles> as DemSo's default build does (USE_BIG_PARTICLE 1,
  // src/main.cuh:29): a <controlled particletype=...> tag is required and its
  // type recorded.  Off: the tag is ignored (USE_BIG_PARTICLE 0).
  bool controlled_particle_build = false;
  // Messages the reference prints on stdout (types defaulting to the first).
  std::ostream* notes = nullptr;
};

// src/datatypes.hpp:96-104
struct DEMSimulation {
  DEMParameters parameters;
  DEMEnvironment environment;
  DEMProperties properties;
  DEMParticles particles;
  void loadFromFile(const std::string& filename, const LoadOptions& opt = {});
  void loadFromString(const std::string& document, const LoadOptions& opt = {});
  void printConfiguration(std::ostream& out) const;  // src/DEMSimulation.cpp:61-133
};

}  // namespace dsml
