// dsml.cpp -- DEMParser's reading rules (src/parser.cpp) over xml.hpp.
#include "dsml.hpp"

#include <cstdlib>
#include <memory>

#include "xml.hpp"

namespace dsml {

int DEMProperties::particleTypeIndexById(const std::string& id) const {
  // src/DEMSimulation.cpp:37-42
  for (size_t i = 0; i < particleTypes.size(); ++i)
    if (particleTypes[i].id == id) return (int)i;
  throw DsmlError("Nonexistant particle id: '" + id + "'");
}

namespace {

// atof/atoi of the node value, as the reference reads every number
float num_f(const XmlNode& n) { return (float)std::atof(n.text.c_str()); }
int num_i(const XmlNode& n) { return std::atoi(n.text.c_str()); }

// DEMParser::readVector (src/parser.cpp:352-361): <x>, <y> children, 0 if absent
Vec2 read_vector(const XmlNode& root) {
  Vec2 v;
  for (const auto& c : root.children) {
    if (c->name == "x") v.x = num_f(*c);
    else if (c->name == "y") v.y = num_f(*c);
  }
  return v;
}

// DEMParser::readCSVLine (src/parser.cpp:364-377): split on ',', keeping empty
// fields (a trailing comma yields an empty last id); no whitespace trimming
std::vector<std::string> read_csv_line(const std::string& line) {
  std::vector<std::string> out;
  size_t b = 0;
  for (;;) {
    const size_t c = line.find(',', b);
    if (c == std::string::npos) {
      out.push_back(line.substr(b));
      return out;
    }
    out.push_back(line.substr(b, c - b));
    b = c + 1;
    if (b == line.size()) {
      out.emplace_back();
      return out;
    }
  }
}

const XmlNode& section(const XmlNode& root, const char* name) {
  const XmlNode* s = root.child(name);
  if (!s) throw DsmlError(std::string("<") + name + "> section missing");
  return *s;
}

DEMParameters load_parameters(const XmlNode& root) {
  // src/parser.cpp:49-75
  DEMParameters p;
  for (const auto& c : section(root, "parameters").children) {
    if (c->name == "timestep") p.timeStep = num_f(*c);
    else if (c->name == "simduration") p.simDuration = num_f(*c);
    else if (c->name == "imageheight") p.imageDIMy = (unsigned)num_i(*c);
    else if (c->name == "follow") p.followedParticles.push_back(num_i(*c));
  }
  if (p.timeStep == -1.0f) throw DsmlError("Simulation timestep not specified");
  return p;
}

DEMEnvironment load_environment(const XmlNode& root) {
  // src/parser.cpp:77-95
  DEMEnvironment e;
  for (const auto& c : section(root, "environment").children) {
    if (c->name == "dimensions") e.dimension = read_vector(*c);
    else if (c->name == "gravity") e.gravity = read_vector(*c);
    else if (c->name == "stiffness") {
      const std::string* dir = c->attr("dir");
      if (!dir) throw DsmlError("Stiffness direction not specified");
      if (*dir == "normal") e.boundaryNormalStiffness = num_f(*c);
      else if (*dir == "shear") e.boundaryShearStiffness = num_f(*c);
      else throw DsmlError("Unrecognized stiffness direction");
    }
  }
  return e;
}

float hex_channel(const std::string& color, size_t at) {
  // "rrggbb": each pair read as the hexadecimal number 0xNN (src/parser.cpp:119-135)
  if (color.size() < at + 2) throw DsmlError("Malformed particle type colour '" + color + "'");
  const std::string pair = color.substr(at, 2);
  char* end = nullptr;
  const long v = std::strtol(pair.c_str(), &end, 16);
  if (end != pair.c_str() + 2) throw DsmlError("Malformed particle type colour '" + color + "'");
  return (float)v;
}

DEMParticleType load_particle_type(const XmlNode& node) {
  // src/parser.cpp:109-167
  DEMParticleType t;
  if (const std::string* a = node.attr("id")) t.id = *a;
  if (const std::string* a = node.attr("name")) t.name = *a;
  if (const std::string* a = node.attr("color")) {
    t.color[0] = hex_channel(*a, 0);
    t.color[1] = hex_channel(*a, 2);
    t.color[2] = hex_channel(*a, 4);
  }
  for (const auto& c : node.children) {
    if (c->name == "mass") t.mass = num_f(*c);
    else if (c->name == "radius") t.radius = num_f(*c);
    else if (c->name == "stiffness") {
      const std::string* dir = c->attr("dir");
      if (!dir) throw DsmlError("Stiffness direction not specified");
      if (*dir == "normal") t.normalStiffness = num_f(*c);
      else if (*dir == "shear") t.shearStiffness = num_f(*c);
      else throw DsmlError("Unrecognized stiffness direction");
    } else if (c->name == "damping") {
      const std::string* type = c->attr("type");
      if (type) {
        if (*type == "boundary") t.boundaryDamping = num_f(*c);
        else throw DsmlError("Unrecognized damping type");
      } else {
        t.normalDamping = num_f(*c);
      }
    } else if (c->name == "friction") t.frictionCoefficient = num_f(*c);
    else if (c->name == "attraction") t.attractCoefficient = num_f(*c);
    else throw DsmlError("Unrecognized tag inside <particletype>");
  }
  return t;
}

DEMProperties load_properties(const XmlNode& root) {
  // src/parser.cpp:97-107
  DEMProperties p;
  for (const auto& c : section(root, "properties").children)
    if (c->name == "particletype") p.addParticleType(load_particle_type(*c));
  return p;
}

std::vector<int> type_list(const XmlNode& node, const DEMProperties& props, const char* shape,
                           const LoadOptions& opt) {
  std::vector<int> idx;
  if (const std::string* a = node.attr("particletype")) {
    for (const std::string& id : read_csv_line(*a)) idx.push_back(props.particleTypeIndexById(id));
  } else {
    idx.push_back(0);
    if (opt.notes)
      *opt.notes << "Particle type not specified for " << shape << ", defaulting to first one: "
                 << props.particleTypes[0].id << "\n";
  }
  return idx;
}

DEMParticles load_particles(const XmlNode& root, const DEMProperties& props, const LoadOptions& opt) {
  // src/parser.cpp:169-350
  DEMParticles parts;
  const XmlNode& sec = section(root, "particles");
  if (props.particleTypes.empty()) throw DsmlError("No particle types defined, aborting.");
  int singles = 0;
  for (const auto& cp : sec.children) {
    const XmlNode& c = *cp;
    if (c.name == "rectangle") {
      parts.types.push_back(type_list(c, props, "rectangle", opt));
      const XmlNode* s = c.child("start");
      if (!s) throw DsmlError("<start> tag not found inside rectangle");
      parts.start.push_back(read_vector(*s));
      const XmlNode* e = c.child("end");
      if (!e) throw DsmlError("<end> tag not found inside rectangle");
      parts.end.push_back(read_vector(*e));
      const XmlNode* n = c.child("num");
      if (!n) throw DsmlError("<num> tag not found inside rectangle");
      const Vec2 v = read_vector(*n);  // read as floats, then converted (make_uint2)
      if (!(v.x >= 0.0f && v.y >= 0.0f && v.x < 4294967296.0f && v.y < 4294967296.0f))
        throw DsmlError("Rectangle <num> out of range");
      parts.num.push_back(UVec2{(unsigned)v.x, (unsigned)v.y});
    }
    if (c.name == "triangle") {
      parts.t_types.push_back(type_list(c, props, "triangle", opt));
      const XmlNode* p = c.child("pos");
      if (!p) throw DsmlError("<pos> tag not found inside triangle");
      parts.t_pos.push_back(read_vector(*p));
      const XmlNode* w = c.child("width");
      if (!w) throw DsmlError("<width> tag not found inside triangle");
      parts.width.push_back(num_f(*w));
      const XmlNode* n = c.child("num");
      if (!n) throw DsmlError("<num> tag not found inside triangle");
      parts.t_num.push_back((unsigned)num_i(*n));
    }
    if (c.name == "particle") {
      int type = 0;
      if (const std::string* a = c.attr("particletype")) {
        type = props.particleTypeIndexById(*a);
      } else if (opt.notes) {
        *opt.notes << "Particle type not specified, defaulting to first one: " << props.particleTypes[0].id << "\n";
      }
      parts.type.push_back(type);
      if (const std::string* f = c.attr("follow"))
        if (*f == "true") parts.followedParticles.push_back(singles);
      const XmlNode* p = c.child("pos");
      if (!p) throw DsmlError("Missing <pos> tag on particle definition");
      parts.pos.push_back(read_vector(*p));
      const XmlNode* v = c.child("vel");
      parts.vel.push_back(v ? read_vector(*v) : Vec2{});
      const XmlNode* th = c.child("theta");
      parts.theta.push_back(th ? num_f(*th) : 0.0f);
      const XmlNode* om = c.child("omega");
      parts.omega.push_back(om ? num_f(*om) : 0.0f);
      ++singles;
    }
  }
  if (opt.controlled_particle_build) {
    // src/parser.cpp:325-347
    const XmlNode* n = sec.child("controlled");
    if (!n) throw DsmlError("Controlled particle tag missing");
    int type = 0;
    if (const std::string* a = n->attr("particletype")) {
      type = props.particleTypeIndexById(*a);
    } else if (opt.notes) {
      *opt.notes << "Controlled particle's type not specified, defaulting to first one: "
                 << props.particleTypes[0].id << "\n";
    }
    parts.controlledType = type;
  }
  return parts;
}

void load_root(DEMSimulation& sim, const XmlNode& root, const LoadOptions& opt) {
  if (root.name != "simulation") throw DsmlError("root element is <" + root.name + ">, not <simulation>");
  // src/DEMSimulation.cpp:45-59 (the same order of sections)
  sim.parameters = load_parameters(root);
  sim.environment = load_environment(root);
  sim.properties = load_properties(root);
  sim.particles = load_particles(root, sim.properties, opt);
}

std::ostream& operator<<(std::ostream& o, const Vec2& v) { return o << "(" << v.x << ", " << v.y << ")"; }
std::ostream& operator<<(std::ostream& o, const UVec2& v) { return o << "(" << v.x << ", " << v.y << ")"; }

}  // namespace

void DEMSimulation::loadFromFile(const std::string& filename, const LoadOptions& opt) {
  std::unique_ptr<XmlNode> root;
  try {
    root = parse_xml_file(filename);
  } catch (const XmlError& e) {
    throw DsmlError(e.what());
  }
  load_root(*this, *root, opt);
}

void DEMSimulation::loadFromString(const std::string& document, const LoadOptions& opt) {
  std::unique_ptr<XmlNode> root;
  try {
    root = parse_xml(document);
  } catch (const XmlError& e) {
    throw DsmlError(e.what());
  }
  load_root(*this, *root, opt);
}

void DEMSimulation::printConfiguration(std::ostream& out) const {
  out << "-- Parameters\n";
  out << "timeStep: " << parameters.timeStep << "\n";
  out << "simDuration: " << parameters.simDuration << "\n\n";
  out << "-- Environment\n";
  out << "dimensions: " << environment.dimension << "\n";
  out << "gravity: " << environment.gravity << "\n";
  out << "boundaryNormalStiffness: " << environment.boundaryNormalStiffness << "\n";
  out << "boundaryShearStiffness: " << environment.boundaryShearStiffness << "\n\n";
  out << "-- Properties\n";
  out << "numParticleTypes: " << properties.particleTypes.size() << "\n\n";
  for (size_t i = 0; i < properties.particleTypes.size(); ++i) {
    const DEMParticleType& t = properties.particleTypes[i];
    out << "\t---- Particle Type " << i << "\n";
    out << "\tid: " << t.id << "\n\tname: " << t.name << "\n";
    out << "\tcolor: (" << t.color[0] << ", " << t.color[1] << ", " << t.color[2] << ")\n";
    out << "\tmass: " << t.mass << "\n\tradius: " << t.radius << "\n";
    out << "\tnormalStiffness: " << t.normalStiffness << "\n\tshearStiffness: " << t.shearStiffness << "\n";
    out << "\tnormalDamping: " << t.normalDamping << "\n\tboundaryDamping: " << t.boundaryDamping << "\n";
    out << "\tfrictionCoefficient: " << t.frictionCoefficient << "\n";
    out << "\tattractCoefficient: " << t.attractCoefficient << "\n\n";
  }
  out << "-- Particle Rectangles\n";
  out << "numRectangles: " << particles.start.size() << "\n";
  for (size_t i = 0; i < particles.start.size(); ++i) {
    out << "\t---- Particle Rectangle " << i << "\n";
    out << "\tstart: " << particles.start[i] << "\n\tend: " << particles.end[i] << "\n";
    out << "\tnum: " << particles.num[i] << "\n\ttypes:";
    for (int t : particles.types[i]) out << " [" << t << "] = " << properties.particleTypes[t].id;
    out << "\n";
  }
  out << "\n-- Particle Triangles\n";
  out << "numTriangles: " << particles.t_pos.size() << "\n";
  for (size_t i = 0; i < particles.t_pos.size(); ++i) {
    out << "\t---- Particle Triangle " << i << "\n";
    out << "\tpos: " << particles.t_pos[i] << "\n\twidth: " << particles.width[i] << "\n";
    out << "\tnum: " << particles.t_num[i] << "\n\ttypes:";
    for (int t : particles.t_types[i]) out << " [" << t << "] = " << properties.particleTypes[t].id;
    out << "\n";
  }
  out << "\n-- Single Particles\n";
  out << "numSingleParticles: " << particles.pos.size() << "\n";
  for (size_t i = 0; i < particles.pos.size(); ++i) {
    out << "\t---- Particle " << i << "\n";
    out << "\ttype: [" << particles.type[i] << "] = " << properties.particleTypes[particles.type[i]].id << "\n";
    out << "\tpos: " << particles.pos[i] << "\n\tvel: " << particles.vel[i] << "\n";
    out << "\ttheta: " << particles.theta[i] << "\n\tomega: " << particles.omega[i] << "\n\n";
  }
  if (particles.controlledType >= 0)
    out << "-- Controlled particle type: [" << particles.controlledType << "] = "
        << properties.particleTypes[particles.controlledType].id << "\n";
}

}  // namespace dsml
