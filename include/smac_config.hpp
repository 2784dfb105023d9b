// smac_config.hpp -- the reference's JSON configuration and VTI output, header-only C++17.
//
// Mirrors CFD/config/SimulationConfig.h, CFD/config/ConfigLoader.cpp:1-882,
// CFD/config/SimulationConfig.cpp and the initial conditions of
// CFD/SimulationSetup.cu:39-160 (StaggeredGrid::place_vof, SMACSolver::set_sphere),
// plus FileInOut::output_vti_binary_cellData (CFD/FileInOut.cpp:145-255). Same struct
// and member names, same required keys, same error texts
// ("configuration error at \"<path>\": <message>", std::runtime_error).
//
// The reference parses with nlohmann/json; this header carries its own small
// recursive-descent reader of RFC 8259 JSON (objects, arrays, strings with
// escapes, numbers, true/false/null), which is all the configuration uses.
// Nothing here touches the GPU: the C-ABI glue (grid descriptor, uploads, VTI
// from a context) lives in smac_facade.hpp.
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace smac_config {

// ------------------------------------------------------------------ JSON value
struct Json {
  enum Kind { NUL, BOOL, INT, UINT, REAL, STRING, ARRAY, OBJECT };
  Kind kind = NUL;
  bool b = false;
  long long i = 0;
  unsigned long long u = 0;
  double d = 0.0;
  std::string s;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;  // in file order

  bool is_object() const { return kind == OBJECT; }
  bool is_array() const { return kind == ARRAY; }
  bool is_string() const { return kind == STRING; }
  bool is_boolean() const { return kind == BOOL; }
  bool is_number() const { return kind == INT || kind == UINT || kind == REAL; }
  bool is_number_integer() const { return kind == INT; }
  bool is_number_unsigned() const { return kind == UINT; }
  size_t size() const { return kind == ARRAY ? arr.size() : (kind == OBJECT ? obj.size() : 0); }
  const Json* find(const std::string& key) const {
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  bool contains(const std::string& key) const { return kind == OBJECT && find(key) != nullptr; }
  const Json& at(size_t k) const { return arr.at(k); }
  double number() const { return kind == REAL ? d : (kind == INT ? (double)i : (double)u); }
};

class JsonReader {
 public:
  explicit JsonReader(const std::string& text) : t_(text) {}
  Json parse() {
    Json v = value();
    ws();
    if (p_ != t_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& t_;
  size_t p_ = 0;

  [[noreturn]] void fail(const std::string& what) const {
    size_t line = 1, col = 1;
    for (size_t k = 0; k < p_ && k < t_.size(); k++) {
      if (t_[k] == '\n') {
        line++;
        col = 1;
      } else {
        col++;
      }
    }
    throw std::runtime_error("JSON parse error at line " + std::to_string(line) + ", column " +
                             std::to_string(col) + ": " + what);
  }
  void ws() {
    while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r')) p_++;
  }
  char peek() {
    ws();
    if (p_ >= t_.size()) fail("unexpected end of input");
    return t_[p_];
  }
  void expect(char c) {
    if (peek() != c) fail(std::string("expected '") + c + "'");
    p_++;
  }
  bool lit(const char* w) {
    size_t n = std::char_traits<char>::length(w);
    if (t_.compare(p_, n, w) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Json value() {
    char c = peek();
    Json v;
    if (c == '{') {
      p_++;
      v.kind = Json::OBJECT;
      if (peek() == '}') {
        p_++;
        return v;
      }
      for (;;) {
        if (peek() != '"') fail("expected a string key");
        std::string k = str();
        expect(':');
        Json m = value();
        v.obj.emplace_back(std::move(k), std::move(m));
        char d = peek();
        p_++;
        if (d == '}') break;
        if (d != ',') fail("expected ',' or '}'");
      }
      return v;
    }
    if (c == '[') {
      p_++;
      v.kind = Json::ARRAY;
      if (peek() == ']') {
        p_++;
        return v;
      }
      for (;;) {
        v.arr.push_back(value());
        char d = peek();
        p_++;
        if (d == ']') break;
        if (d != ',') fail("expected ',' or ']'");
      }
      return v;
    }
    if (c == '"') {
      v.kind = Json::STRING;
      v.s = str();
      return v;
    }
    if (lit("true")) {
      v.kind = Json::BOOL;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.kind = Json::BOOL;
      return v;
    }
    if (lit("null")) return v;
    return num();
  }
  static void utf8(std::string& out, unsigned cp) {
    if (cp < 0x80) {
      out += (char)cp;
    } else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }
  unsigned hex4() {
    if (p_ + 4 > t_.size()) fail("short \\u escape");
    unsigned v = 0;
    for (int k = 0; k < 4; k++) {
      char h = t_[p_++];
      v <<= 4;
      if (h >= '0' && h <= '9') v |= (unsigned)(h - '0');
      else if (h >= 'a' && h <= 'f') v |= (unsigned)(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') v |= (unsigned)(h - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string str() {
    expect('"');
    std::string out;
    for (;;) {
      if (p_ >= t_.size()) fail("unterminated string");
      char c = t_[p_++];
      if (c == '"') break;
      if ((unsigned char)c < 0x20) fail("control character in string");
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= t_.size()) fail("unterminated escape");
      char e = t_[p_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (!(p_ + 1 < t_.size() && t_[p_] == '\\' && t_[p_ + 1] == 'u')) fail("lone surrogate");
            p_ += 2;
            unsigned lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) fail("bad surrogate pair");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
    return out;
  }
  Json num() {
    size_t b = p_;
    if (p_ < t_.size() && t_[p_] == '-') p_++;
    auto digits = [&]() {
      size_t s = p_;
      while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') p_++;
      return p_ - s;
    };
    if (p_ < t_.size() && t_[p_] == '0') {
      p_++;
    } else if (digits() == 0) {
      fail("bad value");
    }
    bool integral = true;
    if (p_ < t_.size() && t_[p_] == '.') {
      p_++;
      integral = false;
      if (digits() == 0) fail("bad number");
    }
    if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
      p_++;
      integral = false;
      if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) p_++;
      if (digits() == 0) fail("bad number");
    }
    std::string tok = t_.substr(b, p_ - b);
    Json v;
    if (integral) {
      errno = 0;
      if (tok[0] == '-') {
        long long x = std::strtoll(tok.c_str(), nullptr, 10);
        if (errno == 0) {
          v.kind = Json::INT;
          v.i = x;
          return v;
        }
      } else {
        unsigned long long x = std::strtoull(tok.c_str(), nullptr, 10);
        if (errno == 0) {
          // non-negative integers are "unsigned" as in nlohmann::json
          v.kind = Json::UINT;
          v.u = x;
          return v;
        }
      }
    }
    v.kind = Json::REAL;
    v.d = std::strtod(tok.c_str(), nullptr);
    return v;
  }
};

// ------------------------------------------------------------------ the configuration
// CFD/Enums.h:10-16 and 18-26
enum BoundaryType { BC_NOSLIP = 0, BC_SLIP = 1, BC_INFLOW = 2, BC_OUTFLOW = 3, BC_OUTLET = 4 };
enum AXIS { AXIS_X = 0, AXIS_Y = 1, AXIS_Z = 2 };

struct ConfigVec3 {
  double x = 0.0, y = 0.0, z = 0.0;
};
struct HardwareConfig {
  bool use_gpu = true;
};
struct OutputConfig {
  std::string directory = "results";
};
struct GridConfig {
  int Nx = 1, Ny = 1, Nz = 1;
  double size_x = 1.0, size_y = 1.0, size_z = 1.0;
  double origin_x_ = 0.0, origin_y_ = 0.0, origin_z_ = 0.0;
  double dx() const { return size_x / static_cast<double>(Nx); }
  double dy() const { return size_y / static_cast<double>(Ny); }
  double dz() const { return size_z / static_cast<double>(Nz); }
};
struct PhaseConfig {
  double density = 1.0;
  double viscosity = 0.0;
};
struct FluidConfig {
  PhaseConfig phase0, phase1;
  double surface_tension = 0.0;
  ConfigVec3 gravity;
};
struct TimeConfig {
  double initial_dt = 1.0e-5;
  double maximum_dt = 1.0e-2;
  double end_time = 1.0;
  double interval = 0.05;
  double cfl = 0.4;
  double vof_cfl = 0.2;
};
struct PressureSolverConfig {
  int gmg_levels = 4;
  double tol = 1e-8;
};
struct BoundaryConditionConfig {
  int id = 0;
  BoundaryType type = BC_NOSLIP;
  ConfigVec3 velocity;
};
enum class ConfigBoundarySide { Min, Max };
struct BoundaryFaceConfig {
  AXIS axis = AXIS_X;
  ConfigBoundarySide side = ConfigBoundarySide::Min;
  unsigned char boundary_id = 0;
};
enum class InitialConditionType { VofBox, VofSphere, SolidBox, SolidCylinder, SolidCylinderInv };
struct InitialConditionConfig {
  InitialConditionType type = InitialConditionType::VofBox;
  ConfigVec3 minimum;
  ConfigVec3 maximum{1.0, 1.0, 1.0};
  ConfigVec3 center;
  double radius = 0.0;
  double vini = 0.0;
  double alpha = 0.0;
  unsigned char solid_id = 0;
};

[[noreturn]] inline void json_error(const std::string& path, const std::string& message) {
  throw std::runtime_error("configuration error at \"" + path + "\": " + message);
}
[[noreturn]] inline void config_error(const std::string& message) {
  throw std::runtime_error("configuration error: " + message);
}

struct SimulationConfig {
  HardwareConfig hardware;
  OutputConfig output;
  GridConfig grid;
  FluidConfig fluid;
  TimeConfig time;
  PressureSolverConfig pressure_solver;
  std::vector<BoundaryConditionConfig> boundary_conditions;
  std::vector<BoundaryFaceConfig> boundary_faces;
  std::vector<InitialConditionConfig> initial_conditions;

  // SimulationConfig::validate (CFD/config/SimulationConfig.cpp)
  void validate() const {
    auto finite = [](double v, const std::string& n) {
      if (!std::isfinite(v)) config_error(n + " must be finite");
    };
    auto positive = [&](double v, const std::string& n) {
      finite(v, n);
      if (v <= 0.0) config_error(n + " must be greater than zero");
    };
    auto nonneg = [&](double v, const std::string& n) {
      finite(v, n);
      if (v < 0.0) config_error(n + " must be nonnegative");
    };
    auto vec3 = [&](const ConfigVec3& v, const std::string& n) {
      finite(v.x, n + "[0]");
      finite(v.y, n + "[1]");
      finite(v.z, n + "[2]");
    };
    if (grid.Nx <= 0) config_error("grid.cells[0] must be positive");
    if (grid.Ny <= 0) config_error("grid.cells[1] must be positive");
    if (grid.Nz <= 0) config_error("grid.cells[2] must be positive");
    positive(grid.size_x, "grid.size[0]");
    positive(grid.size_y, "grid.size[1]");
    positive(grid.size_z, "grid.size[2]");
    positive(fluid.phase0.density, "fluid.phase0.density");
    positive(fluid.phase1.density, "fluid.phase1.density");
    nonneg(fluid.phase0.viscosity, "fluid.phase0.viscosity");
    nonneg(fluid.phase1.viscosity, "fluid.phase1.viscosity");
    nonneg(fluid.surface_tension, "fluid.surface_tension");
    vec3(fluid.gravity, "fluid.gravity");
    positive(time.initial_dt, "time.initial_dt");
    positive(time.maximum_dt, "time.maximum_dt");
    positive(time.end_time, "time.end_time");
    positive(time.cfl, "time.cfl");
    positive(time.interval, "time.interval");
    positive(time.vof_cfl, "time.vof_cfl");
    if (time.initial_dt > time.maximum_dt) config_error("time.initial_dt must not exceed time.maximum_dt");
    if (pressure_solver.gmg_levels <= 0) config_error("pressure_solver.gmg_levels must be positive");
    if (output.directory.empty()) config_error("output.directory must not be empty");
    std::set<int> ids;
    for (const auto& b : boundary_conditions) {
      if (b.id <= 0) config_error("boundary condition IDs must be positive");
      if (!ids.insert(b.id).second)
        config_error("boundary condition ID " + std::to_string(b.id) + " is defined more than once");
      vec3(b.velocity, "boundary velocity");
    }
    std::set<std::pair<int, int>> faces;
    for (const auto& f : boundary_faces) {
      if (f.axis != AXIS_X && f.axis != AXIS_Y && f.axis != AXIS_Z) config_error("boundary face has an invalid axis");
      if (ids.find(f.boundary_id) == ids.end())
        config_error("boundary face refers to undefined boundary ID " + std::to_string(f.boundary_id));
      const int side = f.side == ConfigBoundarySide::Min ? 0 : 1;
      if (!faces.insert({(int)f.axis, side}).second)
        config_error("the same outer boundary face is assigned more than once");
    }
    for (const auto& ic : initial_conditions) {
      vec3(ic.minimum, "initial condition minimum");
      vec3(ic.maximum, "initial condition maximum");
      if (ic.minimum.x > ic.maximum.x || ic.minimum.y > ic.maximum.y || ic.minimum.z > ic.maximum.z)
        config_error("initial condition minimum must not exceed maximum");
      if (ic.type == InitialConditionType::VofBox) {
        finite(ic.alpha, "initial condition alpha");
        if (ic.alpha < 0.0 || ic.alpha > 1.0) config_error("VOF alpha must be between zero and one");
      }
      if (ic.type == InitialConditionType::SolidBox && ic.solid_id <= 0) config_error("solid_id must be positive");
    }
  }
  // one past the largest boundary id (row 0 of the table exists as well)
  int get_num_boundary_ids() const {
    int m = 0;
    for (const auto& b : boundary_conditions) m = b.id > m ? b.id : m;
    return m + 1;
  }
  bool is_pure_neumann() const {
    for (const auto& b : boundary_conditions)
      if (b.type == BC_OUTLET || b.type == BC_INFLOW) return false;
    return true;
  }
  int get_alpha_substeps() const {
    const double ratio = time.cfl / time.vof_cfl;
    const int s = static_cast<int>(std::ceil(ratio));
    return s > 1 ? s : 1;
  }
};

// ------------------------------------------------------------------ the loader
namespace detail {
inline const Json& require_member(const Json& o, const std::string& key, const std::string& parent) {
  if (!o.is_object()) json_error(parent, "must be a JSON object");
  const Json* v = o.find(key);
  if (!v) json_error(parent + "." + key, "required value is missing");
  return *v;
}
inline const Json& require_object(const Json& p, const std::string& key, const std::string& parent) {
  const Json& v = require_member(p, key, parent);
  if (!v.is_object()) json_error(parent + "." + key, "must be an object");
  return v;
}
inline const Json& require_array(const Json& p, const std::string& key, const std::string& parent) {
  const Json& v = require_member(p, key, parent);
  if (!v.is_array()) json_error(parent + "." + key, "must be an array");
  return v;
}
inline std::string read_string(const Json& p, const std::string& key, const std::string& parent) {
  const Json& v = require_member(p, key, parent);
  if (!v.is_string()) json_error(parent + "." + key, "must be a string");
  return v.s;
}
inline bool read_bool(const Json& p, const std::string& key, const std::string& parent) {
  const Json& v = require_member(p, key, parent);
  if (!v.is_boolean()) json_error(parent + "." + key, "must be true or false");
  return v.b;
}
inline double read_number_value(const Json& v, const std::string& path) {
  if (!v.is_number()) json_error(path, "must be a number");
  return v.number();
}
inline double read_number(const Json& p, const std::string& key, const std::string& parent) {
  return read_number_value(require_member(p, key, parent), parent + "." + key);
}
inline int read_integer_value(const Json& v, const std::string& path) {
  if (!v.is_number_integer() && !v.is_number_unsigned()) json_error(path, "must be an integer");
  long long r = 0;
  if (v.is_number_unsigned()) {
    if (v.u > static_cast<unsigned long long>(std::numeric_limits<int>::max()))
      json_error(path, "is outside the range of int");
    r = static_cast<long long>(v.u);
  } else {
    r = v.i;
  }
  if (r < std::numeric_limits<int>::min() || r > std::numeric_limits<int>::max())
    json_error(path, "is outside the range of int");
  return static_cast<int>(r);
}
inline int read_integer(const Json& p, const std::string& key, const std::string& parent) {
  return read_integer_value(require_member(p, key, parent), parent + "." + key);
}
inline ConfigVec3 read_vec3_value(const Json& v, const std::string& path) {
  if (!v.is_array()) json_error(path, "must be an array");
  if (v.size() != 3) json_error(path, "must contain exactly three numbers");
  ConfigVec3 r;
  r.x = read_number_value(v.at(0), path + "[0]");
  r.y = read_number_value(v.at(1), path + "[1]");
  r.z = read_number_value(v.at(2), path + "[2]");
  return r;
}
inline ConfigVec3 read_vec3(const Json& p, const std::string& key, const std::string& parent) {
  return read_vec3_value(require_member(p, key, parent), parent + "." + key);
}
inline BoundaryType parse_boundary_type(const std::string& v, const std::string& path) {
  if (v == "noslip") return BC_NOSLIP;
  if (v == "slip") return BC_SLIP;
  if (v == "inflow") return BC_INFLOW;
  if (v == "outlet") return BC_OUTLET;
  json_error(path, "unknown boundary type \"" + v + "\"");
}
inline AXIS parse_axis(const std::string& v, const std::string& path) {
  if (v == "x") return AXIS_X;
  if (v == "y") return AXIS_Y;
  if (v == "z") return AXIS_Z;
  json_error(path, "unknown axis \"" + v + "\"");
}
inline ConfigBoundarySide parse_boundary_side(const std::string& v, const std::string& path) {
  if (v == "min") return ConfigBoundarySide::Min;
  if (v == "max") return ConfigBoundarySide::Max;
  json_error(path, "boundary side must be \"min\" or \"max\"");
}
inline InitialConditionConfig parse_initial_condition(const Json& v, size_t index) {
  const std::string path = "initial_conditions[" + std::to_string(index) + "]";
  InitialConditionConfig r;
  const std::string type = read_string(v, "type", path);
  if (type == "vof_box") {
    r.type = InitialConditionType::VofBox;
    r.alpha = read_number(v, "alpha", path);
    r.minimum = read_vec3(v, "minimum", path);
    r.maximum = read_vec3(v, "maximum", path);
    return r;
  }
  if (type == "vof_sphere") {
    r.type = InitialConditionType::VofSphere;
    r.vini = read_number(v, "vini", path);
    r.alpha = read_number(v, "alpha", path);
    r.radius = read_number(v, "radius", path);
    r.center = read_vec3(v, "center", path);
    return r;
  }
  if (type == "solid_box") {
    r.type = InitialConditionType::SolidBox;
    r.minimum = read_vec3(v, "minimum", path);
    r.maximum = read_vec3(v, "maximum", path);
    r.solid_id = (unsigned char)read_integer(v, "solid_id", path);
    return r;
  }
  if (type == "solid_cylinder_inv" || type == "solid_cylinder") {
    r.type = type == "solid_cylinder" ? InitialConditionType::SolidCylinder : InitialConditionType::SolidCylinderInv;
    r.center = read_vec3(v, "center", path);
    r.radius = read_number(v, "radius", path);
    r.solid_id = (unsigned char)read_integer(v, "solid_id", path);
    return r;
  }
  json_error(path + ".type", "unknown initial condition type \"" + type + "\"");
}
}  // namespace detail

// ConfigLoader: parse_root over the JSON text, then validate()
inline SimulationConfig parse_config_text(const std::string& text) {
  using namespace detail;
  const Json root = JsonReader(text).parse();
  if (!root.is_object()) json_error("$", "root must be a JSON object");
  SimulationConfig r;
  r.hardware.use_gpu = read_bool(require_object(root, "hardware", "$"), "use_gpu", "hardware");
  r.output.directory = read_string(require_object(root, "output", "$"), "directory", "output");
  {
    const Json& g = require_object(root, "grid", "$");
    const Json& cells = require_array(g, "cells", "grid");
    if (cells.size() != 3) json_error("grid.cells", "must contain exactly three integers");
    r.grid.Nx = read_integer_value(cells.at(0), "grid.cells[0]");
    r.grid.Ny = read_integer_value(cells.at(1), "grid.cells[1]");
    r.grid.Nz = read_integer_value(cells.at(2), "grid.cells[2]");
    const ConfigVec3 size = read_vec3(g, "size", "grid");
    r.grid.size_x = size.x;
    r.grid.size_y = size.y;
    r.grid.size_z = size.z;
    const ConfigVec3 o = read_vec3(g, "origin", "grid");
    r.grid.origin_x_ = o.x;
    r.grid.origin_y_ = o.y;
    r.grid.origin_z_ = o.z;
  }
  {
    const Json& f = require_object(root, "fluid", "$");
    const Json& p0 = require_object(f, "phase0", "fluid");
    r.fluid.phase0.density = read_number(p0, "density", "fluid.phase0");
    r.fluid.phase0.viscosity = read_number(p0, "viscosity", "fluid.phase0");
    const Json& p1 = require_object(f, "phase1", "fluid");
    r.fluid.phase1.density = read_number(p1, "density", "fluid.liquid");
    r.fluid.phase1.viscosity = read_number(p1, "viscosity", "fluid.liquid");
    r.fluid.surface_tension = read_number(f, "surface_tension", "fluid");
    r.fluid.gravity = read_vec3(f, "gravity", "fluid");
  }
  {
    const Json& t = require_object(root, "time", "$");
    r.time.initial_dt = read_number(t, "initial_dt", "time");
    r.time.maximum_dt = read_number(t, "maximum_dt", "time");
    r.time.end_time = read_number(t, "end_time", "time");
    r.time.cfl = read_number(t, "cfl", "time");
    r.time.vof_cfl = read_number(t, "vof_cfl", "time");
    r.time.interval = read_number(t, "interval", "time");
  }
  {
    const Json& ps = require_object(root, "pressure_solver", "$");
    r.pressure_solver.gmg_levels = read_integer(ps, "gmg_levels", "pressure_solver");
    r.pressure_solver.tol = read_number(ps, "tol", "pressure_solver");
  }
  const Json& bcs = require_array(root, "boundary_conditions", "$");
  for (size_t k = 0; k < bcs.size(); k++) {
    const std::string path = "boundary_conditions[" + std::to_string(k) + "]";
    const Json& v = bcs.at(k);
    BoundaryConditionConfig b;
    b.id = read_integer(v, "id", path);
    b.type = parse_boundary_type(read_string(v, "type", path), path + ".type");
    if (v.contains("velocity")) b.velocity = read_vec3(v, "velocity", path);
    r.boundary_conditions.push_back(b);
  }
  const Json& bfs = require_array(root, "boundary_faces", "$");
  for (size_t k = 0; k < bfs.size(); k++) {
    const std::string path = "boundary_faces[" + std::to_string(k) + "]";
    const Json& v = bfs.at(k);
    BoundaryFaceConfig f;
    f.axis = parse_axis(read_string(v, "axis", path), path + ".axis");
    f.side = parse_boundary_side(read_string(v, "side", path), path + ".side");
    f.boundary_id = (unsigned char)read_integer(v, "boundary_id", path);
    r.boundary_faces.push_back(f);
  }
  const Json& ics = require_array(root, "initial_conditions", "$");
  for (size_t k = 0; k < ics.size(); k++) r.initial_conditions.push_back(parse_initial_condition(ics.at(k), k));
  r.validate();
  return r;
}

inline SimulationConfig load_config(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open configuration file \"" + path + "\"");
  std::stringstream ss;
  ss << in.rdbuf();
  return parse_config_text(ss.str());
}

// ------------------------------------------------------------------ set-up on the host
// The bc table rows 0 .. get_num_boundary_ids()-1 (unlisted rows no-slip at
// rest, as the reference's calloc'd table) and the bc id of each side
// (StaggeredGrid::set_boundary_id through SimulationSetup::apply_boundary_conditions).
struct BoundarySetup {
  std::vector<BoundaryConditionConfig> table;  // index = bc id
  unsigned char side_bcid[6] = {0, 0, 0, 0, 0, 0};
};
inline BoundarySetup boundary_setup(const SimulationConfig& c) {
  BoundarySetup s;
  s.table.resize((size_t)c.get_num_boundary_ids());
  for (size_t k = 0; k < s.table.size(); k++) s.table[k].id = (int)k;
  for (const auto& b : c.boundary_conditions) s.table[(size_t)b.id] = b;
  for (const auto& f : c.boundary_faces)
    s.side_bcid[2 * (int)f.axis + (f.side == ConfigBoundarySide::Max ? 1 : 0)] = f.boundary_id;
  return s;
}

// alpha (cells) and v (y-faces) after the VoF initial conditions, in the
// reference padded layout, starting from zero: place_vof
// (CFD/StaggeredGrid.h:144-163) sets alpha on cells whose centre lies strictly
// inside the box; set_sphere (CFD/SMACSolver.cpp:642-693) sets alpha = 1 and
// v = vini on both y-faces of cells whose centre lies inside the sphere (the
// slot of zero width never applies). Solid conditions belong to the immersed
// boundary and are skipped.
struct InitialFields {
  std::vector<double> alpha, v;
};
inline InitialFields initial_fields(const SimulationConfig& c) {
  const GridConfig& g = c.grid;
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const size_t sx = (size_t)Nx + 2, sy = (size_t)Ny + 2, sz = (size_t)Nz + 2;
  InitialFields F;
  F.alpha.assign(sx * sy * sz, 0.0);
  F.v.assign(sx * (sy + 1) * sz, 0.0);
  const double dx = g.dx(), dy = g.dy(), dz = g.dz();
  auto cell = [&](int i, int j, int k) { return ((size_t)k * sy + (size_t)j) * sx + (size_t)i; };
  auto vface = [&](int i, int j, int k) { return ((size_t)k * (sy + 1) + (size_t)j) * sx + (size_t)i; };
  for (const auto& ic : c.initial_conditions) {
    if (ic.type != InitialConditionType::VofBox && ic.type != InitialConditionType::VofSphere) continue;
    const double rsq = ic.radius * ic.radius;
    for (int k = 1; k <= Nz; k++)
      for (int j = 1; j <= Ny; j++)
        for (int i = 1; i <= Nx; i++) {
          // StaggeredGrid::get_cell_coord (CFD/StaggeredGrid.h:121-142)
          const double x = dx * (double)i - 0.5 * dx + g.origin_x_;
          const double y = dy * (double)j - 0.5 * dy + g.origin_y_;
          const double z = dz * (double)k - 0.5 * dz + g.origin_z_;
          if (ic.type == InitialConditionType::VofBox) {
            if (x < ic.maximum.x && x > ic.minimum.x && y < ic.maximum.y && y > ic.minimum.y && z < ic.maximum.z &&
                z > ic.minimum.z)
              F.alpha[cell(i, j, k)] = ic.alpha;
          } else {
            const double ddx = x - ic.center.x, ddy = y - ic.center.y, ddz = z - ic.center.z;
            if (ddx * ddx + ddy * ddy + ddz * ddz < rsq) {
              F.alpha[cell(i, j, k)] = 1.0;
              F.v[vface(i, j, k)] = ic.vini;
              F.v[vface(i, j + 1, k)] = ic.vini;
            }
          }
        }
  }
  return F;
}

// ------------------------------------------------------------------ VTI output
// FileInOut::output_vti_binary_cellData (CFD/FileInOut.cpp:145-255): ImageData
// with the cell arrays pressure, alpha, velocity (3 components) and
// ibm_solid_fraction as raw appended Float32 blocks, each behind a UInt32 byte
// count. `cells` holds, in that order and x fastest over the interior,
// p[n], alpha[n], velocity[3n] (u, v, w interleaved); the solid fraction is 0.
inline std::string vti_filename(const std::string& outdir, int step) {
  char name[64];
  std::snprintf(name, sizeof(name), "/result_%06d.vti", step);
  return outdir + name;
}
inline void write_vti(const std::string& filename, const GridConfig& g, const float* cells) {
  const size_t n = (size_t)g.Nx * (size_t)g.Ny * (size_t)g.Nz;
  const size_t bs = n * sizeof(float), bv = 3 * bs;
  if (bv > 0xFFFFFFFFull) throw std::runtime_error("VTI block is too large for a UInt32 header");
  FILE* fp = std::fopen(filename.c_str(), "wb");
  if (!fp) throw std::runtime_error("cannot open VTI file " + filename);
  const size_t off_p = 0, off_a = off_p + 4 + bs, off_v = off_a + 4 + bs, off_s = off_v + 4 + bv;
  std::fprintf(fp, "<?xml version=\"1.0\"?>\n");
  std::fprintf(fp,
               "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n");
  std::fprintf(fp,
               "  <ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"%.15e %.15e %.15e\" "
               "Spacing=\"%.15e %.15e %.15e\">\n",
               g.Nx, g.Ny, g.Nz, g.origin_x_, g.origin_y_, g.origin_z_, g.dx(), g.dy(), g.dz());
  std::fprintf(fp, "    <Piece Extent=\"0 %d 0 %d 0 %d\">\n", g.Nx, g.Ny, g.Nz);
  std::fprintf(fp, "      <PointData></PointData>\n");
  std::fprintf(fp, "      <CellData Scalars=\"alpha\" Vectors=\"velocity\">\n");
  std::fprintf(fp, "        <DataArray type=\"Float32\" Name=\"pressure\" format=\"appended\" offset=\"%zu\"/>\n", off_p);
  std::fprintf(fp, "        <DataArray type=\"Float32\" Name=\"alpha\" format=\"appended\" offset=\"%zu\"/>\n", off_a);
  std::fprintf(fp,
               "        <DataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"3\" format=\"appended\" "
               "offset=\"%zu\"/>\n",
               off_v);
  std::fprintf(fp,
               "        <DataArray type=\"Float32\" Name=\"ibm_solid_fraction\" format=\"appended\" offset=\"%zu\"/>\n",
               off_s);
  std::fprintf(fp, "      </CellData>\n    </Piece>\n  </ImageData>\n  <AppendedData encoding=\"raw\">\n_");
  auto block = [&](const float* d, size_t cnt) {
    const uint32_t nb = (uint32_t)(cnt * sizeof(float));
    std::fwrite(&nb, sizeof(nb), 1, fp);
    if (cnt) std::fwrite(d, sizeof(float), cnt, fp);
  };
  block(cells, n);
  block(cells + n, n);
  block(cells + 2 * n, 3 * n);
  std::vector<float> zero(n, 0.0f);
  block(zero.data(), n);
  std::fprintf(fp, "\n  </AppendedData>\n</VTKFile>\n");
  const bool bad = std::ferror(fp) != 0;
  if (std::fclose(fp) != 0 || bad) throw std::runtime_error("write error on VTI file " + filename);
}

// the cell arrays of write_vti from host fields in the reference layout
// (p, alpha: cells; u, v, w: faces), cell velocity = mean of the two faces
inline std::vector<float> pack_cells(const GridConfig& g, const double* p, const double* alpha, const double* u,
                                     const double* v, const double* w) {
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const size_t n = (size_t)Nx * Ny * Nz;
  const size_t sx = (size_t)Nx + 2, sy = (size_t)Ny + 2;
  std::vector<float> out(5 * n);
  size_t m = 0;
  for (int k = 1; k <= Nz; k++)
    for (int j = 1; j <= Ny; j++)
      for (int i = 1; i <= Nx; i++, m++) {
        const size_t c = ((size_t)k * sy + j) * sx + i;
        const size_t fu = ((size_t)k * sy + j) * (sx + 1) + i;
        const size_t fv = ((size_t)k * (sy + 1) + j) * sx + i;
        const size_t fw = ((size_t)k * sy + j) * sx + i;
        out[m] = (float)p[c];
        out[n + m] = alpha ? (float)alpha[c] : 0.0f;
        out[2 * n + 3 * m + 0] = (float)(0.5 * (u[fu] + u[fu + 1]));
        out[2 * n + 3 * m + 1] = (float)(0.5 * (v[fv] + v[fv + sx]));
        out[2 * n + 3 * m + 2] = (float)(0.5 * (w[fw] + w[fw + sx * sy]));
      }
  return out;
}

}  // namespace smac_config
