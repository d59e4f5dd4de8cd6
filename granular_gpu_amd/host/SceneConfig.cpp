// Scene JSON reader (own minimal JSON parser) and scene builders.
#include "granular/SceneConfig.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>

namespace {

// Minimal JSON value: null, bool, number (double), string, array, object.
struct JValue {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<JValue> arr;
  std::map<std::string, JValue> obj;
};

class JParser {
 public:
  explicit JParser(const std::string &t) : s_(t), i_(0) {}

  JValue parse() {
    JValue v = value();
    ws();
    if (i_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string &s_;
  size_t i_;

  [[noreturn]] void fail(const char *what) const {
    throw std::runtime_error(std::string("scene JSON: ") + what + " at offset " +
                             std::to_string(i_));
  }
  void ws() {
    while (i_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[i_]))) ++i_;
  }
  bool eat(char c) {
    ws();
    if (i_ < s_.size() && s_[i_] == c) {
      ++i_;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail("unexpected character");
  }
  bool word(const char *w) {
    const size_t n = std::char_traits<char>::length(w);
    if (s_.compare(i_, n, w) == 0) {
      i_ += n;
      return true;
    }
    return false;
  }
  std::string string_lit() {
    expect('"');
    std::string out;
    while (i_ < s_.size() && s_[i_] != '"') {
      char c = s_[i_++];
      if (c == '\\') {
        if (i_ >= s_.size()) fail("bad escape");
        const char e = s_[i_++];
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
            if (i_ + 4 > s_.size()) fail("bad unicode escape");
            const unsigned cp = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
            i_ += 4;
            if (cp < 0x80) {
              out += (char)cp;
            } else if (cp < 0x800) {
              out += (char)(0xC0 | (cp >> 6));
              out += (char)(0x80 | (cp & 0x3F));
            } else {
              out += (char)(0xE0 | (cp >> 12));
              out += (char)(0x80 | ((cp >> 6) & 0x3F));
              out += (char)(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: fail("bad escape");
        }
      } else {
        out += c;
      }
    }
    if (i_ >= s_.size()) fail("unterminated string");
    ++i_;
    return out;
  }
  JValue value() {
    ws();
    if (i_ >= s_.size()) fail("unexpected end");
    JValue v;
    const char c = s_[i_];
    if (c == '{') {
      ++i_;
      v.kind = JValue::Object;
      if (eat('}')) return v;
      do {
        ws();
        std::string k = string_lit();
        expect(':');
        v.obj[k] = value();
      } while (eat(','));
      expect('}');
    } else if (c == '[') {
      ++i_;
      v.kind = JValue::Array;
      if (eat(']')) return v;
      do {
        v.arr.push_back(value());
      } while (eat(','));
      expect(']');
    } else if (c == '"') {
      v.kind = JValue::String;
      v.str = string_lit();
    } else if (word("true")) {
      v.kind = JValue::Bool;
      v.b = true;
    } else if (word("false")) {
      v.kind = JValue::Bool;
    } else if (word("null")) {
      v.kind = JValue::Null;
    } else {
      const char *start = s_.c_str() + i_;
      char *end = nullptr;
      v.num = std::strtod(start, &end);
      if (end == start) fail("bad value");
      i_ += (size_t)(end - start);
      v.kind = JValue::Number;
    }
    return v;
  }
};

const JValue &key(const JValue &o, const char *k) {
  if (o.kind != JValue::Object) throw std::runtime_error(std::string("scene JSON: not an object at ") + k);
  auto it = o.obj.find(k);
  if (it == o.obj.end()) throw std::runtime_error(std::string("scene JSON: missing key ") + k);
  return it->second;
}

double number(const JValue &v, const char *k) {
  if (v.kind == JValue::Number) return v.num;
  if (v.kind == JValue::Bool) return v.b ? 1.0 : 0.0;
  throw std::runtime_error(std::string("scene JSON: key ") + k + " is not a number");
}

float f32(const JValue &o, const char *k) { return (float)number(key(o, k), k); }

float3 vec3(const JValue &o, const char *k) {
  const JValue &v = key(o, k);
  return make_float3(f32(v, "x"), f32(v, "y"), f32(v, "z"));
}

int i32(const JValue &o, const char *k) { return (int)number(key(o, k), k); }

std::string str(const JValue &o, const char *k) {
  const JValue &v = key(o, k);
  if (v.kind != JValue::String) throw std::runtime_error(std::string("scene JSON: key ") + k + " is not a string");
  return v.str;
}

}  // namespace

SceneConfig parseSceneConfig(const std::string &text) {
  const JValue j = JParser(text).parse();
  SceneConfig c;
  c.space_size = vec3(j, "space_size");
  c.dt = f32(j, "dt");
  c.G = vec3(j, "gravity");
  c.box_size = vec3(j, "box_size");
  c.sphSpacing = f32(j, "sph_spacing");
  c.initSpacing = f32(j, "init_spacing");
  c.smoothing_radius = f32(j, "smoothing_radius");
  c.cell_length = f32(j, "cell_length");
  c.density = f32(j, "density");
  c.scene = i32(j, "scene");
  c.boundary_translation = vec3(j, "boundary_translation");
  c.particle_translation = vec3(j, "particle_translation");
  c.cell_size = make_int3((int)std::ceil(c.space_size.x / c.cell_length),
                          (int)std::ceil(c.space_size.y / c.cell_length),
                          (int)std::ceil(c.space_size.z / c.cell_length));
  c.obj_file = str(j, "obj_file");
  c.is_adaptive = i32(j, "is_adaptive");
  c.save_file = str(j, "save_file");
  return c;
}

SceneConfig loadSceneConfig(const std::string &config_file) {
  std::ifstream file(config_file);
  if (!file.is_open()) throw std::runtime_error("Unable to open config file: " + config_file);
  std::stringstream ss;
  ss << file.rdbuf();
  return parseSceneConfig(ss.str());
}

std::vector<float3> readBoundaryParticlesFromFile(const std::string &filename, int &start_index) {
  std::vector<float3> positions;
  std::ifstream file(filename);
  if (!file.is_open()) throw std::runtime_error("Unable to open boundary particles file: " + filename);
  float x, y, z;
  while (file >> x >> y >> z) positions.push_back(make_float3(x, y, z));
  start_index = (int)positions.size();
  return positions;
}

std::vector<float3> granular_lattice(const SceneConfig &c) {
  std::vector<float3> pos;
  const float s = c.initSpacing;
  const float3 t = c.particle_translation;
  for (int i = 0; i < c.box_size.x; ++i)
    for (int j = 0; j < c.box_size.y; ++j)
      for (int k = 0; k < c.box_size.z; ++k) {
        const float x = 0.27f + s * (float)j, y = 0.13f + s * (float)i, z = 0.17f + s * (float)k;
        pos.push_back(make_float3(x + t.x, y + t.y, z + t.z));
      }
  return pos;
}

std::vector<float3> boundary_walls(const SceneConfig &c) {
  const int sc = c.scene;
  if (sc < PILING || sc > BOX_OPENING) throw std::runtime_error("boundary_walls: unknown scene id");
  const bool open_top = sc == EXCAVATOR || sc == FUNNEL;
  const float3 S = c.space_size;
  const int nx = 2 * (int)std::ceil(S.x / c.cell_length);
  const int ny = 2 * (int)std::ceil(S.y / c.cell_length);
  const int nz = 2 * (int)std::ceil(S.z / c.cell_length);
  const float dx = (float)(nx - 1), dy = (float)(ny - 1), dz = (float)(nz - 1);
  std::vector<float3> pos;
  auto put = [&](int i, int j, int k) {
    const float x = ((float)i / dx) * S.x, y = ((float)j / dy) * S.y, z = ((float)k / dz) * S.z;
    pos.push_back(make_float3(0.99f * x + 0.005f * S.x, 0.99f * y + 0.005f * S.y,
                              0.99f * z + 0.005f * S.z));
  };
  const bool sides = sc != PILING;
  const bool far = sc != CORNER;
  if (sides)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        put(i, j, 0);
        if (far) put(i, j, nz - 1);
      }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < nz - 2; ++j) {
      put(i, 0, j + 1);
      if (far && !open_top) put(i, ny - 1, j + 1);
    }
  if (sides)
    for (int i = 0; i < ny - 2; ++i)
      for (int j = 0; j < nz - 2; ++j) {
        put(0, i + 1, j + 1);
        if (far) put(nx - 1, i + 1, j + 1);
      }
  if (open_top) return pos;
  const float3 t = c.boundary_translation;
  for (auto &p : pos) p = make_float3(p.x + t.x, p.y + t.y, p.z + t.z);
  return pos;
}

namespace {

// Sum in insertion order, then divide by the count, in float.
float3 centroid_of(const std::vector<float3> &v) {
  float3 s = make_float3(0.0f, 0.0f, 0.0f);
  for (const auto &p : v) s = make_float3(s.x + p.x, s.y + p.y, s.z + p.z);
  const float n = (float)v.size();
  return make_float3(s.x / n, s.y / n, s.z / n);
}

}  // namespace

SceneBoundary scene_boundary(const SceneConfig &c) {
  SceneBoundary b;
  b.pos = boundary_walls(c);
  b.animated.assign(b.pos.size(), 0);
  if (c.scene != EXCAVATOR && c.scene != FUNNEL) return b;
  if (c.obj_file.empty() || !std::filesystem::exists(c.obj_file)) return b;
  int count = 0;
  std::vector<float3> cloud = readBoundaryParticlesFromFile(c.obj_file, count);
  const float3 centre = centroid_of(cloud);
  b.anim.start_index = (int)b.pos.size();
  b.anim.num_particles = (int)cloud.size();
  b.anim.is_animating = true;
  b.anim.animation_time = 0.0f;
  if (c.scene == EXCAVATOR) {
    b.anim.rotation_center = centre;
    for (auto &p : cloud) p = make_float3(p.x + 2.5f, p.y + 0.0f, p.z + 0.5f);
  } else {
    const float angle = (float)(90.0 * M_PI / 180.0);
    const float cs = std::cos(angle), sn = std::sin(angle);
    for (auto &p : cloud) {
      const float x = p.x - centre.x, y = p.y - centre.y, z = p.z - centre.z;
      const float ry = y * cs - z * sn;
      const float rz = y * sn + z * cs;
      p = make_float3((x + centre.x) + 1.25f, (ry + centre.y) + 0.3f, (rz + centre.z) + 1.5f);
    }
  }
  b.pos.insert(b.pos.end(), cloud.begin(), cloud.end());
  b.animated.resize(b.pos.size(), 1);
  return b;
}

std::vector<float3> upsampled_cloud(const std::vector<float3> &coarse, float radius, int per) {
  std::vector<float3> out;
  out.reserve(coarse.size() * (size_t)(per > 0 ? per : 0));
  std::default_random_engine generator;
  std::uniform_real_distribution<float> distribution(-radius, radius);
  for (const auto &p : coarse)
    for (int n = 0; n < per; ++n) {
      const float ox = distribution(generator);
      const float oy = distribution(generator);
      const float oz = distribution(generator);
      out.push_back(make_float3(p.x + ox, p.y + oy, p.z + oz));
    }
  return out;
}
