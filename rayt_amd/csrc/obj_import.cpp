// Wavefront OBJ scenes as octree sources; see obj_import.h.
//
// Arithmetic follows the reference's types and evaluation order: vertex
// positions and node boxes in double, everything sampled or accumulated per
// node in float, `v / d` as `v * (1 / d)` (util/vec.h:23,31), dot products
// left to right, and the triangle/box test on the float-rounded box and
// vertices (the reference converts its dvec3 arguments to fvec3 at the call,
// import-obj-scene.cpp:570). The library is built with -ffp-contract=off, so
// every operation is one rounded IEEE operation.
#include "obj_import.h"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <thread>
#include <type_traits>
#include <vector>

namespace rayt {
namespace {

// ------------------------------------------------------------------ vectors
template <class T>
struct V3 {
  T x = 0, y = 0, z = 0;
  V3() = default;
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  template <class U>
  explicit V3(const V3<U>& v) : x(T(v.x)), y(T(v.y)), z(T(v.z)) {}
  V3 operator+(const V3& v) const { return V3(x + v.x, y + v.y, z + v.z); }
  V3 operator-(const V3& v) const { return V3(x - v.x, y - v.y, z - v.z); }
  V3 operator*(T d) const { return V3(x * d, y * d, z * d); }
  V3 Mul(const V3& v) const { return V3(x * v.x, y * v.y, z * v.z); }
  V3 Div(T d) const {  // vec.h:23: d = 1.0f / d, then three products
    d = T(1.0f) / d;
    return V3(x * d, y * d, z * d);
  }
  T Dot(const V3& v) const { return x * v.x + y * v.y + z * v.z; }
  V3 Cross(const V3& v) const { return V3(y * v.z - z * v.y, z * v.x - x * v.z, x * v.y - y * v.x); }
  T LengthSquare() const { return Dot(*this); }
  T Length() const { return std::sqrt(LengthSquare()); }
  V3 Normalized() const { return Div(Length()); }
  V3 Min(const V3& v) const { return V3(v.x < x ? v.x : x, v.y < y ? v.y : y, v.z < z ? v.z : z); }
  V3 Max(const V3& v) const { return V3(x < v.x ? v.x : x, y < v.y ? v.y : y, z < v.z ? v.z : z); }
  bool AllLessThan(const V3& v) const { return x < v.x && y < v.y && z < v.z; }
  bool AllGreaterThan(const V3& v) const { return x > v.x && y > v.y && z > v.z; }
};
using F3 = V3<float>;
using D3 = V3<double>;

// tvec3::ToBarycentric (util/vec.h:55-74).
template <class T>
void ToBarycentric(const V3<T>& p, const V3<T>* tri, T* out, bool clamp) {
  const V3<T> n = (tri[1] - tri[0]).Cross(tri[2] - tri[0]);
  T sum = 0;
  for (int i = 0; i < 3; ++i) {
    out[i] = (tri[(i + 2) % 3] - tri[(i + 1) % 3]).Cross(p - tri[(i + 1) % 3]).Dot(n);
    if (clamp && out[i] < 0) out[i] = 0;
    sum += out[i];
  }
  if (double(std::fabs(sum)) >= 1e-5) {
    for (int i = 0; i < 3; ++i) out[i] /= sum;
  } else {
    out[0] = out[1] = out[2] = static_cast<T>(1. / 3);
  }
}

// ------------------------------------------------------- triangle vs box
// Separation of box and triangle along `axis` (intersections.cpp:7-27).
bool Divides(const F3& axis, const F3& box_min, const F3& box_max, const F3* t) {
  const float d0 = axis.Dot(t[0]), d1 = axis.Dot(t[1]), d2 = axis.Dot(t[2]);
  float tmin = d1 < d0 ? d1 : d0;
  tmin = d2 < tmin ? d2 : tmin;
  float tmax = d0 < d1 ? d1 : d0;
  tmax = tmax < d2 ? d2 : tmax;
  F3 pmin = box_min, pmax = box_max;
  if (axis.x < 0) std::swap(pmin.x, pmax.x);
  if (axis.y < 0) std::swap(pmin.y, pmax.y);
  if (axis.z < 0) std::swap(pmin.z, pmax.z);
  const float bmin = axis.Dot(pmin), bmax = axis.Dot(pmax);
  return bmin > tmax || tmin > bmax;
}

struct FaceGeom {
  F3 v[3];            // float-rounded vertices
  F3 bmin, bmax;      // their bounding box
};

// TriangleBoxIntersection (intersections.cpp:29-59).
bool TriangleBoxIntersection(const F3& box_min, const F3& box_max, const FaceGeom& g) {
  if (!g.bmin.AllLessThan(box_max) || !g.bmax.AllGreaterThan(box_min)) return false;
  for (int k = 0; k < 3; ++k)
    if (g.v[k].AllGreaterThan(box_min) && g.v[k].AllLessThan(box_max)) return true;
  if (Divides((g.v[1] - g.v[0]).Cross(g.v[2] - g.v[0]), box_min, box_max, g.v)) return false;
  const F3 box_sides[3] = {F3(0, 0, 1), F3(0, 1, 0), F3(1, 0, 0)};
  const F3 tri_sides[3] = {g.v[1] - g.v[0], g.v[2] - g.v[1], g.v[0] - g.v[2]};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (Divides(box_sides[i].Cross(tri_sides[j]), box_min, box_max, g.v)) return false;
  return true;
}

// ------------------------------------------------------------------ textures
struct Texture {
  bool valid = false;
  int width = 0, height = 0, channels = 0;
  std::vector<float> px;  // channel planes, rows of `width` (CImg's layout)

  float At(unsigned x, unsigned y, int c) const { return px[(size_t(c) * height + y) * width + x]; }

  // CImg<float>::_linear_atXY (libs/CImg.h:11725-11742): coordinates clamped
  // to the image, bilinear weights from the fractional parts.
  float Linear(float fx, float fy, int c) const {
    float nfx = fx < 0 ? 0.0f : (fx > float(width - 1) ? float(width - 1) : fx);
    float nfy = fy < 0 ? 0.0f : (fy > float(height - 1) ? float(height - 1) : fy);
    if (nfx != nfx) nfx = 0;  // NaN coordinates (undefined in CImg) read the first column / row
    if (nfy != nfy) nfy = 0;
    const unsigned x = unsigned(nfx), y = unsigned(nfy);
    const float dx = nfx - float(x), dy = nfy - float(y);
    const unsigned nx = dx > 0 ? x + 1 : x, ny = dy > 0 ? y + 1 : y;
    const float icc = At(x, y, c), inc = At(nx, y, c), icn = At(x, ny, c), inn = At(nx, ny, c);
    return icc + dx * (inc - icc + dy * (icc + inn - icn - inc)) + dy * (icn - icc);
  }

  // Texture::ColorAt / XyAt (import-obj-scene.cpp:48-66).
  F3 ColorAt(float u, float v) const {
    if (!valid) return F3(1, 1, 1);
    u -= std::floor(u);
    v -= std::floor(v);
    u *= float(width);
    v *= float(height);
    return F3(Linear(u, v, 0), Linear(u, v, 1), Linear(u, v, 2)).Div(255.0f);
  }
  F3 XyAt(float u, float v) const {
    if (!valid) return F3(1, 1, 1);
    u -= std::floor(u);
    v -= std::floor(v);
    u *= float(width);
    v *= float(height);
    return F3(Linear(u, v, 0), Linear(u, v, 1), 0).Div(255.0f);
  }
};

bool ReadWholeFile(const std::string& path, std::string* out) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  std::ostringstream ss;
  ss << in.rdbuf();
  *out = ss.str();
  return true;
}

// CImg 1.4.9 load_pnm (libs/CImg.h:33979-34140) for P2, P3, P5 and P6: the
// header is read line by line (comment lines skipped), "W H" then the
// maximum on its own line or "W H max" on one line; binary samples above 255
// are 16-bit big-endian. Values are stored unscaled.
bool LoadPnm(const std::string& path, Texture* t, std::string* err) {
  std::string buf;
  if (!ReadWholeFile(path, &buf)) {
    *err = "cannot open texture '" + path + "'";
    return false;
  }
  size_t pos = 0;
  auto next_line = [&](bool skip_space, std::string* item) -> bool {
    for (;;) {
      if (skip_space)
        while (pos < buf.size() && std::isspace(static_cast<unsigned char>(buf[pos]))) ++pos;
      if (pos >= buf.size()) return false;
      size_t e = buf.find('\n', pos);
      if (e == std::string::npos) e = buf.size();
      *item = buf.substr(pos, e - pos);
      pos = e;
      if (!item->empty() && (*item)[0] != '#') return true;
      if (pos < buf.size()) ++pos;  // the newline after a comment or an empty line
    }
  };
  std::string item;
  unsigned type = 0, w = 0, h = 0, d = 1, cmax = 255;
  if (!next_line(false, &item) || std::sscanf(item.c_str(), " P%u", &type) != 1) {
    *err = "PNM header not found in '" + path + "'";
    return false;
  }
  if (type != 2 && type != 3 && type != 5 && type != 6) {
    *err = "PNM type P" + std::to_string(type) + " not supported ('" + path + "')";
    return false;
  }
  int n = next_line(true, &item) ? std::sscanf(item.c_str(), " %u %u %u %u", &w, &h, &d, &cmax) : 0;
  if (n < 2) {
    *err = "PNM width and height undefined in '" + path + "'";
    return false;
  }
  if (n == 2) {
    if (!next_line(true, &item) || std::sscanf(item.c_str(), "%u", &cmax) != 1) cmax = 255;
  } else if (n == 3 && type != 5) {
    cmax = d;
  } else {
    *err = "3-D PNM header not supported ('" + path + "')";
    return false;
  }
  if (pos < buf.size()) ++pos;  // the byte after the header
  if (w == 0 || h == 0 || uint64_t(w) * h > (1ull << 28)) {
    *err = "PNM size out of range in '" + path + "'";
    return false;
  }
  t->width = int(w);
  t->height = int(h);
  t->channels = (type == 3 || type == 6) ? 3 : 1;
  const size_t plane = size_t(w) * h;
  t->px.assign(plane * t->channels, 0.0f);
  if (type == 2 || type == 3) {
    const char* p = buf.c_str() + pos;
    for (size_t i = 0; i < plane; ++i)
      for (int c = 0; c < t->channels; ++c) {
        char* e = nullptr;
        long v = std::strtol(p, &e, 10);
        if (e == p) {
          *err = "PNM samples missing in '" + path + "'";
          return false;
        }
        p = e;
        t->px[c * plane + i] = float(int(v));
      }
  } else {
    const int bytes = cmax < 256 ? 1 : 2;
    if (buf.size() - std::min(pos, buf.size()) < plane * t->channels * bytes) {
      *err = "PNM samples missing in '" + path + "'";
      return false;
    }
    const auto* s = reinterpret_cast<const uint8_t*>(buf.data() + pos);
    for (size_t i = 0; i < plane; ++i)
      for (int c = 0; c < t->channels; ++c) {
        const size_t k = i * t->channels + c;
        t->px[c * plane + i] = bytes == 1 ? float(s[k]) : float((unsigned(s[2 * k]) << 8) | s[2 * k + 1]);
      }
  }
  t->valid = true;
  return true;
}

// Texture::Load: CImg picks the reader by extension; the PNM family is the
// one read here.
bool LoadTexture(const std::string& path, Texture* t, std::string* err) {
  const size_t dot = path.find_last_of('.');
  std::string ext = dot == std::string::npos ? "" : path.substr(dot + 1);
  for (auto& ch : ext) ch = char(std::tolower(static_cast<unsigned char>(ch)));
  if (ext != "ppm" && ext != "pgm" && ext != "pnm") {
    *err = "texture format not supported (PNM only): '" + path + "'";
    return false;
  }
  return LoadPnm(path, t, err);
}

// ------------------------------------------------------------------ materials
struct ColorMap {
  F3 multiplier;
  Texture texture;
  F3 ColorAt(float u, float v) const { return texture.valid ? texture.ColorAt(u, v).Mul(multiplier) : multiplier; }
};
struct ScalarMap {
  float multiplier = 0;
  Texture texture;
};
struct NormalMap {
  Texture texture;
  // NormalMap::Build (import-obj-scene.cpp:109-130): central differences of
  // the bump map's first channel, wrapped at the edges, times multiplier / 2.
  void Build(const ScalarMap& bump) {
    const Texture& b = bump.texture;
    const int w = b.width, h = b.height;
    texture.valid = true;
    texture.width = w;
    texture.height = h;
    texture.channels = 2;
    texture.px.assign(size_t(w) * h * 2, 0.0f);
    const size_t plane = size_t(w) * h;
    for (int i = 0; i < h; ++i)
      for (int j = 0; j < w; ++j) {
        texture.px[size_t(i) * w + j] = (b.At((j + w - 1) % w, i, 0) - b.At((j + 1) % w, i, 0)) * bump.multiplier / 2;
        texture.px[plane + size_t(i) * w + j] =
            (b.At(j, (i + h - 1) % h, 0) - b.At(j, (i + 1) % h, 0)) * bump.multiplier / 2;
      }
  }
  F3 NormalAt(float u, float v) const {
    if (!texture.valid) return F3(0, 0, 1);
    F3 t = texture.XyAt(u, v);
    t.z = 1;
    return t;
  }
};
struct Material {
  ColorMap ambient, diffuse, specular, transmission_filter;
  ScalarMap specular_exponent, alpha, optical_density, bump;
  NormalMap normal;
};

// Lines as std::istream::getline into a 2048-byte buffer reads them: a line
// of 2048 or more characters ends the file, a final line without a newline
// counts, and each line ends at its first NUL.
std::vector<std::string> SplitLines(const std::string& buf) {
  std::vector<std::string> out;
  size_t pos = 0;
  while (pos < buf.size()) {
    size_t e = buf.find('\n', pos);
    const size_t end = e == std::string::npos ? buf.size() : e;
    if (end - pos >= 2048) break;
    std::string line = buf.substr(pos, end - pos);
    const size_t nul = line.find('\0');
    if (nul != std::string::npos) line.resize(nul);
    out.push_back(std::move(line));
    if (e == std::string::npos) break;
    pos = e + 1;
  }
  return out;
}

// A whitespace tokenizer with the failure rules of std::istream >> on the
// reference's stringstream: reading at the end fails and leaves the value;
// a token that does not start with a number stores 0 and fails the stream.
struct LineStream {
  const std::string& s;
  size_t p = 0;
  bool failed = false;
  explicit LineStream(const std::string& line) : s(line) {}
  void Skip() {
    while (p < s.size() && std::isspace(static_cast<unsigned char>(s[p]))) ++p;
  }
  bool Token(std::string* out) {
    if (failed) return false;
    Skip();
    if (p >= s.size()) return failed = true, false;
    size_t b = p;
    while (p < s.size() && !std::isspace(static_cast<unsigned char>(s[p]))) ++p;
    *out = s.substr(b, p - b);
    return true;
  }
  template <class T>
  bool Number(T* out) {
    if (failed) return false;
    Skip();
    if (p >= s.size()) return failed = true, false;
    const char* b = s.c_str() + p;
    char* e = nullptr;
    const double v = std::strtod(b, &e);
    if (e == b) {
      *out = 0;
      return failed = true, false;
    }
    // a float reads its own nearest value, not the double's rounding
    *out = std::is_same<T, float>::value ? T(std::strtof(b, nullptr)) : T(v);
    p += size_t(e - b);
    return true;
  }
};

struct Model {
  std::vector<std::unique_ptr<Material>> materials;
  std::map<std::string, int> index;
  std::vector<D3> vert_pos;
  std::vector<F3> vert_tex, vert_norm;
  struct Face {
    int pos[3] = {0, 0, 0}, tex[3] = {0, 0, 0}, normal[3] = {0, 0, 0};
    int material = 0;
    F3 tangent_u, tangent_v, tangent_n;
  };
  std::vector<Face> faces;
  uint64_t unsupported = 0;

  Model() {  // MaterialLibrary() (import-obj-scene.cpp:157-164)
    index[""] = 0;
    materials.emplace_back(new Material());
    materials[0]->diffuse.multiplier = F3(1, 1, 1);
    materials[0]->ambient.multiplier = F3(float(.1), float(.1), float(.1));
    materials[0]->alpha.multiplier = 1;
  }
};

std::string DirOf(const std::string& path) {  // RemoveFileNameFromPath (util/string-util.cpp:52-57)
  const size_t p = path.find_last_of("/\\");
  return p == std::string::npos ? "" : path.substr(0, p + 1);
}
std::string CombinePaths(std::string a, const std::string& b) {  // string-util.cpp:59-66
  if (a.empty()) return b;
  const char l = a.back();
  if (l != '/' && l != '\\') a += "/";
  return a + b;
}

// MaterialLibrary::Load (import-obj-scene.cpp:166-279). A file that cannot be
// opened adds nothing (the reference's stream check does not catch it).
bool LoadMtl(const std::string& path, Model* m, std::string* err) {
  std::string buf;
  if (!ReadWholeFile(path, &buf)) return true;
  const std::string prefix = DirOf(path);
  Material* cur = m->materials[0].get();
  for (const std::string& line : SplitLines(buf)) {
    if (!line.empty() && line[0] == '#') continue;
    LineStream ss(line);
    std::string s;
    ss.Token(&s);
    if (s == "newmtl") {
      std::string name;
      ss.Token(&name);
      if (name.empty()) return *err = "newmtl without name", false;
      if (m->index.count(name)) return *err = "duplicate material name", false;
      m->index[name] = int(m->materials.size());
      m->materials.emplace_back(new Material());
      cur = m->materials.back().get();
      continue;
    }
    bool is_map = false;
    std::string map_file;
    double bump_mult = 1;
    if (s == "bump" || s.compare(0, 4, "map_") == 0) {
      is_map = true;
      if (s.compare(0, 4, "map_") == 0) s = s.substr(4);
      std::string tok;
      while (ss.Token(&tok)) {
        if (tok[0] == '-') {
          if (tok == "-bm")
            ss.Number(&bump_mult);
          else
            ++m->unsupported;
        } else {
          map_file = tok;
        }
      }
      map_file = prefix + map_file;
    }
    if (s == "Ns" || s == "d" || s == "Ni") {
      ScalarMap* sm = s == "Ns" ? &cur->specular_exponent : s == "Ni" ? &cur->optical_density : &cur->alpha;
      if (is_map) {
        if (!LoadTexture(map_file, &sm->texture, err)) return false;
      } else {
        float v;
        if (!ss.Number(&v)) return *err = "invalid value after " + s, false;
        sm->multiplier = v;
      }
    } else if (s == "bump") {
      if (!LoadTexture(map_file, &cur->bump.texture, err)) return false;
      cur->bump.multiplier = float(bump_mult);
      cur->normal.Build(cur->bump);
    } else if (s == "Ka" || s == "Kd" || s == "Ks" || s == "Tf") {
      ColorMap* cm = s == "Ka" ? &cur->ambient : s == "Kd" ? &cur->diffuse : s == "Ks" ? &cur->specular
                                                                                  : &cur->transmission_filter;
      if (is_map) {
        if (!LoadTexture(map_file, &cm->texture, err)) return false;
        if ((s == "Ka" || s == "Kd") && cm->texture.channels < 3)
          return *err = "colour texture with fewer than 3 channels: '" + map_file + "'", false;
      } else {
        F3 v;
        if (!(ss.Number(&v.x) && ss.Number(&v.y) && ss.Number(&v.z))) return *err = "invalid value after " + s, false;
        cm->multiplier = v;
      }
    } else {
      ++m->unsupported;
    }
  }
  return true;
}

// ReadObjFile (import-obj-scene.cpp:334-479) and CalculateTangentSpaces
// (:301-331). `vt` keeps z = 0 (the reference leaves it unset).
bool ReadObj(const std::string& path, Model* m, std::string* err) {
  std::string buf;
  if (!ReadWholeFile(path, &buf)) return *err = "failed to open obj file '" + path + "'", false;
  m->vert_pos.push_back(D3(0, 0, 0));
  m->vert_norm.push_back(F3(0, 0, 0));
  m->vert_tex.push_back(F3(0, 0, 0));
  int cur_mtl = 0;
  for (const std::string& line : SplitLines(buf)) {
    const char* l = line.c_str();
    if (l[0] == 'v') {
      if (l[1] == ' ') {
        D3 v;
        if (std::sscanf(l + 2, "%lf %lf %lf", &v.x, &v.y, &v.z) != 3) return *err = "bad coordinates in obj file", false;
        m->vert_pos.push_back(v);
      } else if (l[1] == 'n') {
        F3 v;
        if (std::sscanf(l + 2, "%f %f %f", &v.x, &v.y, &v.z) != 3) return *err = "bad normal in obj file", false;
        m->vert_norm.push_back(v);
      } else if (l[1] == 't') {
        F3 v;
        if (std::sscanf(l + 2, "%f %f", &v.x, &v.y) != 2) return *err = "bad texture coordinate in obj file", false;
        m->vert_tex.push_back(v);
      } else {
        ++m->unsupported;
      }
    } else if (l[0] == 'f') {
      if (l[1] != ' ') {
        ++m->unsupported;
        continue;
      }
      // convex polygons are split into a fan around their first vertex
      Model::Face face;
      int vert = 0, vert_cnt = 0;
      const size_t n = line.size();
      for (size_t i = 2;;) {
        while (i < n && l[i] == ' ') ++i;
        if (i >= n) break;
        size_t j = i;
        int chan = 0;
        while (j < n && l[j] != ' ') {
          while (j < n && l[j] == '/') ++chan, ++j;
          long long v = 0;
          while (j < n && l[j] >= '0' && l[j] <= '9') {
            v = v * 10 + (l[j] - '0');
            if (v > 0x7fffffff) return *err = "face vertex index out of range in obj file", false;
            ++j;
          }
          if (chan >= 3 || !v) return *err = "invalid face in obj file", false;
          (chan == 0 ? face.pos : chan == 1 ? face.tex : face.normal)[vert] = int(v);
        }
        i = j;
        ++vert;
        ++vert_cnt;
        if (vert == 3) {
          face.material = cur_mtl;
          for (int k = 0; k < 3; ++k) {
            if (face.pos[k] == 0) return *err = "no face vertex position in obj file", false;
            if (face.pos[k] >= int(m->vert_pos.size()))
              return *err = "face vertex position out of range in obj file", false;
            if (face.normal[k] >= int(m->vert_norm.size()))
              return *err = "face vertex normal out of range in obj file", false;
            if (face.tex[k] >= int(m->vert_tex.size()))
              return *err = "face vertex texture coordinate out of range in obj file", false;
          }
          m->faces.push_back(face);
          Model::Face nf;
          nf.pos[0] = face.pos[0], nf.tex[0] = face.tex[0], nf.normal[0] = face.normal[0];
          nf.pos[1] = face.pos[2], nf.tex[1] = face.tex[2], nf.normal[1] = face.normal[2];
          face = nf;
          vert = 2;
        }
      }
      if (vert_cnt < 3) return *err = "less than 3 vertices in a face in obj file", false;
    } else if (line.size() >= 7 && l[6] == ' ' && line.compare(0, 6, "mtllib") == 0) {
      if (!LoadMtl(CombinePaths(DirOf(path), line.substr(7)), m, err)) return false;
    } else if (line.size() >= 7 && l[6] == ' ' && line.compare(0, 6, "usemtl") == 0) {
      auto it = m->index.find(line.substr(7));
      if (it == m->index.end()) return *err = "usemtl with undefined name in obj file", false;
      cur_mtl = it->second;
    } else {
      ++m->unsupported;
    }
  }
  for (Model::Face& f : m->faces) {
    D3 tv[3];
    F3 tt[3];
    for (int k = 0; k < 3; ++k) tv[k] = m->vert_pos[f.pos[k]], tt[k] = m->vert_tex[f.tex[k]];
    float w0[3], wu[3], wv[3];
    ToBarycentric(F3(0, 0, 0), tt, w0, false);
    ToBarycentric(F3(1, 0, 0), tt, wu, false);
    ToBarycentric(F3(0, 1, 0), tt, wv, false);
    f.tangent_u = F3(tv[0] * double(wu[0] - w0[0]) + tv[1] * double(wu[1] - w0[1]) + tv[2] * double(wu[2] - w0[2]));
    f.tangent_v = F3(tv[0] * double(wv[0] - w0[0]) + tv[1] * double(wv[1] - w0[1]) + tv[2] * double(wv[2] - w0[2]));
    f.tangent_n = F3((tv[1] - tv[0]).Cross(tv[2] - tv[0]).Normalized());
  }
  return true;
}

// ------------------------------------------------------------------ node data
// The common tail of GetNodeData (import-obj-scene.cpp:627-635): a normal
// shorter than sqrt(1e-5) becomes +z, others are normalised.
void FinishNode(F3 n, const F3& c, float out[6]) {
  if (double(n.LengthSquare()) < 1e-5)
    n = F3(0, 0, 1);
  else
    n = n.Div(n.Length());
  out[0] = n.x, out[1] = n.y, out[2] = n.z, out[3] = c.x, out[4] = c.y, out[5] = c.z;
}

// GetNodeData of a node without children (import-obj-scene.cpp:580-612).
void LeafData(const Model& m, const D3& lo, const D3& hi, const uint32_t* tris, size_t n, float out[6]) {
  F3 normal(0, 0, 0), color(0, 0, 0);
  const D3 center = (lo + hi).Div(2.0);
  for (size_t q = 0; q < n; ++q) {
    const Model::Face& f = m.faces[tris[q]];
    const Material& mat = *m.materials[f.material];
    D3 tv[3];
    F3 tt[3];
    for (int k = 0; k < 3; ++k) tv[k] = m.vert_pos[f.pos[k]], tt[k] = m.vert_tex[f.tex[k]];
    double w[3];
    ToBarycentric(center, tv, w, true);
    const float fw[3] = {float(w[0]), float(w[1]), float(w[2])};
    const F3 tc = tt[0] * fw[0] + tt[1] * fw[1] + tt[2] * fw[2];
    if (mat.normal.texture.valid) {
      const F3 tn = mat.normal.NormalAt(tc.x, tc.y);
      normal = normal + (f.tangent_u * tn.x + f.tangent_v * tn.y + f.tangent_n * tn.z).Normalized();
    } else if (f.normal[0] && f.normal[1] && f.normal[2]) {
      normal = normal + (m.vert_norm[f.normal[0]] * fw[0] + m.vert_norm[f.normal[1]] * fw[1] +
                         m.vert_norm[f.normal[2]] * fw[2]).Normalized();
    } else {
      normal = normal + f.tangent_n;
    }
    color = color + (mat.ambient.ColorAt(tc.x, tc.y) + mat.diffuse.ColorAt(tc.x, tc.y));
  }
  color = color.Div(float(n));
  FinishNode(normal, color, out);
}

// (char)(x * 127) as x86 converts: through int32, NaN and overflow to INT32_MIN.
inline uint8_t NormalByte(float x) {
  const float v = x * 127;
  const int32_t i = (v > -2147483904.0f && v < 2147483648.0f) ? int32_t(v) : INT32_MIN;
  return uint8_t(i & 0xff);
}

// NodeData::GetChannelsData (import-obj-scene.cpp:516-526).
inline void PackChannels(const float nc[6], uint8_t* out) {
  for (int a = 0; a < 3; ++a) out[a] = NormalByte(nc[a]);
  out[3] = 0;
  for (int a = 0; a < 3; ++a) {
    float v = nc[3 + a] * 255;
    v = v < 255.0f ? v : 255.0f;
    v = 0.0f < v ? v : 0.0f;
    out[4 + a] = uint8_t(v);
  }
  out[7] = 255;
}

// One level of the tree under construction, in BFS order.
struct Level {
  std::vector<D3> lo, hi;
  std::vector<uint64_t> tri_begin{0};  // node k's triangles: tris[tri_begin[k], tri_begin[k + 1])
  std::vector<uint32_t> tris;
};

template <class F>
void ParallelChunks(int threads, size_t chunks, F&& f) {
  if (threads <= 1 || chunks <= 1) {
    for (size_t c = 0; c < chunks; ++c) f(c);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < threads && size_t(t) < chunks; ++t)
    th.emplace_back([&] {
      for (size_t c; (c = next.fetch_add(1)) < chunks;) f(c);
    });
  for (auto& x : th) x.join();
}

}  // namespace

bool ImportObjTree(const std::string& obj_path, int max_level, int threads, NodeTree* out, ObjImportReport* report,
                   std::string* err) {
  if (max_level < 0 || max_level > 20) return *err = "max_level must be in [0, 20]", false;
  Model m;
  if (!ReadObj(obj_path, &m, err)) return false;
  if (m.faces.empty()) return *err = "obj file has no faces", false;
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  const bool log = std::getenv("RAYT_BUILD_LOG") != nullptr;

  const size_t nf = m.faces.size();
  std::vector<FaceGeom> geom(nf);
  // BoundingCube (import-obj-scene.cpp:481-493)
  D3 bmin(1e30, 1e30, 1e30), bmax = D3(0, 0, 0) - bmin;
  for (size_t i = 0; i < nf; ++i) {
    FaceGeom& g = geom[i];
    for (int k = 0; k < 3; ++k) {
      const D3& p = m.vert_pos[m.faces[i].pos[k]];
      bmax = bmax.Max(p);
      bmin = bmin.Min(p);
      g.v[k] = F3(p);
    }
    g.bmin = g.v[0].Min(g.v[1]).Min(g.v[2]);
    g.bmax = g.v[0].Max(g.v[1]).Max(g.v[2]);
  }
  const D3 ext = bmax - bmin;
  const double s = ext.x > ext.y ? (ext.x > ext.z ? ext.x : ext.z) : (ext.y > ext.z ? ext.y : ext.z);
  bmax = bmin + D3(s, s, s);

  // Level-synchronous expansion (GetChildren, import-obj-scene.cpp:548-577);
  // leaf data is computed while the leaf's triangles are at hand.
  std::vector<std::vector<uint8_t>> masks(max_level + 1);
  std::vector<std::vector<float>> nc(max_level + 1);  // normal xyz + colour rgb per node
  Level cur;
  cur.lo.push_back(bmin);
  cur.hi.push_back(bmax);
  cur.tris.resize(nf);
  for (size_t i = 0; i < nf; ++i) cur.tris[i] = uint32_t(i);
  cur.tri_begin.push_back(nf);
  uint64_t refs = nf;
  constexpr size_t kChunk = 64;
  for (int L = 0; L <= max_level; ++L) {
    const size_t n = cur.lo.size();
    masks[L].assign(n, 0);
    nc[L].assign(n * 6, 0.0f);
    const bool last = L == max_level;
    const size_t chunks = (n + kChunk - 1) / kChunk;
    std::vector<Level> parts(last ? 0 : chunks);
    ParallelChunks(threads, chunks, [&](size_t c) {
      const size_t b = c * kChunk, e = std::min(n, b + kChunk);
      Level* part = last ? nullptr : &parts[c];
      std::vector<uint32_t> kept;
      for (size_t x = b; x < e; ++x) {
        const uint32_t* tris = cur.tris.data() + cur.tri_begin[x];
        const size_t cnt = size_t(cur.tri_begin[x + 1] - cur.tri_begin[x]);
        uint8_t mask = 0;
        if (!last) {
          const D3& lo = cur.lo[x];
          const D3& hi = cur.hi[x];
          const D3 mid = (lo + hi).Div(2.0);
          for (int ci = 0; ci < 8; ++ci) {
            D3 cl = lo, ch = hi;
            if (ci & 1) cl.x = mid.x; else ch.x = mid.x;
            if (ci & 2) cl.y = mid.y; else ch.y = mid.y;
            if (ci & 4) cl.z = mid.z; else ch.z = mid.z;
            const F3 fl(cl), fh(ch);
            kept.clear();
            for (size_t q = 0; q < cnt; ++q)
              if (TriangleBoxIntersection(fl, fh, geom[tris[q]])) kept.push_back(tris[q]);
            if (kept.empty()) continue;
            mask |= uint8_t(1u << ci);
            part->lo.push_back(cl);
            part->hi.push_back(ch);
            part->tris.insert(part->tris.end(), kept.begin(), kept.end());
            part->tri_begin.push_back(part->tris.size());
          }
        }
        masks[L][x] = mask;
        if (!mask) LeafData(m, cur.lo[x], cur.hi[x], tris, cnt, &nc[L][x * 6]);
      }
    });
    if (last) break;
    Level next;
    size_t nodes = 0, total = 0;
    for (const Level& p : parts) nodes += p.lo.size(), total += p.tris.size();
    next.lo.reserve(nodes);
    next.hi.reserve(nodes);
    next.tri_begin.reserve(nodes + 1);
    next.tris.reserve(total);
    for (Level& p : parts) {
      const uint64_t base = next.tris.size();
      next.lo.insert(next.lo.end(), p.lo.begin(), p.lo.end());
      next.hi.insert(next.hi.end(), p.hi.begin(), p.hi.end());
      for (size_t k = 1; k < p.tri_begin.size(); ++k) next.tri_begin.push_back(base + p.tri_begin[k]);
      next.tris.insert(next.tris.end(), p.tris.begin(), p.tris.end());
      p = Level();
    }
    refs += total;
    cur = std::move(next);
    if (log) {
      std::fprintf(stderr, "[rayt obj] level %d: %zu nodes, %zu triangle refs\n", L + 1, nodes, total);
      std::fflush(stderr);
    }
    if (nodes == 0) {
      masks.resize(L + 1);
      nc.resize(L + 1);
      break;
    }
  }
  const int depth = int(masks.size()) - 1;

  // Flatten and compute inner node data bottom-up (children in bit order).
  std::vector<uint64_t> level_off(depth + 2, 0);
  for (int L = 0; L <= depth; ++L) level_off[L + 1] = level_off[L] + masks[L].size();
  const uint64_t total = level_off[depth + 1];
  if (total > 0xffffffffull) return *err = "tree has more than 2^32 nodes", false;
  NodeTree t;
  t.max_depth = max_level;
  t.channels = {{4, "normals"}, {4, "colors"}};
  t.data_size = 8;
  t.mask.resize(total);
  t.first_child.resize(total);
  t.data.resize(total * 8);
  for (int L = 0; L <= depth; ++L) {
    uint64_t next = level_off[L + 1];
    for (size_t x = 0; x < masks[L].size(); ++x) {
      const uint8_t mk = masks[L][x];
      t.mask[level_off[L] + x] = mk;
      t.first_child[level_off[L] + x] = mk ? uint32_t(next) : 0;
      next += __builtin_popcount(mk);
    }
  }
  for (int L = depth; L >= 0; --L) {
    const size_t n = masks[L].size();
    const size_t chunks = (n + 4095) / 4096;
    ParallelChunks(threads, chunks, [&](size_t c) {
      for (size_t x = c * 4096, e = std::min(n, x + 4096); x < e; ++x) {
        const uint64_t node = level_off[L] + x;
        float* o = &nc[L][x * 6];
        const uint8_t mk = masks[L][x];
        if (mk) {
          const uint64_t c0 = t.first_child[node] - level_off[L + 1];
          const int k = __builtin_popcount(mk);
          F3 sn(0, 0, 0), sc(0, 0, 0);
          for (int q = 0; q < k; ++q) {
            const float* cc = &nc[L + 1][(c0 + q) * 6];
            sn = sn + F3(cc[0], cc[1], cc[2]);
            sc = sc + F3(cc[3], cc[4], cc[5]);
          }
          FinishNode(sn, sc.Div(float(k)), o);
        }
        PackChannels(o, &t.data[node * 8]);
      }
    });
    if (L + 1 <= depth) std::vector<float>().swap(nc[L + 1]);
  }
  *out = std::move(t);
  if (report) {
    report->faces = nf;
    report->materials = m.materials.size();
    report->unsupported = m.unsupported;
    report->triangle_refs = refs;
  }
  return true;
}

}  // namespace rayt
