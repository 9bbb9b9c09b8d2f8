// ORACLE - test infrastructure only.
//
// A CPU restatement of the reference's OBJ scene source, used by tests/ as
// the checker for the product importer (rayt_amd/csrc/obj_import.cpp).
// Nothing in the product links or calls it. Built with g++ -O2
// -ffp-contract=off (SSE2, IEEE fp32/fp64, correctly rounded division and
// sqrt).
//
// Restated, and from where (reference paths):
//   ReadObjFile + CalculateTangentSpaces  scene sources/import-obj-scene.cpp:301-479
//   MaterialLibrary::Load                 import-obj-scene.cpp:153-280
//   Texture / ColorMap / NormalMap        import-obj-scene.cpp:32-142
//   BoundingCube, GetRoot, GetChildren,   import-obj-scene.cpp:481-638
//   GetNodeData, GetChannelsData
//   TriangleBoxIntersection               util/intersections.cpp:7-59
//   tvec3 arithmetic, ToBarycentric       util/vec.h:9-74
//   CImg<float>::_linear_atXY, load_pnm   libs/CImg.h:11725-11742, 33979-34140 (CImg 1.4.9)
//   RemoveFileNameFromPath, CombinePaths  util/string-util.cpp:52-66
//
// Unlike the product (level by level over threads), this follows the
// reference's shape: a depth-first recursion over nodes that own their
// triangle lists, node data computed after the children. The MTL reader uses
// std::stringstream and the OBJ reader std::istream::getline into one
// reused 2048-byte buffer, as the reference does.
//
// Fixed where the reference leaves it open: `vt` z is 0 (the reference does
// not set it); (char) of a NaN normal component is 0 (x86's conversion
// through int32); NaN texture coordinates sample column / row 0.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc_obj {

template <typename T>
struct vec3 {
  T x, y, z;
  vec3() : x(0), y(0), z(0) {}
  vec3(T a, T b, T c) : x(a), y(b), z(c) {}
  template <typename U>
  vec3(const vec3<U>& o) : x(static_cast<T>(o.x)), y(static_cast<T>(o.y)), z(static_cast<T>(o.z)) {}
  vec3 operator+(const vec3& o) const { return vec3(x + o.x, y + o.y, z + o.z); }
  vec3 operator-(const vec3& o) const { return vec3(x - o.x, y - o.y, z - o.z); }
  vec3 operator*(T d) const { return vec3(x * d, y * d, z * d); }
  vec3 operator*(const vec3& o) const { return vec3(x * o.x, y * o.y, z * o.z); }
  vec3 operator/(T d) const {
    d = 1.0f / d;
    return vec3(x * d, y * d, z * d);
  }
  vec3& operator+=(const vec3& o) {
    x += o.x;
    y += o.y;
    z += o.z;
    return *this;
  }
  vec3& operator/=(T d) {
    d = 1.0f / d;
    x *= d;
    y *= d;
    z *= d;
    return *this;
  }
  T dot(const vec3& o) const { return x * o.x + y * o.y + z * o.z; }
  vec3 cross(const vec3& o) const { return vec3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
  T len2() const { return dot(*this); }
  T len() const { return std::sqrt(len2()); }
  vec3 normalized() const { return *this / len(); }
  bool all_gt(const vec3& o) const { return x > o.x && y > o.y && z > o.z; }
  bool all_lt(const vec3& o) const { return x < o.x && y < o.y && z < o.z; }
  static T mn(T a, T b) { return (b < a) ? b : a; }  // std::min
  static T mx(T a, T b) { return (a < b) ? b : a; }  // std::max
  vec3 vmin(const vec3& o) const { return vec3(mn(x, o.x), mn(y, o.y), mn(z, o.z)); }
  vec3 vmax(const vec3& o) const { return vec3(mx(x, o.x), mx(y, o.y), mx(z, o.z)); }
  void barycentric(const vec3* t, T* out, bool clamp) const {
    vec3 n = (t[1] - t[0]).cross(t[2] - t[0]);
    T sum = 0;
    for (int i = 0; i < 3; ++i) {
      out[i] = (t[(i + 2) % 3] - t[(i + 1) % 3]).cross(*this - t[(i + 1) % 3]).dot(n);
      if (clamp && out[i] < 0) out[i] = 0;
      sum += out[i];
    }
    if (std::fabs(static_cast<double>(sum)) >= 1e-5) {
      for (int i = 0; i < 3; ++i) out[i] /= sum;
    } else {
      out[0] = out[1] = out[2] = static_cast<T>(1. / 3);
    }
  }
};
typedef vec3<float> fvec3;
typedef vec3<double> dvec3;

static void fail(const std::string& m) { throw std::runtime_error(m); }

// ----------------------------------------------------------- intersections
static bool separated(const fvec3& axis, const fvec3& bmin, const fvec3& bmax, const fvec3& a, const fvec3& b,
                      const fvec3& c) {
  float lo = fvec3::mn(fvec3::mn(axis.dot(a), axis.dot(b)), axis.dot(c));
  float hi = fvec3::mx(fvec3::mx(axis.dot(a), axis.dot(b)), axis.dot(c));
  fvec3 p = bmin, q = bmax;
  if (axis.x < 0) std::swap(p.x, q.x);
  if (axis.y < 0) std::swap(p.y, q.y);
  if (axis.z < 0) std::swap(p.z, q.z);
  float pb = axis.dot(p), qb = axis.dot(q);
  return pb > hi || lo > qb;
}

static bool tri_box(const fvec3& bmin, const fvec3& bmax, const fvec3& a, const fvec3& b, const fvec3& c) {
  fvec3 tmin = a.vmin(b).vmin(c), tmax = a.vmax(b).vmax(c);
  if (!tmin.all_lt(bmax) || !tmax.all_gt(bmin)) return false;
  const fvec3* v[3] = {&a, &b, &c};
  for (int k = 0; k < 3; ++k)
    if (v[k]->all_gt(bmin) && v[k]->all_lt(bmax)) return true;
  if (separated((b - a).cross(c - a), bmin, bmax, a, b, c)) return false;
  fvec3 box_axes[3] = {fvec3(0, 0, 1), fvec3(0, 1, 0), fvec3(1, 0, 0)};
  fvec3 edges[3] = {b - a, c - b, a - c};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (separated(box_axes[i].cross(edges[j]), bmin, bmax, a, b, c)) return false;
  return true;
}

// ------------------------------------------------------------------ images
struct Image {
  unsigned w = 0, h = 0, c = 0;
  std::vector<float> v;
  float at(unsigned x, unsigned y, unsigned ch) const { return v[x + y * w + ch * w * h]; }
};

static float linear_at(const Image& im, float fx, float fy, unsigned ch) {
  float nfx = fx < 0 ? 0 : (fx > im.w - 1 ? im.w - 1 : fx);
  float nfy = fy < 0 ? 0 : (fy > im.h - 1 ? im.h - 1 : fy);
  if (std::isnan(nfx)) nfx = 0;
  if (std::isnan(nfy)) nfy = 0;
  unsigned x = (unsigned)nfx, y = (unsigned)nfy;
  float dx = nfx - x, dy = nfy - y;
  unsigned nx = dx > 0 ? x + 1 : x, ny = dy > 0 ? y + 1 : y;
  float Icc = im.at(x, y, ch), Inc = im.at(nx, y, ch), Icn = im.at(x, ny, ch), Inn = im.at(nx, ny, ch);
  return Icc + dx * (Inc - Icc + dy * (Icc + Inn - Icn - Inc)) + dy * (Icn - Icc);
}

// load_pnm with FILE* as CImg reads it (P2, P3, P5, P6).
static void load_pnm(const std::string& fn, Image* im) {
  FILE* f = std::fopen(fn.c_str(), "rb");
  if (!f) fail("cannot open texture " + fn);
  char item[1024] = {0};
  int err;
  unsigned type = 0, W = 0, H = 0, D = 1, cmax = 255;
  while ((err = std::fscanf(f, "%1023[^\n]", item)) != EOF && (*item == '#' || !err)) std::fgetc(f);
  if (std::sscanf(item, " P%u", &type) != 1) {
    std::fclose(f);
    fail("PNM header not found");
  }
  while ((err = std::fscanf(f, " %1023[^\n]", item)) != EOF && (*item == '#' || !err)) std::fgetc(f);
  if ((err = std::sscanf(item, " %u %u %u %u", &W, &H, &D, &cmax)) < 2) {
    std::fclose(f);
    fail("PNM size missing");
  }
  if (err == 2) {
    while ((err = std::fscanf(f, " %1023[^\n]", item)) != EOF && (*item == '#' || !err)) std::fgetc(f);
    if (std::sscanf(item, "%u", &cmax) != 1) cmax = 255;
  } else if (err == 3 && type != 5) {
    cmax = D;
  } else {
    std::fclose(f);
    fail("3-D PNM");
  }
  std::fgetc(f);
  if (type != 2 && type != 3 && type != 5 && type != 6) {
    std::fclose(f);
    fail("PNM type");
  }
  im->w = W;
  im->h = H;
  im->c = (type == 3 || type == 6) ? 3 : 1;
  im->v.assign(size_t(W) * H * im->c, 0.0f);
  const size_t plane = size_t(W) * H;
  bool ok = true;
  for (size_t i = 0; i < plane && ok; ++i) {
    for (unsigned ch = 0; ch < im->c && ok; ++ch) {
      float val = 0;
      if (type == 2 || type == 3) {
        int r;
        ok = std::fscanf(f, "%d", &r) == 1;
        val = (float)r;
      } else if (cmax < 256) {
        int b = std::fgetc(f);
        ok = b != EOF;
        val = (float)b;
      } else {
        int hi = std::fgetc(f), lo = std::fgetc(f);
        ok = lo != EOF;
        val = (float)((hi << 8) | lo);
      }
      im->v[i + ch * plane] = val;
    }
  }
  std::fclose(f);
  if (!ok) fail("PNM samples missing");
}

static void load_image(const std::string& fn, Image* im) {
  std::string ext = fn.substr(fn.find_last_of('.') == std::string::npos ? fn.size() : fn.find_last_of('.') + 1);
  for (char& ch : ext) ch = (char)std::tolower((unsigned char)ch);
  if (ext != "ppm" && ext != "pgm" && ext != "pnm") fail("texture format");
  load_pnm(fn, im);
}

struct Texture {
  bool valid = false;
  Image image;
  void Load(const std::string& fn) {
    load_image(fn, &image);
    valid = image.w > 0 && image.h > 0;
  }
  fvec3 sample3(float u, float v, bool third) const {
    u -= std::floor(u);
    v -= std::floor(v);
    u *= image.w;
    v *= image.h;
    return fvec3(linear_at(image, u, v, 0), linear_at(image, u, v, 1), third ? linear_at(image, u, v, 2) : 0) /
           255.;
  }
  fvec3 ColorAt(float u, float v) const { return valid ? sample3(u, v, true) : fvec3(1, 1, 1); }
  fvec3 XyAt(float u, float v) const { return valid ? sample3(u, v, false) : fvec3(1, 1, 1); }
};

struct ColorMap {
  fvec3 multiplier;
  Texture texture;
  fvec3 ColorAt(float u, float v) const { return texture.valid ? texture.ColorAt(u, v) * multiplier : multiplier; }
};
struct ScalarMap {
  float multiplier = 0;
  Texture texture;
};
struct NormalMap {
  Texture texture;
  void Build(const ScalarMap& bump) {
    const Image& b = bump.texture.image;
    int w = int(b.w), h = int(b.h);
    texture.valid = true;
    texture.image.w = w;
    texture.image.h = h;
    texture.image.c = 2;
    texture.image.v.assign(size_t(w) * h * 2, 0.0f);
    for (int i = 0; i < h; ++i)
      for (int j = 0; j < w; ++j) {
        texture.image.v[j + i * w] = (b.at((j + w - 1) % w, i, 0) - b.at((j + 1) % w, i, 0)) * bump.multiplier / 2;
        texture.image.v[j + i * w + w * h] =
            (b.at(j, (i + h - 1) % h, 0) - b.at(j, (i + 1) % h, 0)) * bump.multiplier / 2;
      }
  }
  fvec3 NormalAt(float u, float v) const {
    if (!texture.valid) return fvec3(0, 0, 1);
    fvec3 t = texture.XyAt(u, v);
    t.z = 1;
    return t;
  }
};
struct Material {
  ColorMap ambient, diffuse, specular, transmission_filter;
  ScalarMap specular_exponent, alpha, optical_density, bump;
  NormalMap normal;
};

static std::string dir_of(const std::string& s) {
  std::string::size_type p = s.find_last_of("/\\");
  return p == std::string::npos ? std::string() : s.substr(0, p + 1);
}
static std::string join(std::string a, const std::string& b) {
  if (a.empty()) return b;
  if (a[a.size() - 1] != '/' && a[a.size() - 1] != '\\') a += "/";
  return a + b;
}

struct Library {
  std::map<std::string, int> index;
  std::vector<std::shared_ptr<Material>> materials;
  Library() {
    index[""] = 0;
    auto m = std::make_shared<Material>();
    materials.push_back(m);
    m->diffuse.multiplier = fvec3(1, 1, 1);
    m->ambient.multiplier = fvec3(.1, .1, .1);
    m->alpha.multiplier = 1;
  }
  void Load(const std::string& fn) {
    std::ifstream in(fn.c_str());
    static char line[2048];
    std::string prefix = dir_of(fn);
    std::shared_ptr<Material> cur = materials[0];
    while (in.getline(line, sizeof(line))) {
      if (line[0] == '#') continue;
      std::stringstream ss(line);
      std::string s;
      ss >> s;
      if (s == "newmtl") {
        std::string name;
        ss >> name;
        if (name.empty()) fail("newmtl without name");
        if (index.count(name)) fail("duplicate material name");
        index[name] = int(materials.size());
        cur = std::make_shared<Material>();
        materials.push_back(cur);
        continue;
      }
      bool is_map = false;
      std::string map_file;
      double bump_mult = 1;
      if (s == "bump" || s.substr(0, 4) == "map_") {
        is_map = true;
        if (s.substr(0, 4) == "map_") s = s.substr(4);
        std::string tok;
        while (ss >> tok) {
          if (tok[0] == '-') {
            if (tok == "-bm") ss >> bump_mult;
          } else {
            map_file = tok;
          }
        }
        map_file = prefix + map_file;
      }
      if (s == "Ns" || s == "d" || s == "Ni") {
        ScalarMap* m = s == "Ns" ? &cur->specular_exponent : (s == "Ni" ? &cur->optical_density : &cur->alpha);
        if (is_map) {
          m->texture.Load(map_file);
        } else {
          float val;
          if (!(ss >> val)) fail("invalid value after " + s);
          m->multiplier = val;
        }
      } else if (s == "bump") {
        cur->bump.texture.Load(map_file);
        cur->bump.multiplier = (float)bump_mult;
        cur->normal.Build(cur->bump);
      } else if (s == "Ka" || s == "Kd" || s == "Ks" || s == "Tf") {
        ColorMap* m = s == "Ka" ? &cur->ambient
                      : s == "Kd" ? &cur->diffuse
                      : s == "Ks" ? &cur->specular
                                  : &cur->transmission_filter;
        if (is_map) {
          m->texture.Load(map_file);
          if ((s == "Ka" || s == "Kd") && m->texture.image.c < 3) fail("grey colour texture");
        } else {
          fvec3 val;
          if (!(ss >> val.x >> val.y >> val.z)) fail("invalid value after " + s);
          m->multiplier = val;
        }
      }
    }
  }
};

struct Face {
  int pos[3] = {0, 0, 0}, tex[3] = {0, 0, 0}, nrm[3] = {0, 0, 0};
  int material = 0;
  fvec3 tu, tv, tn;
};

struct Model {
  Library lib;
  std::vector<dvec3> pos;
  std::vector<fvec3> tex, nrm;
  std::vector<Face> faces;
};

static void read_obj(const std::string& fn, Model& m) {
  std::ifstream in(fn.c_str());
  if (!in) fail("failed to open obj file");
  m.pos.push_back(dvec3(0, 0, 0));
  m.nrm.push_back(fvec3(0, 0, 0));
  m.tex.push_back(fvec3(0, 0, 0));
  int cur = 0;
  static char line[2048];
  std::memset(line, 0, sizeof(line));
  while (in.getline(line, sizeof(line))) {
    if (line[0] == 'v') {
      if (line[1] == ' ') {
        dvec3 v;
        if (std::sscanf(line + 2, "%lf %lf %lf", &v.x, &v.y, &v.z) != 3) fail("bad coordinates");
        m.pos.push_back(v);
      } else if (line[1] == 'n') {
        fvec3 v;
        if (std::sscanf(line + 2, "%f %f %f", &v.x, &v.y, &v.z) != 3) fail("bad normal");
        m.nrm.push_back(v);
      } else if (line[1] == 't') {
        fvec3 v;
        if (std::sscanf(line + 2, "%f %f", &v.x, &v.y) != 2) fail("bad texture coordinate");
        v.z = 0;
        m.tex.push_back(v);
      }
    } else if (line[0] == 'f') {
      if (line[1] != ' ') continue;
      Face face;
      int vert = 0, count = 0;
      int i = 2;
      for (;;) {
        while (line[i] == ' ') ++i;
        if (!line[i]) break;
        int j = i, chan = 0;
        while (line[j] && line[j] != ' ') {
          while (line[j] == '/') {
            ++chan;
            ++j;
          }
          long long v = 0;
          while (line[j] >= '0' && line[j] <= '9') {
            v = v * 10 + (line[j] - '0');
            if (v > 2147483647LL) fail("index overflow");
            ++j;
          }
          if (chan >= 3 || !v) fail("invalid face");
          if (chan == 0) face.pos[vert] = int(v);
          if (chan == 1) face.tex[vert] = int(v);
          if (chan == 2) face.nrm[vert] = int(v);
        }
        i = j;
        ++vert;
        ++count;
        if (vert == 3) {
          face.material = cur;
          for (int k = 0; k < 3; ++k) {
            if (face.pos[k] == 0) fail("no face vertex position");
            if (face.pos[k] >= int(m.pos.size())) fail("position out of range");
            if (face.nrm[k] >= int(m.nrm.size())) fail("normal out of range");
            if (face.tex[k] >= int(m.tex.size())) fail("texture coordinate out of range");
          }
          m.faces.push_back(face);
          Face nf;
          nf.pos[0] = face.pos[0];
          nf.tex[0] = face.tex[0];
          nf.nrm[0] = face.nrm[0];
          nf.pos[1] = face.pos[2];
          nf.tex[1] = face.tex[2];
          nf.nrm[1] = face.nrm[2];
          face = nf;
          vert = 2;
        }
      }
      if (count < 3) fail("less than 3 vertices in a face");
    } else if (line[6] == ' ') {
      line[6] = 0;
      if (!std::strcmp(line, "mtllib")) {
        m.lib.Load(join(dir_of(fn), std::string(line + 7)));
      } else if (!std::strcmp(line, "usemtl")) {
        std::string name = line + 7;
        if (!m.lib.index.count(name)) fail("usemtl with undefined name");
        cur = m.lib.index[name];
      }
      line[6] = ' ';
    }
  }
  for (Face& f : m.faces) {
    dvec3 tv[3] = {m.pos[f.pos[0]], m.pos[f.pos[1]], m.pos[f.pos[2]]};
    fvec3 tt[3] = {m.tex[f.tex[0]], m.tex[f.tex[1]], m.tex[f.tex[2]]};
    float w0[3], wu[3], wv[3];
    fvec3(0, 0, 0).barycentric(tt, w0, false);
    fvec3(1, 0, 0).barycentric(tt, wu, false);
    fvec3(0, 1, 0).barycentric(tt, wv, false);
    f.tu = tv[0] * (wu[0] - w0[0]) + tv[1] * (wu[1] - w0[1]) + tv[2] * (wu[2] - w0[2]);
    f.tv = tv[0] * (wv[0] - w0[0]) + tv[1] * (wv[1] - w0[1]) + tv[2] * (wv[2] - w0[2]);
    f.tn = (tv[1] - tv[0]).cross(tv[2] - tv[0]).normalized();
  }
}

// ------------------------------------------------------------------ the tree
struct Node {
  dvec3 lo, hi;
  int depth = 0;
  std::vector<int> tris;
  fvec3 normal, color;
  int mask = 0;
  std::vector<std::unique_ptr<Node>> kids;
};

struct Source {
  Model m;
  int max_depth = 0;

  void children(Node& n) {
    if (n.depth == max_depth) return;
    for (int ci = 0; ci < 8; ++ci) {
      dvec3 lo = n.lo, hi = n.hi, mid = (n.lo + n.hi) / 2;
      if (ci & 1) lo.x = mid.x; else hi.x = mid.x;
      if (ci & 2) lo.y = mid.y; else hi.y = mid.y;
      if (ci & 4) lo.z = mid.z; else hi.z = mid.z;
      std::unique_ptr<Node> k(new Node());
      k->lo = lo;
      k->hi = hi;
      k->depth = n.depth + 1;
      for (int t : n.tris) {
        const Face& f = m.faces[t];
        if (tri_box(fvec3(lo), fvec3(hi), fvec3(m.pos[f.pos[0]]), fvec3(m.pos[f.pos[1]]), fvec3(m.pos[f.pos[2]])))
          k->tris.push_back(t);
      }
      if (k->tris.empty()) continue;
      n.mask |= 1 << ci;
      n.kids.push_back(std::move(k));
    }
  }

  void node_data(Node& n) {
    n.normal = fvec3(0, 0, 0);
    n.color = fvec3(0, 0, 0);
    if (n.kids.empty()) {
      for (int t : n.tris) {
        const Face& f = m.faces[t];
        const Material& mat = *m.lib.materials[f.material];
        dvec3 tv[3] = {m.pos[f.pos[0]], m.pos[f.pos[1]], m.pos[f.pos[2]]};
        double w[3];
        ((n.lo + n.hi) / 2).barycentric(tv, w, true);
        float fw[3] = {(float)w[0], (float)w[1], (float)w[2]};
        fvec3 tt[3] = {m.tex[f.tex[0]], m.tex[f.tex[1]], m.tex[f.tex[2]]};
        fvec3 tc = tt[0] * fw[0] + tt[1] * fw[1] + tt[2] * fw[2];
        if (mat.normal.texture.valid) {
          fvec3 tn = mat.normal.NormalAt(tc.x, tc.y);
          n.normal += (f.tu * tn.x + f.tv * tn.y + f.tn * tn.z).normalized();
        } else if (f.nrm[0] && f.nrm[1] && f.nrm[2]) {
          n.normal += (m.nrm[f.nrm[0]] * fw[0] + m.nrm[f.nrm[1]] * fw[1] + m.nrm[f.nrm[2]] * fw[2]).normalized();
        } else {
          n.normal += f.tn;
        }
        n.color += mat.ambient.ColorAt(tc.x, tc.y) + mat.diffuse.ColorAt(tc.x, tc.y);
      }
      n.color /= (float)n.tris.size();
    } else {
      for (auto& k : n.kids) {
        n.normal += k->normal;
        n.color += k->color;
      }
      n.color /= (float)n.kids.size();
    }
    if (n.normal.len2() < 1e-5)
      n.normal = fvec3(0, 0, 1);
    else
      n.normal /= n.normal.len();
  }

  void expand(Node& n) {  // GetChildren, recursion, then GetNodeData
    children(n);
    for (auto& k : n.kids) expand(*k);
    node_data(n);
  }
};

static uint8_t char_of(float v) {
  int32_t i = (v != v || v >= 2147483648.0f || v <= -2147483904.0f) ? INT32_MIN : (int32_t)v;
  return (uint8_t)(i & 0xff);
}
static uint8_t uchar_of(float v) {
  float c = fvec3::mx(0.0f, fvec3::mn(255.0f, v));
  return (uint8_t)c;
}

static std::vector<uint8_t> g_masks, g_data;
static std::string g_error;

}  // namespace orc_obj

extern "C" {

// TriangleBoxIntersection on float boxes and vertices (3 floats each).
int orc_tri_box(const float* bmin, const float* bmax, const float* a, const float* b, const float* c) {
  using orc_obj::fvec3;
  return orc_obj::tri_box(fvec3(bmin[0], bmin[1], bmin[2]), fvec3(bmax[0], bmax[1], bmax[2]),
                          fvec3(a[0], a[1], a[2]), fvec3(b[0], b[1], b[2]), fvec3(c[0], c[1], c[2]));
}

// CImg's _linear_atXY on a one-channel w x h image (row-major).
float orc_linear_at(const float* img, int w, int h, float fx, float fy) {
  orc_obj::Image im;
  im.w = unsigned(w);
  im.h = unsigned(h);
  im.c = 1;
  im.v.assign(img, img + size_t(w) * h);
  return orc_obj::linear_at(im, fx, fy, 0);
}

// Imports an OBJ to depth max_level. Returns the node count (then
// orc_import_obj_fetch copies the BFS masks and 8-byte node data) or -1 with
// orc_import_obj_error() set.
int64_t orc_import_obj(const char* path, int max_level) {
  using namespace orc_obj;
  g_masks.clear();
  g_data.clear();
  try {
    Source src;
    src.max_depth = max_level;
    read_obj(path, src.m);
    if (src.m.faces.empty()) fail("obj file has no faces");
    dvec3 bmin(1e30, 1e30, 1e30), bmax = dvec3(0, 0, 0) - bmin;
    for (const Face& f : src.m.faces)
      for (int k = 0; k < 3; ++k) {
        bmax = bmax.vmax(src.m.pos[f.pos[k]]);
        bmin = bmin.vmin(src.m.pos[f.pos[k]]);
      }
    dvec3 e = bmax - bmin;
    double s = e.x > e.y ? (e.x > e.z ? e.x : e.z) : (e.y > e.z ? e.y : e.z);
    bmax = bmin + dvec3(s, s, s);
    Node root;
    root.lo = bmin;
    root.hi = bmax;
    for (int i = 0; i < int(src.m.faces.size()); ++i) root.tris.push_back(i);
    src.expand(root);
    std::vector<const Node*> q{&root};
    for (size_t i = 0; i < q.size(); ++i) {
      const Node* n = q[i];
      g_masks.push_back(uint8_t(n->mask));
      g_data.push_back(char_of(n->normal.x * 127));
      g_data.push_back(char_of(n->normal.y * 127));
      g_data.push_back(char_of(n->normal.z * 127));
      g_data.push_back(0);
      g_data.push_back(uchar_of(n->color.x * 255));
      g_data.push_back(uchar_of(n->color.y * 255));
      g_data.push_back(uchar_of(n->color.z * 255));
      g_data.push_back(255);
      for (auto& k : n->kids) q.push_back(k.get());
    }
    return int64_t(g_masks.size());
  } catch (const std::exception& ex) {
    g_error = ex.what();
    return -1;
  }
}

void orc_import_obj_fetch(uint8_t* masks, uint8_t* data) {
  std::memcpy(masks, orc_obj::g_masks.data(), orc_obj::g_masks.size());
  std::memcpy(data, orc_obj::g_data.data(), orc_obj::g_data.size());
}

const char* orc_import_obj_error() { return orc_obj::g_error.c_str(); }

}  // extern "C"
