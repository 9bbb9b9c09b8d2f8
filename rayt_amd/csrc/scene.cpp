// Procedural octree sources (see scene.h).
#include "scene.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <thread>

namespace rayt {

namespace {

// ---------------------------------------------------------------- sphere test
struct F3 {
  float x, y, z;
};

static inline float DistSq(const F3& a, const F3& c) {
  // fvec3::DistanceSquare: (b - a).LengthSquare() = (dx*dx + dy*dy) + dz*dz
  float dx = a.x - c.x, dy = a.y - c.y, dz = a.z - c.z;
  return dx * dx + dy * dy + dz * dz;
}

struct SphereNode {
  F3 minp, maxp;
  int32_t node_id;
};

}  // namespace

NodeTree GenerateSphereTree(float cx, float cy, float cz, float radius, int level) {
  NodeTree t;
  t.max_depth = level;
  t.channels = {{4, "node index"}, {1, "node index % 211"}};
  t.data_size = 5;
  const F3 center{cx, cy, cz};
  const float radius_square = radius * radius;

  std::vector<SphereNode> cur{{{0, 0, 0}, {1, 1, 1}, 1}}, next;
  uint64_t level_first = 0;  // index of cur[0] in the tree
  auto emit_data = [&](int32_t node_id) {
    size_t p = t.data.size();
    t.data.resize(p + 5);
    uint32_t u = static_cast<uint32_t>(node_id);
    memcpy(&t.data[p], &u, 4);
    t.data[p + 4] = static_cast<uint8_t>(static_cast<char>(node_id % 211));
  };
  t.mask.push_back(0);
  t.first_child.push_back(0);
  emit_data(1);
  for (int depth = 0; depth < level && !cur.empty(); ++depth) {
    next.clear();
    uint64_t next_first = level_first + cur.size();
    for (size_t i = 0; i < cur.size(); ++i) {
      const SphereNode& n = cur[i];
      uint8_t mask = 0;
      uint32_t first = static_cast<uint32_t>(next_first + next.size());
      for (int ch = 0; ch < 8; ++ch) {
        F3 minp = n.minp, maxp = n.maxp;
        F3 midp{(minp.x + maxp.x) * 0.5f, (minp.y + maxp.y) * 0.5f, (minp.z + maxp.z) * 0.5f};
        if (ch & 1) minp.x = midp.x; else maxp.x = midp.x;
        if (ch & 2) minp.y = midp.y; else maxp.y = midp.y;
        if (ch & 4) minp.z = midp.z; else maxp.z = midp.z;
        bool inside_box = center.x > minp.x && center.y > minp.y && center.z > minp.z &&
                          center.x < maxp.x && center.y < maxp.y && center.z < maxp.z;
        if (!inside_box) {
          F3 v[8] = {{minp.x, minp.y, minp.z}, {maxp.x, minp.y, minp.z}, {minp.x, maxp.y, minp.z},
                     {maxp.x, maxp.y, minp.z}, {minp.x, minp.y, maxp.z}, {maxp.x, minp.y, maxp.z},
                     {minp.x, maxp.y, maxp.z}, {maxp.x, maxp.y, maxp.z}};
          bool was_inside = false, was_outside = false;
          for (int k = 0; k < 8; ++k) {
            if (DistSq(v[k], center) < radius_square) {
              was_inside = true;
              if (was_outside) break;
            } else {
              was_outside = true;
              if (was_inside) break;
            }
          }
          if (!was_inside || !was_outside) continue;
        }
        mask |= uint8_t(1u << ch);
        int32_t id = static_cast<int32_t>(static_cast<uint32_t>(n.node_id) * 8u + uint32_t(ch));
        next.push_back({minp, maxp, id});
      }
      t.mask[level_first + i] = mask;
      t.first_child[level_first + i] = mask ? first : 0;
    }
    for (auto& c : next) {
      t.mask.push_back(0);
      t.first_child.push_back(0);
      emit_data(c.node_id);
    }
    level_first = next_first;
    std::swap(cur, next);
  }
  return t;
}

// -------------------------------------------------------------------- terrain
namespace {

static inline uint64_t SplitMix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline double Uniform(uint64_t& s, double lo, double hi) {
  return lo + (hi - lo) * (double(SplitMix(s) >> 11) * (1.0 / 9007199254740992.0));
}
static inline double LatticeValue(uint64_t seed, int64_t ix, int64_t iz, int octave) {
  uint64_t h = seed ^ (uint64_t(ix) * 0x8CB92BA72F3D8DD7ull) ^ (uint64_t(iz) * 0xD6E8FEB86659FD93ull) ^
               (uint64_t(octave) * 0xA0761D6478BD642Full);
  uint64_t s = h;
  return double(SplitMix(s) >> 11) * (1.0 / 9007199254740992.0);
}
static inline double Smooth(double t) { return t * t * (3.0 - 2.0 * t); }

// Seeded value-noise fBm (lacunarity 2, gain 0.5) over x, z in [0, 1]. The
// lattice values of each octave whose lattice is small (at most 2^22 points)
// are tabulated once; larger lattices are hashed per lookup. Both give the
// same values.
class Fbm {
 public:
  Fbm(uint64_t seed, int octaves, double base_freq) : seed_(seed), octaves_(octaves), base_freq_(base_freq) {
    double freq = base_freq;
    for (int o = 0; o < octaves; ++o, freq *= 2.0) {
      const int64_t w = int64_t(std::floor(freq)) + 2;  // lattice points 0 .. floor(freq) + 1 per axis
      width_.push_back(w);
      tab_.emplace_back();
      if (w * w > (int64_t(1) << 22)) continue;
      tab_.back().resize(size_t(w * w));
      for (int64_t iz = 0; iz < w; ++iz)
        for (int64_t ix = 0; ix < w; ++ix) tab_.back()[size_t(iz * w + ix)] = LatticeValue(seed, ix, iz, o);
    }
  }
  double operator()(double x, double z) const {
    double sum = 0, amp = 1, freq = base_freq_;
    for (int o = 0; o < octaves_; ++o) {
      double fx = x * freq, fz = z * freq;
      double flx = std::floor(fx), flz = std::floor(fz);
      int64_t ix = int64_t(flx), iz = int64_t(flz);
      double tx = Smooth(fx - flx), tz = Smooth(fz - flz);
      double v00, v10, v01, v11;
      const std::vector<double>& t = tab_[o];
      const int64_t w = width_[o];
      if (!t.empty() && ix >= 0 && iz >= 0 && ix + 1 < w && iz + 1 < w) {
        const double* r0 = &t[size_t(iz * w + ix)];
        v00 = r0[0], v10 = r0[1], v01 = r0[w], v11 = r0[w + 1];
      } else {
        v00 = LatticeValue(seed_, ix, iz, o), v10 = LatticeValue(seed_, ix + 1, iz, o);
        v01 = LatticeValue(seed_, ix, iz + 1, o), v11 = LatticeValue(seed_, ix + 1, iz + 1, o);
      }
      double a = v00 + (v10 - v00) * tx, b = v01 + (v11 - v01) * tx;
      sum += amp * (a + (b - a) * tz);
      amp *= 0.5;
      freq *= 2.0;
    }
    return sum;
  }

 private:
  uint64_t seed_;
  int octaves_;
  double base_freq_;
  std::vector<int64_t> width_;
  std::vector<std::vector<double>> tab_;
};

struct Sphere {
  double c[3], r;
};
struct Box {
  double lo[3], hi[3];
};

template <class F>
static void ParallelFor(int threads, size_t n, F&& f) {
  if (threads <= 1 || n < 4096) {
    f(size_t(0), n, 0);
    return;
  }
  std::vector<std::thread> th;
  size_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([&, b, e, t] { f(b, e, t); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

struct TerrainSource::Impl {
  explicit Impl(const TerrainParams& p) : p_(p) {
    threads_ = p.threads > 0 ? p.threads : std::max(1u, std::thread::hardware_concurrency());
    N_ = 1u << p.depth;
    uint64_t s = p.seed * 0x2545F4914F6CDD1Dull + 0x1234567ull;
    for (int i = 0; i < p.spheres; ++i) {
      Sphere sp;
      sp.r = Uniform(s, p.sphere_rmin, p.sphere_rmax);
      for (int a = 0; a < 3; ++a) sp.c[a] = Uniform(s, 0.1, 0.9);
      spheres_.push_back(sp);
    }
    for (int i = 0; i < p.boxes; ++i) {
      Box b;
      for (int a = 0; a < 3; ++a) {
        double e = Uniform(s, p.box_emin, p.box_emax);
        b.lo[a] = Uniform(s, 0.0, 1.0 - e);
        b.hi[a] = b.lo[a] + e;
      }
      boxes_.push_back(b);
    }
    channels_ = TerrainSource::Channels();
    BuildHeightmap();
  }

  TerrainParams p_;
  int threads_;
  uint32_t N_;
  std::vector<float> H_;                       // (N+1)^2 corner heights, row = z
  std::vector<std::vector<float>> hmin_, hmax_;  // per level L: 2^L x 2^L cells
  std::vector<Sphere> spheres_;
  std::vector<Box> boxes_;
  std::vector<StoredOctreeChannel> channels_;

  float Hat(uint32_t i, uint32_t k) const { return H_[size_t(k) * (N_ + 1) + i]; }
  int nprims() const { return 1 + int(spheres_.size()) + int(boxes_.size()); }

  void BuildHeightmap() {
    const uint32_t M = N_ + 1;
    H_.assign(size_t(M) * M, 0.f);
    std::vector<double> raw(size_t(M) * M);
    const Fbm fbm(p_.seed * 0x9E3779B97F4A7C15ull + 77, p_.octaves, p_.base_frequency);
    ParallelFor(threads_, M, [&](size_t b, size_t e, int) {
      for (size_t k = b; k < e; ++k)
        for (uint32_t i = 0; i < M; ++i) raw[k * M + i] = fbm(double(i) / N_, double(k) / N_);
    });
    double lo = raw[0], hi = raw[0];
    for (double v : raw) lo = std::min(lo, v), hi = std::max(hi, v);
    double scale = hi > lo ? (double(p_.height_hi) - p_.height_lo) / (hi - lo) : 0.0;
    for (size_t x = 0; x < raw.size(); ++x) H_[x] = float(p_.height_lo + (raw[x] - lo) * scale);
    // min/max pyramid over cell footprints (bilinear surface stays within corner range)
    const int D = p_.depth;
    hmin_.resize(D + 1);
    hmax_.resize(D + 1);
    hmin_[D].resize(size_t(N_) * N_);
    hmax_[D].resize(size_t(N_) * N_);
    ParallelFor(threads_, N_, [&](size_t b, size_t e, int) {
      for (size_t k = b; k < e; ++k)
        for (uint32_t i = 0; i < N_; ++i) {
          float a = Hat(i, uint32_t(k)), bb = Hat(i + 1, uint32_t(k)), c = Hat(i, uint32_t(k) + 1),
                d = Hat(i + 1, uint32_t(k) + 1);
          hmin_[D][k * N_ + i] = std::min(std::min(a, bb), std::min(c, d));
          hmax_[D][k * N_ + i] = std::max(std::max(a, bb), std::max(c, d));
        }
    });
    for (int L = D - 1; L >= 0; --L) {
      uint32_t n = 1u << L;
      hmin_[L].resize(size_t(n) * n);
      hmax_[L].resize(size_t(n) * n);
      const auto& cm = hmin_[L + 1];
      const auto& cM = hmax_[L + 1];
      for (uint32_t k = 0; k < n; ++k)
        for (uint32_t i = 0; i < n; ++i) {
          size_t c0 = size_t(2 * k) * (2 * n) + 2 * i, c1 = c0 + 2 * n;
          hmin_[L][size_t(k) * n + i] = std::min(std::min(cm[c0], cm[c0 + 1]), std::min(cm[c1], cm[c1 + 1]));
          hmax_[L][size_t(k) * n + i] = std::max(std::max(cM[c0], cM[c0 + 1]), std::max(cM[c1], cM[c1 + 1]));
        }
    }
  }

  // Does the cell (i,j,k) at level L straddle primitive `prim`?
  bool Straddles(int prim, uint32_t i, uint32_t j, uint32_t k, int L) const {
    if (prim == 0) {
      double s = std::ldexp(1.0, -L);
      size_t n = size_t(1) << L;
      float lo = hmin_[L][size_t(k) * n + i], hi = hmax_[L][size_t(k) * n + i];
      return double(lo) <= (j + 1) * s && double(hi) >= j * s;
    }
    double s = std::ldexp(1.0, -L);
    double cmin[3] = {i * s, j * s, k * s}, cmax[3] = {(i + 1) * s, (j + 1) * s, (k + 1) * s};
    int q = prim - 1;
    if (q < int(spheres_.size())) {
      const Sphere& sp = spheres_[q];
      double dmin = 0, dmax = 0;
      for (int a = 0; a < 3; ++a) {
        double c = sp.c[a];
        double nearest = std::max(cmin[a], std::min(c, cmax[a]));
        double far = (c - cmin[a] > cmax[a] - c) ? cmin[a] : cmax[a];
        dmin += (c - nearest) * (c - nearest);
        dmax += (c - far) * (c - far);
      }
      double r2 = sp.r * sp.r;
      return dmin <= r2 && dmax >= r2;
    }
    const Box& b = boxes_[q - spheres_.size()];
    bool overlap = true, inside = true;
    for (int a = 0; a < 3; ++a) {
      if (cmin[a] > b.hi[a] || cmax[a] < b.lo[a]) overlap = false;
      if (!(cmin[a] > b.lo[a] && cmax[a] < b.hi[a])) inside = false;
    }
    return overlap && !inside;
  }

  // Distance from the leaf centre to primitive `prim`'s surface and its outward normal there.
  double SurfaceDistance(int prim, const double p[3], uint32_t i, uint32_t k, double n[3]) const {
    if (prim == 0) {
      float h00 = Hat(i, k), h10 = Hat(i + 1, k), h01 = Hat(i, k + 1), h11 = Hat(i + 1, k + 1);
      double dhdx = ((double(h10) + h11) - (double(h00) + h01)) * 0.5 * N_;
      double dhdz = ((double(h01) + h11) - (double(h00) + h10)) * 0.5 * N_;
      n[0] = -dhdx; n[1] = 1.0; n[2] = -dhdz;
      double hc = (double(h00) + h10 + h01 + h11) * 0.25;
      return std::fabs(p[1] - hc);
    }
    int q = prim - 1;
    if (q < int(spheres_.size())) {
      const Sphere& sp = spheres_[q];
      double d2 = 0;
      for (int a = 0; a < 3; ++a) n[a] = p[a] - sp.c[a], d2 += n[a] * n[a];
      return std::fabs(std::sqrt(d2) - sp.r);
    }
    const Box& b = boxes_[q - spheres_.size()];
    double best = 1e300;
    int best_face = 0;
    for (int a = 0; a < 3; ++a) {
      double dl = std::fabs(p[a] - b.lo[a]), dh = std::fabs(p[a] - b.hi[a]);
      if (dl < best) best = dl, best_face = 2 * a;
      if (dh < best) best = dh, best_face = 2 * a + 1;
    }
    n[0] = n[1] = n[2] = 0;
    n[best_face / 2] = (best_face & 1) ? 1.0 : -1.0;
    return best;
  }

  void Gradient(double y, float rgb[3]) const {
    double t = (y - p_.height_lo) / std::max(1e-9, double(p_.height_hi) - p_.height_lo);
    t = std::min(1.0, std::max(0.0, t));
    static const double stops[4][4] = {{0.0, 0.16, 0.40, 0.12},
                                       {0.45, 0.46, 0.38, 0.22},
                                       {0.75, 0.55, 0.50, 0.45},
                                       {1.0, 0.95, 0.95, 0.97}};
    for (int s = 0; s < 3; ++s) {
      if (t <= stops[s + 1][0]) {
        double u = (t - stops[s][0]) / (stops[s + 1][0] - stops[s][0]);
        for (int c = 0; c < 3; ++c) rgb[c] = float(stops[s][c + 1] + (stops[s + 1][c + 1] - stops[s][c + 1]) * u);
        return;
      }
    }
    for (int c = 0; c < 3; ++c) rgb[c] = float(stops[3][c + 1]);
  }

  int Children(const TerrainCell& n, TerrainCell out[8], uint8_t* mask) const {
    *mask = 0;
    if (n.level >= p_.depth) return 0;
    int cnt = 0;
    // the height field's test (Straddles, prim 0) inlined: the children share
    // four footprints of the level's min/max pyramid
    const int L = n.level + 1;
    const double sc = std::ldexp(1.0, -L);
    const size_t w = size_t(1) << L;
    const float* mn = hmin_[L].data();
    const float* mx = hmax_[L].data();
    const bool terrain = n.prims[0] & 1u;
    for (int ch = 0; ch < 8; ++ch) {
      TerrainCell c;
      c.i = 2 * n.i + (ch & 1);
      c.j = 2 * n.j + ((ch >> 1) & 1);
      c.k = 2 * n.k + ((ch >> 2) & 1);
      c.level = L;
      if (terrain) {
        const size_t f = size_t(c.k) * w + c.i;
        if (double(mn[f]) <= (c.j + 1) * sc && double(mx[f]) >= c.j * sc) c.prims[0] |= 1u;
      }
      for (int wd = 0; wd < 2; ++wd) {
        uint64_t m = n.prims[wd] & (wd == 0 ? ~uint64_t(1) : ~uint64_t(0));
        while (m) {
          int bit = __builtin_ctzll(m);
          m &= m - 1;
          if (Straddles(wd * 64 + bit, c.i, c.j, c.k, c.level)) c.prims[wd] |= 1ull << bit;
        }
      }
      if (c.prims[0] | c.prims[1]) {
        *mask |= uint8_t(1u << ch);
        out[cnt++] = c;
      }
    }
    return cnt;
  }

  // Leaves take the nearest surface among the primitives they straddle.
  void LeafValues(const TerrainCell& c, float nc[6]) const {
    const double s = std::ldexp(1.0, -c.level);
    double p[3] = {(c.i + 0.5) * s, (c.j + 0.5) * s, (c.k + 0.5) * s};
    double best = 1e300, bn[3] = {0, 0, 1};
    uint32_t fi = c.i << (p_.depth - c.level), fk = c.k << (p_.depth - c.level);
    for (int w = 0; w < 2; ++w) {
      uint64_t mm = c.prims[w];
      while (mm) {
        int q = w * 64 + __builtin_ctzll(mm);
        mm &= mm - 1;
        double n[3];
        double d = SurfaceDistance(q, p, fi, fk, n);
        if (d < best) best = d, bn[0] = n[0], bn[1] = n[1], bn[2] = n[2];
      }
    }
    double len = std::sqrt(bn[0] * bn[0] + bn[1] * bn[1] + bn[2] * bn[2]);
    for (int a = 0; a < 3; ++a) nc[a] = len > 0 ? float(bn[a] / len) : (a == 2 ? 1.f : 0.f);
    Gradient(p[1], nc + 3);
  }

  NodeTree Subtree(const TerrainCell& top, int threads, float root_values[6]) const {
    NodeTree t;
    t.max_depth = p_.depth;
    t.channels = channels_;
    t.data_size = 8;
    const int L0 = top.level, D = p_.depth;
    // Level-synchronous expansion. levels[L - L0] holds the cells of level L in BFS order.
    std::vector<std::vector<TerrainCell>> levels(D - L0 + 1);
    levels[0].push_back(top);
    std::vector<std::vector<uint8_t>> masks(D - L0 + 1);
    for (int L = L0; L < D; ++L) {
      const auto& cur = levels[L - L0];
      masks[L - L0].assign(cur.size(), 0);
      std::vector<std::vector<TerrainCell>> parts(threads);
      ParallelFor(threads, cur.size(), [&](size_t b, size_t e, int tid) {
        auto& out = parts[tid];
        TerrainCell kids[8];
        for (size_t x = b; x < e; ++x) {
          const int n = Children(cur[x], kids, &masks[L - L0][x]);
          out.insert(out.end(), kids, kids + n);
        }
      });
      size_t total = 0;
      for (auto& pp : parts) total += pp.size();
      levels[L + 1 - L0].reserve(total);
      for (auto& pp : parts) levels[L + 1 - L0].insert(levels[L + 1 - L0].end(), pp.begin(), pp.end());
      if (threads > 1 && std::getenv("RAYT_BUILD_LOG")) {  // progress of long builds
        std::fprintf(stderr, "[rayt build] level %d: %zu cells\n", L + 1, total);
        std::fflush(stderr);
      }
    }
    masks[D - L0].assign(levels[D - L0].size(), 0);

    // Flatten: node index = level offset + position; children contiguous.
    std::vector<uint64_t> level_off(D - L0 + 2, 0);
    for (int L = L0; L <= D; ++L) level_off[L - L0 + 1] = level_off[L - L0] + levels[L - L0].size();
    const uint64_t total = level_off[D - L0 + 1];
    if (total > 0xffffffffull) throw std::runtime_error("subtree has more than 2^32 nodes");
    t.mask.resize(total);
    t.first_child.resize(total);
    t.data.resize(total * 8);
    for (int L = L0; L <= D; ++L) {
      uint64_t next = level_off[L - L0 + 1];
      for (size_t x = 0; x < levels[L - L0].size(); ++x) {
        uint8_t m = masks[L - L0][x];
        t.mask[level_off[L - L0] + x] = m;
        t.first_child[level_off[L - L0] + x] = m ? uint32_t(next) : 0;
        next += __builtin_popcount(m);
      }
    }

    // Node data bottom-up: leaves from the nearest primitive surface; inner
    // nodes average their children (normal sum re-normalised, colour mean).
    std::vector<float> child_nc, cur_nc;  // 6 floats per node (normal xyz, colour rgb)
    for (int L = D; L >= L0; --L) {
      const auto& lv = levels[L - L0];
      cur_nc.assign(lv.size() * 6, 0.f);
      ParallelFor(threads, lv.size(), [&](size_t b, size_t e, int) {
        for (size_t x = b; x < e; ++x) {
          uint64_t node = level_off[L - L0] + x;
          float* nc = &cur_nc[x * 6];
          uint8_t m = t.mask[node];
          if (!m)
            LeafValues(lv[x], nc);
          else
            InnerValues(&child_nc[(t.first_child[node] - level_off[L - L0 + 1]) * 6], __builtin_popcount(m), nc);
          PackValues(nc, &t.data[node * 8]);
        }
      });
      std::swap(child_nc, cur_nc);
    }
    for (int a = 0; a < 6; ++a) root_values[a] = child_nc[a];
    return t;
  }

  static void InnerValues(const float* children, int cnt, float nc[6]) {
    float n[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    for (int q = 0; q < cnt; ++q) {
      const float* cc = &children[q * 6];
      for (int a = 0; a < 3; ++a) n[a] += cc[a], c[a] += cc[3 + a];
    }
    float inv = 1.0f / float(cnt);
    for (int a = 0; a < 3; ++a) c[a] *= inv;
    float ls = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (ls < 1e-5f) {
      n[0] = 0; n[1] = 0; n[2] = 1;
    } else {
      float il = 1.0f / std::sqrt(ls);
      for (int a = 0; a < 3; ++a) n[a] *= il;
    }
    for (int a = 0; a < 3; ++a) nc[a] = n[a], nc[3 + a] = c[a];
  }

  // Quantised channel bytes: normal char4 = (char)(n*127), w 0; colour uchar4 =
  // (uchar)clamp(c*255), alpha 255 (import-obj-scene.cpp:516-526).
  static void PackValues(const float v[6], uint8_t* out) {
    for (int a = 0; a < 3; ++a) out[a] = static_cast<uint8_t>(static_cast<char>(v[a] * 127));
    out[3] = 0;
    for (int a = 0; a < 3; ++a)
      out[4 + a] = static_cast<uint8_t>(std::max(0.0f, std::min(255.0f, v[3 + a] * 255)));
    out[7] = 255;
  }

  TerrainCell Root() const {
    TerrainCell r;
    for (int q = 0; q < nprims(); ++q)
      if (Straddles(q, 0, 0, 0, 0)) r.prims[q >> 6] |= 1ull << (q & 63);
    return r;
  }
};

TerrainSource::TerrainSource(const TerrainParams& p) : impl_(new Impl(p)) {}
TerrainSource::~TerrainSource() { delete impl_; }
int TerrainSource::depth() const { return impl_->p_.depth; }
int TerrainSource::threads() const { return impl_->threads_; }
const std::vector<StoredOctreeChannel>& TerrainSource::channels() const { return impl_->channels_; }
std::vector<StoredOctreeChannel> TerrainSource::Channels() { return {{4, "normals"}, {4, "colors"}}; }
TerrainCell TerrainSource::Root() const { return impl_->Root(); }
int TerrainSource::Children(const TerrainCell& c, TerrainCell out[8], uint8_t* mask) const {
  return impl_->Children(c, out, mask);
}
NodeTree TerrainSource::Subtree(const TerrainCell& c, int threads, float root_values[6]) const {
  return impl_->Subtree(c, std::max(1, threads), root_values);
}
void TerrainSource::LeafValues(const TerrainCell& c, float out[6]) const { impl_->LeafValues(c, out); }
void TerrainSource::InnerValues(const float* children, int count, float out[6]) {
  Impl::InnerValues(children, count, out);
}
void TerrainSource::PackValues(const float v[6], uint8_t out[8]) { Impl::PackValues(v, out); }

NodeTree GenerateTerrainTree(const TerrainParams& p) {
  TerrainSource src(p);
  float root[6];
  return src.Subtree(src.Root(), src.threads(), root);
}

}  // namespace rayt
