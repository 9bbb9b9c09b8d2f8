// ORACLE - test infrastructure only.
//
// A CPU restatement of the reference renderer's hot path, used as the checker
// for the HIP implementation by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py. Nothing in the product links or calls it.
// Built with g++ -O2 -ffp-contract=off (SSE2, IEEE fp32, correctly rounded
// division and sqrt), so every fp32 operation is the single rounded IEEE op.
//
// What is restated, and from where (reference paths):
//   CastRay                 GPU_renderer/kernels/raycast.cl:33-256
//   InitTracingFrame + mul  GPU_renderer/kernels/init_tracing_frame.cl:3-45
//   RaytracingPass          GPU_renderer/kernels/raytracing_pass.cl:4-61
//   ProcessHit              GPU_renderer/kernels/shading.cl:4-33
//   FinishTracingFrame      GPU_renderer/kernels/finish_tracing_frame.cl:3-24
//   RaytraceKernel (depth)  GPU_renderer/kernels/raytracer.cl:16-72
//   restart CastRay         raytracer_volatile.cl:25-177
//   NodeToBlock/Diff/Compress GPU_renderer/kernels/feedback_compression.cl:4-34
//   feedback lists          GPU_renderer/ray-tracing-feedback-extractor.cpp:38-185
//   RenderFrame pass loop   GPU_renderer/gpu-ray-tracer.cpp:66-227
//   block cache + LRU       GPU data/gpu-octree-cache.cpp:11-313,
//                           GPU data/gpu-octree-cache-manager.cpp:11-212
//   .tree reader            CPU data/stored-octree-loader.cpp:31-98
//   link codecs             CPU data/octree-constants.h:13-53
//   Camera / fmat4          scene/camera.cpp:23-28, util/mat.cpp:6-194
//   sphere checkers         scene sources/stored-octree-test.cpp:39-233
//
// FP32 definitions fixed here (OpenCL leaves them to the implementation):
// dot = ((x*a + y*b) + z*c) + w*d; normalize(v) = v / sqrt(dot(v,v));
// distance(a,b) = sqrt(dot(a-b, a-b)); fmax(x,y) = y if x<y else x (NaN x ->
// y); fmin(x,y) = y if y<x else x (NaN x -> y); max/min the same without the
// NaN rule; float->uchar truncates. rand() of the hit window is replaced by a
// seeded splitmix64 draw (>> 33). The shadow ray is an extension (no
// reference counterpart): from hit_t - 2^-(depth+2) along the primary ray,
// towards -normalize(2,-2,3), without LOD; it drops light 1 when it hits.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <list>
#include <map>
#include <memory>
#include <queue>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace orc {

static const int S_MAX = 23;
static const float EPSILON = 0.00000011920928955078125f;
static const int NO_HIT = (1 << 30) - 1 + (1 << 30);
static const uint32_t ROOT_BLOCK_PARENT = (1u << 24) - 1;

static inline uint32_t as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline float cl_fmax(float x, float y) { if (x != x) return y; return (x < y) ? y : x; }
static inline float cl_fmin(float x, float y) { if (x != x) return y; return (y < x) ? y : x; }
static inline float cl_max(float x, float y) { return (x < y) ? y : x; }
static inline float cl_min(float x, float y) { return (y < x) ? y : x; }
static inline int popcount8(uint32_t a) {
  static const int l[16] = {0, 1, 1, 2, 1, 2, 2, 3, 1, 2, 2, 3, 2, 3, 3, 4};
  a &= 255u;
  return l[a & 15] + l[a >> 4];
}

struct F4 { float x, y, z, w; };
static inline F4 sub4(F4 a, F4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
static inline float dot4(F4 a, F4 b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
static inline F4 normalize4(F4 v) {
  float l = sqrtf(dot4(v, v));
  return {v.x / l, v.y / l, v.z / l, v.w / l};
}
static inline float distance4(F4 a, F4 b) { F4 d = sub4(a, b); return sqrtf(dot4(d, d)); }

// mul(float16, float3) of init_tracing_frame.cl:3-9 (float3 is float4 with w 0).
static F4 mul(const float* m, F4 v3) {
  F4 vec = {v3.x, v3.y, v3.z, 1.0f};
  float d = dot4(vec, F4{m[12], m[13], m[14], m[15]});
  F4 r = {dot4(vec, F4{m[0], m[1], m[2], m[3]}), dot4(vec, F4{m[4], m[5], m[6], m[7]}),
          dot4(vec, F4{m[8], m[9], m[10], m[11]}), 0.0f};
  return {r.x / d, r.y / d, r.z / d, r.w / d};
}

struct RayCastResult {
  float hit_t;
  int hit_node;
  int fault_block;
  int hit_duplicate;
  int error_code;
  int node_depth;
};

// raycast.cl:33-256. `fetches` counts node-link and far-pointer loads.
// Extension (no reference counterpart; the product kernels do the same, and
// on a valid pool it never triggers): a far-pointer index outside the far
// table (far_count entries), a child offset outside the pool (node_count
// links), or a push below the finest scale ends the ray with error code 2
// instead of reading outside the arrays; the reference follows such links
// unchecked.
static void CastRay(const uint32_t* node_links, uint64_t node_count, const uint32_t* far_pointers,
                    uint64_t far_count, int root_node, F4 p, F4 d,
                    float ray_size_coef, float ray_size_bias, RayCastResult* res, uint64_t* fetches) {
  p.x += 1.0f; p.y += 1.0f; p.z += 1.0f;  // p += (1,1,1,0)
  res->fault_block = -1;
  res->error_code = 0;
  res->hit_duplicate = 0;
  uint32_t stack_parent[S_MAX + 1];
  float stack_tmax[S_MAX + 1];
  if (fabsf(d.x) < EPSILON) d.x = copysignf(EPSILON, d.x);
  if (fabsf(d.y) < EPSILON) d.y = copysignf(EPSILON, d.y);
  if (fabsf(d.z) < EPSILON) d.z = copysignf(EPSILON, d.z);
  float tx_coef = 1.0f / -fabsf(d.x);
  float ty_coef = 1.0f / -fabsf(d.y);
  float tz_coef = 1.0f / -fabsf(d.z);
  float tx_bias = tx_coef * p.x;
  float ty_bias = ty_coef * p.y;
  float tz_bias = tz_coef * p.z;
  unsigned octant_mask = 0;
  if (d.x > 0.0f) octant_mask ^= 1, tx_bias = 3.0f * tx_coef - tx_bias;
  if (d.y > 0.0f) octant_mask ^= 2, ty_bias = 3.0f * ty_coef - ty_bias;
  if (d.z > 0.0f) octant_mask ^= 4, tz_bias = 3.0f * tz_coef - tz_bias;
  float t_min = cl_fmax(cl_fmax(2.0f * tx_coef - tx_bias, 2.0f * ty_coef - ty_bias), 2.0f * tz_coef - tz_bias);
  float t_max = cl_fmin(cl_fmin(tx_coef - tx_bias, ty_coef - ty_bias), tz_coef - tz_bias);
  float h = t_max;
  t_min = cl_fmax(t_min, 0.0f);
  uint32_t parent = uint32_t(root_node);
  uint32_t child_descriptor = 0;
  unsigned idx = 0;
  float posx = 1.0f, posy = 1.0f, posz = 1.0f;
  int scale = S_MAX - 1;
  float scale_exp2 = 0.5f;
  if (1.5f * tx_coef - tx_bias > t_min) idx ^= 1, posx = 1.5f;
  if (1.5f * ty_coef - ty_bias > t_min) idx ^= 2, posy = 1.5f;
  if (1.5f * tz_coef - tz_bias > t_min) idx ^= 4, posz = 1.5f;
  unsigned iter = 0;
  const unsigned iter_limit = 300;
  while (scale < S_MAX) {
    if (++iter > iter_limit) {
      res->error_code = 1;
      res->hit_t = t_min;
      res->hit_node = int(parent);
      res->node_depth = S_MAX - scale - 1;
      return;
    }
    if (child_descriptor == 0) {
      child_descriptor = node_links[parent];
      ++*fetches;
    }
    float tx_corner = posx * tx_coef - tx_bias;
    float ty_corner = posy * ty_coef - ty_bias;
    float tz_corner = posz * tz_coef - tz_bias;
    float tc_max = cl_fmin(cl_fmin(tx_corner, ty_corner), tz_corner);
    unsigned child_shift = idx ^ octant_mask;
    unsigned child_masks = (1u << child_shift) & 255u;
    if ((child_descriptor & child_masks) != 0 && t_min <= t_max) {
      if ((child_descriptor & (1u << 8)) != 0) {
        res->hit_node = int(parent);
        res->hit_t = t_min;
        res->fault_block = int(child_descriptor >> 11);
        res->node_depth = S_MAX - scale - 1;
        return;
      }
      uint32_t ofs;
      bool bad = false;
      if (child_descriptor & (1u << 10)) {
        if ((child_descriptor >> 11) >= far_count) {
          bad = true;
          ofs = 0;
        } else {
          ofs = far_pointers[child_descriptor >> 11];
          ++*fetches;
        }
      } else {
        ofs = (child_descriptor >> 11) + parent - (1u << 20);
      }
      ofs += uint32_t(popcount8(child_descriptor & (child_masks - 1)));
      if (bad || ofs >= node_count || scale < 1) {
        res->error_code = 2;
        res->hit_t = t_min;
        res->hit_node = int(parent);
        res->node_depth = S_MAX - scale - 1;
        return;
      }
      child_descriptor = node_links[ofs];
      ++*fetches;
      if (!(child_descriptor & 255u) || tc_max * ray_size_coef + ray_size_bias >= scale_exp2) {
        res->hit_t = t_min;
        res->hit_node = int(ofs);
        res->hit_duplicate = int((child_descriptor >> 9) & 1u);
        res->node_depth = S_MAX - scale;
        return;
      }
      float half = scale_exp2 * 0.5f;
      float tx_center = half * tx_coef + tx_corner;
      float ty_center = half * ty_coef + ty_corner;
      float tz_center = half * tz_coef + tz_corner;
      if (tc_max < h) {
        stack_parent[scale] = parent;
        stack_tmax[scale] = t_max;
      }
      h = tc_max;
      parent = ofs;
      idx = 0;
      scale--;
      scale_exp2 = half;
      if (tx_center > t_min) idx ^= 1, posx += scale_exp2;
      if (ty_center > t_min) idx ^= 2, posy += scale_exp2;
      if (tz_center > t_min) idx ^= 4, posz += scale_exp2;
      t_max = tc_max;
      continue;
    }
    unsigned step_mask = 0;
    if (tx_corner <= tc_max) step_mask ^= 1, posx -= scale_exp2;
    if (ty_corner <= tc_max) step_mask ^= 2, posy -= scale_exp2;
    if (tz_corner <= tc_max) step_mask ^= 4, posz -= scale_exp2;
    t_min = tc_max;
    idx ^= step_mask;
    if ((idx & step_mask) != 0) {
      uint32_t differing_bits = 0;
      if ((step_mask & 1) != 0) differing_bits |= as_uint(posx) ^ as_uint(posx + scale_exp2);
      if ((step_mask & 2) != 0) differing_bits |= as_uint(posy) ^ as_uint(posy + scale_exp2);
      if ((step_mask & 4) != 0) differing_bits |= as_uint(posz) ^ as_uint(posz + scale_exp2);
      scale = int(as_uint(float(differing_bits)) >> 23) - 127;  // position of the highest bit
      scale_exp2 = as_float(uint32_t(scale - S_MAX + 127) << 23);
      if (scale >= S_MAX) break;  // the reference reads stack[scale] here and then leaves the loop
      parent = stack_parent[scale];
      t_max = stack_tmax[scale];
      uint32_t shx = as_uint(posx) >> scale;
      uint32_t shy = as_uint(posy) >> scale;
      uint32_t shz = as_uint(posz) >> scale;
      posx = as_float(shx << scale);
      posy = as_float(shy << scale);
      posz = as_float(shz << scale);
      idx = (shx & 1) | ((shy & 1) << 1) | ((shz & 1) << 2);
      h = 0.0f;
      child_descriptor = 0;
    }
  }
  res->hit_t = 1e30f;
  res->hit_node = -1;
}

struct Ray {
  F4 origin;     // w: ray size bias
  F4 direction;  // w: ray size coefficient
};

// InitTracingFrame for one pixel.
static Ray InitRay(const float* m, int x, int y, int W, int H, float lod) {
  F4 tmp = {(x + 0.5f) / float(W) * 2 - 1, (y + 0.5f) / float(H) * 2 - 1, -1, 0};
  F4 origin = {m[2] / m[14], m[6] / m[14], m[10] / m[14], 0.0f / m[14]};
  tmp.z = 0;
  F4 direction = sub4(mul(m, tmp), origin);
  direction = normalize4(direction);
  tmp = {float(x) / float(W) * 2 - 1, float(y) / float(H) * 2 - 1, 0, 0};
  F4 p1far = mul(m, tmp);
  tmp = {float(x + 1) / float(W) * 2 - 1, float(y + 1) / float(H) * 2 - 1, 0, 0};
  F4 p2far = mul(m, tmp);
  float ray_size_bias = 0;
  float ray_size_coef = lod * distance4(p1far, p2far) / distance4(origin, p1far);
  Ray r;
  r.origin = {origin.x, origin.y, origin.z, ray_size_bias};
  r.direction = {direction.x, direction.y, direction.z, ray_size_coef};
  return r;
}

struct Lights {
  F4 l1, l2, l3;
  F4 shadow_dir;
};
static Lights MakeLights() {
  auto n = [](float x, float y, float z, float k) {
    F4 v = normalize4(F4{x, y, z, 0});
    return F4{v.x * k, v.y * k, v.z * k, v.w * k};
  };
  Lights L;
  L.l1 = n(2, -2, 3, .6f);
  L.l2 = n(-3, -3, -2, .5f);
  L.l3 = n(1, 2.5f, -1, .3f);
  F4 u = n(2, -2, 3, 1.0f);
  L.shadow_dir = {-u.x, -u.y, -u.z, 0};
  return L;
}

static inline F4 DecodeNormal(uint32_t c) {
  return {float(int8_t(c & 255)) / 127.f, float(int8_t((c >> 8) & 255)) / 127.f,
          float(int8_t((c >> 16) & 255)) / 127.f, 0.0f / 127.f};
}
static inline float LightDot(F4 n, F4 L) { return dot4(F4{-n.x, -n.y, -n.z, -n.w}, L); }

static inline uint32_t rgba(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  return r | (g << 8) | (b << 16) | (a << 24);
}

// ProcessHit colour (shading.cl:4-33); occluded1 drops light 1.
static uint32_t ProcessHitColor(uint32_t cnormal, uint32_t cncol, const Lights& L, bool occluded1) {
  F4 normal = DecodeNormal(cnormal);
  float m1 = occluded1 ? 0.0f : cl_max(0.f, LightDot(normal, L.l1));
  float m2 = cl_max(0.f, LightDot(normal, L.l2));
  float m3 = cl_max(0.f, LightDot(normal, L.l3));
  float light = ((.1f + m1 * 1.0f) + m2 * 1.0f) + m3 * 1.0f;
  float cr = cl_min(255.f, float(cncol & 255) * light);
  float cg = cl_min(255.f, float((cncol >> 8) & 255) * light);
  float cb = cl_min(255.f, float((cncol >> 16) & 255) * light);
  return rgba(uint32_t(cr), uint32_t(cg), uint32_t(cb), 255);
}

// finish_tracing_frame.cl:21-22.
static uint32_t Blend(uint32_t c, const float* bg) {
  float w = float(c >> 24);
  float k = 255.f - w;
  float t[4] = {(float(c & 255) + bg[0] * k) + .5f, (float((c >> 8) & 255) + bg[1] * k) + .5f,
                (float((c >> 16) & 255) + bg[2] * k) + .5f, (w + bg[3] * k) + .5f};
  uint32_t o[4];
  for (int i = 0; i < 4; ++i) o[i] = uint32_t(cl_max(0.f, cl_min(255.f, t[i])));
  return rgba(o[0], o[1], o[2], o[3]);
}

struct Pool {
  const uint32_t* links;
  const uint32_t* far_ptrs;
  uint64_t nodes, far_count;  // element counts of links and far_ptrs
  int root;
  const uint32_t* normals;  // may be null (treated as 0)
  const uint32_t* colors;   // may be null (treated as 0xffffffff)
};

static uint32_t ShadeNode(const Pool& P, int node, const Lights& L, bool occluded) {
  uint32_t n = P.normals ? P.normals[node] : 0u;
  uint32_t c = P.colors ? P.colors[node] : 0xffffffffu;
  return ProcessHitColor(n, c, L, occluded);
}

static bool ShadowOccluded(const Pool& P, F4 o, F4 d, const RayCastResult& hit, const Lights& L,
                           uint64_t* fetches) {
  float eps = as_float(uint32_t(127 - hit.node_depth - 2) << 23);
  float tb = cl_max(0.0f, hit.hit_t - eps);
  F4 s = {o.x + d.x * tb, o.y + d.y * tb, o.z + d.z * tb, 0};
  RayCastResult r;
  CastRay(P.links, P.nodes, P.far_ptrs, P.far_count, P.root, s, L.shadow_dir, 0.0f, 0.0f, &r, fetches);
  return r.error_code == 0 && r.fault_block == -1 && r.hit_node != -1;
}

static int HitCode(int hit_node, int dup, int nib) {
  int unit = nib > 0 ? hit_node / nib : hit_node;
  return ((unit + 1) << 1) | dup;
}

// Per-pixel tracing state (common.cl:25-32).
struct TracingState {
  F4 ray_origin;
  F4 ray_direction;
  int fault_parent_node;
  uint32_t color;
  uint32_t color_multiplier;
  int reflections_count;
};

static const int FLAG_SHADOWS = 1;

// RaytracingPass for one pixel; returns the faults_and_hits code.
static int RaytracingPassPixel(const Pool& P, TracingState& s, int nib, int flags, const Lights& L,
                               uint64_t* fetches, uint64_t* shadow_rays, RayCastResult* out_res) {
  if (s.color_multiplier == 0) {
    if (out_res) out_res->hit_node = -2;
    return NO_HIT;
  }
  RayCastResult res;
  CastRay(P.links, P.nodes, P.far_ptrs, P.far_count, P.root, s.ray_origin, s.ray_direction, s.ray_direction.w,
          s.ray_origin.w, &res,
          fetches);
  if (out_res) *out_res = res;
  if (res.error_code) {
    s.color = rgba(255, 0, 0, 255);
    s.color_multiplier = 0;
    return NO_HIT;
  } else if (res.fault_block != -1) {
    F4& o = s.ray_origin;
    const F4& d = s.ray_direction;
    o = {o.x + d.x * res.hit_t, o.y + d.y * res.hit_t, o.z + d.z * res.hit_t, o.w + d.w * res.hit_t};
    s.fault_parent_node = res.hit_node;
    return -res.fault_block - 1;
  } else if (res.hit_node == -1) {
    s.color_multiplier = 0;
    return NO_HIT;
  }
  bool occluded = false;
  if ((flags & FLAG_SHADOWS) && P.normals) {
    F4 n = DecodeNormal(P.normals[res.hit_node]);
    if (cl_max(0.f, LightDot(n, L.l1)) > 0.0f) {
      occluded = ShadowOccluded(P, s.ray_origin, s.ray_direction, res, L, fetches);
      ++*shadow_rays;
    }
  }
  s.color = ShadeNode(P, res.hit_node, L, occluded);
  s.color_multiplier = 0;
  return HitCode(res.hit_node, res.hit_duplicate, nib);
}

static uint32_t FinishPixel(const Pool& P, const TracingState& s, const float* bg, const Lights& L) {
  uint32_t c = s.color;
  if (s.color_multiplier != 0) c = s.fault_parent_node >= 0 ? ShadeNode(P, s.fault_parent_node, L, false) : 0u;
  return Blend(c, bg);
}

template <class F>
static void ParallelRows(int H, int threads, F&& f) {
  if (threads <= 1) {
    for (int y = 0; y < H; ++y) f(y, 0);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> th;
  for (int t = 0; t < threads; ++t)
    th.emplace_back([&, t] {
      for (int y; (y = next.fetch_add(1)) < H;) f(y, t);
    });
  for (auto& x : th) x.join();
}

static int Threads(int t) {
  if (t > 0) return t;
  unsigned h = std::thread::hardware_concurrency();
  return h ? int(h) : 1;
}

// ------------------------------------------------------------------ .tree
struct Channel {
  int bytes;
  std::string name;
};
struct Tree {
  FILE* f = nullptr;
  uint32_t nib = 0, blocks = 0, root_block = 0, stride = 0;
  std::vector<Channel> channels;
  uint64_t header_size = 0;
  uint64_t content = 0;
  ~Tree() { if (f) fclose(f); }
};
struct Block {
  uint32_t index, parent, roots_count;
  uint16_t parent_ptr_idx[8], child_idx[8];
  uint32_t parent_ptr_value[8];
  std::vector<uint8_t> data;
  uint32_t link(int n) const { uint32_t v; memcpy(&v, &data[size_t(n) * 4], 4); return v; }
  void set_link(int n, uint32_t v) { memcpy(&data[size_t(n) * 4], &v, 4); }
};
static bool ReadAt(FILE* f, uint64_t pos, size_t n, void* dst) {
  if (fseeko(f, off_t(pos), SEEK_SET)) return false;
  return fread(dst, 1, n, f) == n;
}
static uint32_t u32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint16_t u16(const uint8_t* p) { uint16_t v; memcpy(&v, p, 2); return v; }

static Tree* OpenTree(const char* path) {
  std::unique_ptr<Tree> t(new Tree());
  t->f = fopen(path, "rb");
  if (!t->f) return nullptr;
  uint8_t h[20];
  if (!ReadAt(t->f, 0, 20, h)) return nullptr;
  t->nib = u32(h);
  t->blocks = u32(h + 4);
  t->root_block = u32(h + 8);
  t->stride = u32(h + 12);
  uint32_t nc = u32(h + 16);
  uint64_t pos = 20;
  for (uint32_t i = 0; i < nc; ++i) {
    uint8_t r[8];
    if (!ReadAt(t->f, pos, 8, r)) return nullptr;
    pos += 8;
    Channel c;
    c.bytes = int(u32(r));
    uint32_t l = u32(r + 4);
    c.name.resize(l);
    if (l && !ReadAt(t->f, pos, l, &c.name[0])) return nullptr;
    pos += l;
    t->channels.push_back(c);
  }
  t->header_size = pos;
  for (auto& c : t->channels) t->content += uint64_t(c.bytes) * t->nib;
  return t.release();
}

static bool LoadBlock(Tree* t, uint32_t index, Block* b) {
  if (index >= t->blocks) return false;
  uint64_t pos = t->header_size + uint64_t(t->stride) * index;
  uint8_t h[76];
  if (!ReadAt(t->f, pos, 76, h)) return false;
  b->index = u32(h);
  b->parent = u32(h + 4);
  b->roots_count = u32(h + 8);
  for (int i = 0; i < 8; ++i) {
    b->parent_ptr_idx[i] = u16(h + 12 + i * 8);
    b->child_idx[i] = u16(h + 14 + i * 8);
    b->parent_ptr_value[i] = u32(h + 16 + i * 8);
  }
  b->data.resize(t->content);
  return ReadAt(t->f, pos + 76, t->content, b->data.data());
}

// ---------------------------------------------------------- block cache
// GPUOctreeCache + GPUOctreeCacheManager restated over host arrays.
static inline uint32_t PackCache(uint32_t mask, bool fault, bool dup, bool far_flag, uint32_t ptr) {
  return (ptr << 11) | ((far_flag ? 1u : 0u) << 10) | ((dup ? 1u : 0u) << 9) | ((fault ? 1u : 0u) << 8) | (mask & 255u);
}
static inline bool NeedFar(int64_t a, int64_t b) { int64_t d = a - b; return (d < 0 ? -d : d) >= (1 << 20); }
static inline uint32_t NearValue(int64_t pointer_index, int64_t pointed) { return uint32_t((1 << 20) + (pointed - pointer_index)); }

struct Cache;
struct Marker {
  bool is_hit = false;
};
struct LRUEntry {
  bool is_marker;
  int block;       // slot when !is_marker
  Marker* marker;  // when is_marker
};

struct CachedBlockInfo {
  int block_index = -1;
  int parent_block_index = -1;  // slot of the parent, ROOT_BLOCK_PARENT for the root
  int roots_count = 0;
  uint16_t pointer_index_in_parent[8];
  uint32_t pointer_value_in_parent[8];
  std::list<LRUEntry>::iterator lru;
  std::vector<int> child_blocks;
};

struct Cache {
  Tree* tree;
  int nib, max_blocks;
  int first_free = 0;
  int root_node_index = -1;
  std::vector<uint32_t> links, far_ptrs;
  std::vector<std::vector<uint8_t>> chans;  // channels 1.. (index c-1)
  std::map<int, int> block_to_cache;
  std::map<int, int> block_to_parent_cache;
  std::vector<CachedBlockInfo> contents;
  std::list<LRUEntry> lru;
  std::unique_ptr<Marker> session_marker, transaction_marker;
  std::list<LRUEntry>::iterator session_it, transaction_it;
  bool session = false, transaction = false;
  std::vector<int> transaction_blocks;
  uint64_t uploaded = 0, evicted = 0, patches = 0, far_writes = 0;

  Cache(Tree* t, int maxb) : tree(t), nib(int(t->nib)), max_blocks(maxb) {
    links.assign(size_t(maxb) * nib, 0);
    far_ptrs.assign(size_t(maxb) * 8, 0);
    for (size_t c = 1; c < t->channels.size(); ++c)
      chans.emplace_back(size_t(maxb) * nib * t->channels[c].bytes, 0);
    contents.resize(maxb);
  }

  void DequeueMarkers() {
    while (!lru.empty() && lru.front().is_marker) {
      lru.front().marker->is_hit = true;
      lru.pop_front();
    }
  }
  // gpu-octree-cache.cpp:58-66
  void MarkBlockAsUsed(int index) {
    for (int idx = index; uint32_t(idx) != ROOT_BLOCK_PARENT; idx = contents[idx].parent_block_index) {
      lru.erase(contents[idx].lru);
      contents[idx].lru = lru.insert(lru.end(), LRUEntry{false, idx, nullptr});
    }
    DequeueMarkers();
  }
  void MarkParentAsUsed(int index, bool global) {
    if (global)
      MarkBlockAsUsed(block_to_parent_cache.at(index));
    else
      MarkBlockAsUsed(contents[index].parent_block_index);
  }
  std::list<LRUEntry>::iterator InsertMarker(Marker* m) { return lru.insert(lru.end(), LRUEntry{true, -1, m}); }
  void RemoveMarker(Marker* m, std::list<LRUEntry>::iterator it) {
    if (!m->is_hit) lru.erase(it);
  }

  void WriteLink(int slot, int node, uint32_t v) { links[size_t(slot) * nib + node] = v; }

  // gpu-octree-cache.cpp:243-293
  int UnloadLRUBlock() {
    int index = lru.front().block;
    CachedBlockInfo& info = contents[index];
    int parent_index = info.parent_block_index;
    for (int i = 0; i < info.roots_count; ++i) {
      uint32_t v = info.pointer_value_in_parent[i];
      uint32_t mask = v & 255, ptr = v >> 11;
      bool fault = (v >> 8) & 1, dup = (v >> 9) & 1;
      if (mask && !fault) v = PackCache(mask, false, dup, false, NearValue(info.pointer_index_in_parent[i], ptr));
      WriteLink(parent_index, info.pointer_index_in_parent[i], v);
      ++patches;
    }
    for (int c : info.child_blocks) block_to_parent_cache.erase(c);
    info.child_blocks.clear();
    lru.pop_front();
    block_to_cache.erase(info.block_index);
    DequeueMarkers();
    ++evicted;
    return index;
  }

  // gpu-octree-cache.cpp:102-234
  void UploadBlock(Block& block) {
    int index;
    if (first_free == max_blocks) {
      if (block.parent != ROOT_BLOCK_PARENT) MarkParentAsUsed(int(block.index), true);
      index = UnloadLRUBlock();
    } else {
      index = first_free++;
    }
    CachedBlockInfo& info = contents[index];
    for (int n = 0; n < nib; ++n) {
      uint32_t link = block.link(n);
      uint32_t mask = link & 255, ptr = link >> 11;
      bool fault = (link >> 8) & 1, dup = (link >> 9) & 1;
      if (mask) {
        if (!fault) {
          block.set_link(n, PackCache(mask, false, dup, false, NearValue(n, ptr)));
        } else if (!block_to_parent_cache.count(int(ptr))) {
          block_to_parent_cache[int(ptr)] = index;
          info.child_blocks.push_back(int(ptr));
        }
      }
    }
    // channel upload
    memcpy(&links[size_t(index) * nib], block.data.data(), size_t(nib) * 4);
    size_t off = size_t(nib) * 4;
    for (size_t c = 1; c < tree->channels.size(); ++c) {
      size_t s = size_t(tree->channels[c].bytes) * nib;
      memcpy(&chans[c - 1][size_t(index) * s], &block.data[off], s);
      off += s;
    }
    ++uploaded;
    if (block.parent != ROOT_BLOCK_PARENT) {
      int parent_index = block_to_parent_cache.at(int(block.index));
      info.parent_block_index = parent_index;
      info.roots_count = int(block.roots_count);
      for (uint32_t i = 0; i < block.roots_count; ++i) {
        info.pointer_index_in_parent[i] = block.parent_ptr_idx[i];
        info.pointer_value_in_parent[i] = block.parent_ptr_value[i];
        int64_t absolute_ptr = int64_t(index) * nib + block.child_idx[i];
        int64_t ptr_index = int64_t(parent_index) * nib + block.parent_ptr_idx[i];
        bool farr;
        uint32_t ptr;
        if (NeedFar(ptr_index, absolute_ptr)) {
          uint32_t fi = uint32_t(index) * 8 + i;
          farr = true;
          ptr = fi;
          far_ptrs[fi] = uint32_t(absolute_ptr);
          ++far_writes;
        } else {
          farr = false;
          ptr = NearValue(ptr_index, absolute_ptr);
        }
        uint32_t pv = block.parent_ptr_value[i];
        uint32_t nv = PackCache(pv & 255, false, !((pv >> 8) & 1), farr, ptr);
        WriteLink(parent_index, block.parent_ptr_idx[i], nv);
        ++patches;
      }
    } else {
      info.parent_block_index = int(ROOT_BLOCK_PARENT);
      info.roots_count = 0;
    }
    block_to_cache[int(block.index)] = index;
    info.block_index = int(block.index);
    info.lru = lru.insert(lru.begin(), LRUEntry{false, index, nullptr});
    MarkBlockAsUsed(index);
  }

  void LoadAndUpload(int index) {
    Block b;
    if (!LoadBlock(tree, uint32_t(index), &b)) throw std::runtime_error("failed to load block");
    bool its_root = b.parent == ROOT_BLOCK_PARENT;
    int root_in_block = b.child_idx[0];
    UploadBlock(b);
    if (its_root) root_node_index = block_to_cache.at(index) * nib + root_in_block;
  }

  // gpu-octree-cache-manager.cpp:60-108
  void InitialFill() {
    std::queue<int> q;
    q.push(int(tree->root_block));
    while (first_free != max_blocks && !q.empty()) {
      int index = q.front();
      q.pop();
      if (block_to_cache.count(index)) continue;
      Block b;
      if (!LoadBlock(tree, uint32_t(index), &b)) throw std::runtime_error("failed to load block");
      bool its_root = b.parent == ROOT_BLOCK_PARENT;
      int root_in_block = b.child_idx[0];
      for (int i = 0; i < nib; ++i) {
        uint32_t link = b.link(i);
        if ((link & 255) && ((link >> 8) & 1)) q.push(int(link >> 11));
      }
      UploadBlock(b);
      if (its_root) root_node_index = block_to_cache.at(index) * nib + root_in_block;
    }
  }
  void StartSession() {
    session = true;
    session_marker.reset(new Marker());
    session_it = InsertMarker(session_marker.get());
  }
  void StartTransaction() {
    transaction = true;
    transaction_marker.reset(new Marker());
    transaction_it = InsertMarker(transaction_marker.get());
  }
  void RequestBlock(int index) {
    if (block_to_cache.count(index)) return;
    transaction_blocks.push_back(index);
    MarkParentAsUsed(index, true);
  }
  bool TransactionFilled() const { return transaction_marker->is_hit; }
  bool SessionFilled() const { return session_marker->is_hit; }
  void EndTransaction() {
    for (size_t i = 0; i < transaction_blocks.size(); ++i) {
      if (TransactionFilled()) break;
      LoadAndUpload(transaction_blocks[i]);
    }
    transaction_blocks.clear();
    transaction = false;
    RemoveMarker(transaction_marker.get(), transaction_it);
    transaction_marker.reset();
  }
  void EndSession() {
    session = false;
    RemoveMarker(session_marker.get(), session_it);
    session_marker.reset();
  }
  const uint32_t* Channel(const char* name) const {
    for (size_t c = 1; c < tree->channels.size(); ++c)
      if (tree->channels[c].name == name && tree->channels[c].bytes == 4)
        return reinterpret_cast<const uint32_t*>(chans[c - 1].data());
    return nullptr;
  }
};

// ------------------------------------------------------------ feedback
static inline uint64_t SplitMix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Feedback {
  int expected_faults = 1 << 16;
  int hits_to_extract = 4000;
  uint64_t rng = 0;
  int unique_count = 0;
  std::vector<int> faults, hits, dup_hits;

  void AddHit(int e) {
    if (e & 1)
      dup_hits.push_back((e >> 1) - 1);
    else
      hits.push_back((e >> 1) - 1);
  }
  // Process() on block-level codes (NodeToBlock already applied).
  void Process(std::vector<int>& codes) {
    std::sort(codes.begin(), codes.end());
    std::vector<int> sorted;
    for (size_t i = 0; i < codes.size(); ++i)
      if (i == 0 || codes[i - 1] != codes[i]) sorted.push_back(codes[i]);
    unique_count = int(sorted.size());
    faults.clear();
    hits.clear();
    dup_hits.clear();
    int stride = expected_faults, size = 0;
    while (size < unique_count && (size == 0 || sorted[size - 1] < 0)) {
      stride = std::min(stride, unique_count - size);
      for (int i = size; i < size + stride; ++i) {
        if (sorted[i] < 0)
          faults.push_back(-sorted[i] - 1);
        else if (sorted[i] != NO_HIT)
          AddHit(sorted[i]);
      }
      size += stride;
      stride *= 2;
    }
    if (size < unique_count) {
      int hits_length = unique_count - size;
      int start = int((SplitMix(rng) >> 33) % uint64_t(hits_length));
      int len = std::min(hits_to_extract, hits_length);
      std::vector<std::pair<int, int>> segs;
      if (start + len <= hits_length)
        segs.push_back({start, len});
      else {
        segs.push_back({0, start + len - hits_length});
        segs.push_back({start, hits_length - start});
      }
      for (auto& s : segs)
        for (int i = 0; i < s.second; ++i) {
          int v = sorted[size + s.first + i];
          if (v != NO_HIT) AddHit(v);
        }
    }
    if (expected_faults > 1 && int(faults.size()) < expected_faults)
      expected_faults /= 2;
    else if (int(faults.size()) > expected_faults)
      expected_faults *= 2;
  }
};

// ------------------------------------------------------------ camera
struct M4 {
  float m[16];
};
static M4 Mul(const M4& a, const M4& b) {
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      float t = 0;
      for (int k = 0; k < 4; ++k) t += a.m[(i << 2) + k] * b.m[(k << 2) + j];
      r.m[(i << 2) + j] = t;
    }
  return r;
}
static M4 Identity() {
  M4 r;
  for (int i = 0; i < 16; ++i) r.m[i] = 0;
  r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1;
  return r;
}
static const float kPI = 3.14159265358979323846f;
static const float kDegToRad = kPI / 180.0f;
static M4 RotX(float a) { float c = cosf(a * kDegToRad), s = sinf(a * kDegToRad); M4 r = Identity(); r.m[5] = c; r.m[6] = -s; r.m[9] = s; r.m[10] = c; return r; }
static M4 RotY(float a) { float c = cosf(a * kDegToRad), s = sinf(a * kDegToRad); M4 r = Identity(); r.m[0] = c; r.m[2] = s; r.m[8] = -s; r.m[10] = c; return r; }
static M4 RotZ(float a) { float c = cosf(a * kDegToRad), s = sinf(a * kDegToRad); M4 r = Identity(); r.m[0] = c; r.m[1] = -s; r.m[4] = s; r.m[5] = c; return r; }
static M4 Transposed(const M4& a) {
  M4 r;
  for (int i = 0; i < 4; i++) {
    r.m[i] = a.m[(i << 2) + 0];
    r.m[i + 4] = a.m[(i << 2) + 1];
    r.m[i + 8] = a.m[(i << 2) + 2];
    r.m[i + 12] = a.m[(i << 2) + 3];
  }
  return r;
}
static M4 Perspective(float fov, float aspect, float nearv, float farv) {
  M4 r;
  for (int i = 0; i < 16; ++i) r.m[i] = 0;
  float f = 1.0f / tanf(fov * (kDegToRad / 2));
  r.m[0] = f;
  r.m[5] = f * aspect;
  r.m[14] = -1;
  if (farv == 0) {
    r.m[10] = -1;
    r.m[11] = -2 * nearv;
  } else {
    r.m[10] = -(farv + nearv) / (farv - nearv);
    r.m[11] = -2 * farv * nearv / (farv - nearv);
  }
  return r;
}
// Cofactor inverse of util/mat.cpp:20-107 (Intel SSE note layout).
static M4 Inverse(const M4& in) {
  M4 res;
  const float* mat = in.m;
  float* dst = res.m;
  float tmp[12], src[16], det;
  for (int i = 0; i < 4; i++) {
    src[i] = mat[(i << 2) + 0];
    src[i + 4] = mat[(i << 2) + 1];
    src[i + 8] = mat[(i << 2) + 2];
    src[i + 12] = mat[(i << 2) + 3];
  }
  tmp[0] = src[10] * src[15]; tmp[1] = src[11] * src[14]; tmp[2] = src[9] * src[15];
  tmp[3] = src[11] * src[13]; tmp[4] = src[9] * src[14]; tmp[5] = src[10] * src[13];
  tmp[6] = src[8] * src[15]; tmp[7] = src[11] * src[12]; tmp[8] = src[8] * src[14];
  tmp[9] = src[10] * src[12]; tmp[10] = src[8] * src[13]; tmp[11] = src[9] * src[12];
  dst[0] = tmp[0] * src[5] + tmp[3] * src[6] + tmp[4] * src[7];
  dst[0] -= tmp[1] * src[5] + tmp[2] * src[6] + tmp[5] * src[7];
  dst[1] = tmp[1] * src[4] + tmp[6] * src[6] + tmp[9] * src[7];
  dst[1] -= tmp[0] * src[4] + tmp[7] * src[6] + tmp[8] * src[7];
  dst[2] = tmp[2] * src[4] + tmp[7] * src[5] + tmp[10] * src[7];
  dst[2] -= tmp[3] * src[4] + tmp[6] * src[5] + tmp[11] * src[7];
  dst[3] = tmp[5] * src[4] + tmp[8] * src[5] + tmp[11] * src[6];
  dst[3] -= tmp[4] * src[4] + tmp[9] * src[5] + tmp[10] * src[6];
  dst[4] = tmp[1] * src[1] + tmp[2] * src[2] + tmp[5] * src[3];
  dst[4] -= tmp[0] * src[1] + tmp[3] * src[2] + tmp[4] * src[3];
  dst[5] = tmp[0] * src[0] + tmp[7] * src[2] + tmp[8] * src[3];
  dst[5] -= tmp[1] * src[0] + tmp[6] * src[2] + tmp[9] * src[3];
  dst[6] = tmp[3] * src[0] + tmp[6] * src[1] + tmp[11] * src[3];
  dst[6] -= tmp[2] * src[0] + tmp[7] * src[1] + tmp[10] * src[3];
  dst[7] = tmp[4] * src[0] + tmp[9] * src[1] + tmp[10] * src[2];
  dst[7] -= tmp[5] * src[0] + tmp[8] * src[1] + tmp[11] * src[2];
  tmp[0] = src[2] * src[7]; tmp[1] = src[3] * src[6]; tmp[2] = src[1] * src[7];
  tmp[3] = src[3] * src[5]; tmp[4] = src[1] * src[6]; tmp[5] = src[2] * src[5];
  tmp[6] = src[0] * src[7]; tmp[7] = src[3] * src[4]; tmp[8] = src[0] * src[6];
  tmp[9] = src[2] * src[4]; tmp[10] = src[0] * src[5]; tmp[11] = src[1] * src[4];
  dst[8] = tmp[0] * src[13] + tmp[3] * src[14] + tmp[4] * src[15];
  dst[8] -= tmp[1] * src[13] + tmp[2] * src[14] + tmp[5] * src[15];
  dst[9] = tmp[1] * src[12] + tmp[6] * src[14] + tmp[9] * src[15];
  dst[9] -= tmp[0] * src[12] + tmp[7] * src[14] + tmp[8] * src[15];
  dst[10] = tmp[2] * src[12] + tmp[7] * src[13] + tmp[10] * src[15];
  dst[10] -= tmp[3] * src[12] + tmp[6] * src[13] + tmp[11] * src[15];
  dst[11] = tmp[5] * src[12] + tmp[8] * src[13] + tmp[11] * src[14];
  dst[11] -= tmp[4] * src[12] + tmp[9] * src[13] + tmp[10] * src[14];
  dst[12] = tmp[2] * src[10] + tmp[5] * src[11] + tmp[1] * src[9];
  dst[12] -= tmp[4] * src[11] + tmp[0] * src[9] + tmp[3] * src[10];
  dst[13] = tmp[8] * src[11] + tmp[0] * src[8] + tmp[7] * src[10];
  dst[13] -= tmp[6] * src[10] + tmp[9] * src[11] + tmp[1] * src[8];
  dst[14] = tmp[6] * src[9] + tmp[11] * src[11] + tmp[3] * src[8];
  dst[14] -= tmp[10] * src[11] + tmp[2] * src[8] + tmp[7] * src[9];
  dst[15] = tmp[10] * src[10] + tmp[4] * src[8] + tmp[9] * src[9];
  dst[15] -= tmp[8] * src[9] + tmp[11] * src[10] + tmp[5] * src[8];
  det = src[0] * dst[0] + src[1] * dst[1] + src[2] * dst[2] + src[3] * dst[3];
  det = 1.0f / det;
  for (int j = 0; j < 16; j++) dst[j] *= det;
  return res;
}

// ------------------------------------------------ restart traversal (legacy)
static inline float min_component(float x, float y, float z) { return cl_fmin(cl_fmin(x, y), z); }
static inline float max_component(float x, float y, float z) { return cl_fmax(cl_fmax(x, y), z); }

static void CastRayRestart(const uint32_t* node_links, int root, F4 origin, F4 dir, float ray_size_coef,
                           float ray_size_bias, float* hit_t, int* hit_node, int* fault_block, int* error_code) {
  *fault_block = -1;
  *error_code = 0;
  const float epsilon = 1e-6f;
  if (fabsf(dir.x) < epsilon) dir.x = copysignf(epsilon, dir.x);
  if (fabsf(dir.y) < epsilon) dir.y = copysignf(epsilon, dir.y);
  if (fabsf(dir.z) < epsilon) dir.z = copysignf(epsilon, dir.z);
  float cx = 1.0f / fabsf(dir.x), cy = 1.0f / fabsf(dir.y), cz = 1.0f / fabsf(dir.z);
  float bx = cx * origin.x, by = cy * origin.y, bz = cz * origin.z;
  unsigned octant_mask = 0;
  if (dir.x < 0.0f) octant_mask ^= 1, bx = cx - bx;
  if (dir.y < 0.0f) octant_mask ^= 2, by = cy - by;
  if (dir.z < 0.0f) octant_mask ^= 4, bz = cz - bz;
  float t_max = min_component(cx - bx, cy - by, cz - bz);
  float t_min = cl_fmax(0.0f, max_component(-bx, -by, -bz));
  int steps = 0;
  while (t_max > t_min) {
    uint32_t node = uint32_t(root);
    float px = 0, py = 0, pz = 0, size = 1;
    float ray_size = ray_size_bias + ray_size_coef * t_min;
    while (1) {
      if (++steps > 5000) {
        *error_code = 1;
        return;
      }
      if (size < ray_size) {
        *hit_t = t_min;
        *hit_node = int(node);
        return;
      }
      uint32_t link = node_links[node];
      uint32_t children_mask = link & 255;
      if (!children_mask) {
        *hit_t = t_min;
        *hit_node = int(node);
        return;
      }
      size = float(double(size) * 0.5);
      float nx = px + size, ny = py + size, nz = pz + size;
      float tcx = nx * cx - bx, tcy = ny * cy - by, tcz = nz * cz - bz;
      unsigned index = octant_mask;
      if (tcx < t_min) index ^= 1, px = nx;
      if (tcy < t_min) index ^= 2, py = ny;
      if (tcz < t_min) index ^= 4, pz = nz;
      unsigned index_mask = (1u << index) & 255u;
      if (!(children_mask & index_mask)) {
        t_min = min_component((px + size) * cx - bx, (py + size) * cy - by, (pz + size) * cz - bz) + epsilon;
        break;
      }
      if (link & (1 << 8)) {
        *hit_t = t_min;
        *hit_node = int(node);
        *fault_block = int(link >> 9);
        return;
      }
      node = (link >> 9) + uint32_t(popcount8(children_mask & (index_mask - 1)));
    }
  }
  *hit_t = 1e30f;
  *hit_node = -1;
}

struct NearFar {
  F4 origin, dir;
  float coef, bias;
};
// Near-plane ray of raytracer.cl:29-50 / raytracer_volatile.cl:134-150.
static NearFar NearPlaneRay(const float* m, int x, int y, int W, int H, float lod) {
  F4 tmp = {(x + 0.5f) / float(W) * 2 - 1, (y + 0.5f) / float(H) * 2 - 1, -1, 0};
  F4 origin = mul(m, tmp);
  tmp.z = 0;
  F4 direction = normalize4(sub4(mul(m, tmp), origin));
  tmp = {float(x) / float(W) * 2 - 1, float(y) / float(H) * 2 - 1, -1, 0};
  F4 p1near = mul(m, tmp);
  tmp.z = 0;
  F4 p1far = mul(m, tmp);
  tmp = {float(x + 1) / float(W) * 2 - 1, float(y + 1) / float(H) * 2 - 1, -1, 0};
  F4 p2near = mul(m, tmp);
  tmp.z = 0;
  F4 p2far = mul(m, tmp);
  NearFar r;
  r.origin = origin;
  r.dir = direction;
  r.bias = lod * distance4(p1near, p2near);
  r.coef = (r.bias < 1e-10f) ? 0 : (lod * (distance4(p1far, p2far) - distance4(p1near, p2near)) / distance4(p1near, p1far));
  return r;
}

// ln t in double with a fixed operation order, rounded to float: the
// reduction t = m * 2^e, m in [sqrt(1/2), sqrt(2)), and the series
// ln m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^23/23), s = (m - 1) / (m + 1)
// (Abramowitz & Stegun 4.1.27). The fixed order makes host and GPU agree to
// the bit (a libm log may differ by an ulp); tests pin it against libm.
static float LnFixed(float t) {
  if (!(t > 0.0f)) return t == 0.0f ? -std::numeric_limits<float>::infinity() : std::nanf("");
  if (std::isinf(t)) return t;
  double dt = double(t);
  uint64_t bits;
  memcpy(&bits, &dt, 8);
  int e = int((bits >> 52) & 0x7ff) - 1023;
  uint64_t mb = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m;
  memcpy(&m, &mb, 8);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 23.0;
  const double k[11] = {1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0, 1.0 / 15.0, 1.0 / 13.0, 1.0 / 11.0,
                        1.0 / 9.0,  1.0 / 7.0,  1.0 / 5.0,  1.0 / 3.0,  1.0};
  for (double c : k) p = p * z + c;
  return float(double(e) * 0.6931471805599453094 + 2.0 * (s * p));
}

static uint32_t DepthGrey(float t) {
  const float ln_near = -4.60517025f, ln_far = 1.60943794f;
  float lt = LnFixed(t);
  float d = (lt - ln_near) / (ln_far - ln_near);
  d = (1.0f - d) * 255.0f;
  uint32_t g = uint32_t(cl_max(0.0f, cl_min(255.0f, d)));
  return g | (g << 8) | (g << 16) | (255u << 24);
}

}  // namespace orc

using namespace orc;

extern "C" {

// ---- trees
void* orc_tree_open(const char* path) { return OpenTree(path); }
void orc_tree_close(void* t) { delete static_cast<Tree*>(t); }
int orc_tree_info(void* tp, uint32_t out[4]) {
  Tree* t = static_cast<Tree*>(tp);
  out[0] = t->nib; out[1] = t->blocks; out[2] = t->root_block; out[3] = uint32_t(t->channels.size());
  return 0;
}
// Loads block `index`: header76 fields decoded into hdr[3 + 8*3] ints, payload into data.
int orc_tree_load_block(void* tp, uint32_t index, uint32_t* hdr, uint8_t* data, uint64_t cap) {
  Tree* t = static_cast<Tree*>(tp);
  Block b;
  if (!LoadBlock(t, index, &b)) return -1;
  if (cap < b.data.size()) return -2;
  hdr[0] = b.index; hdr[1] = b.parent; hdr[2] = b.roots_count;
  for (int i = 0; i < 8; ++i) {
    hdr[3 + i * 3] = b.parent_ptr_idx[i];
    hdr[4 + i * 3] = b.child_idx[i];
    hdr[5 + i * 3] = b.parent_ptr_value[i];
  }
  memcpy(data, b.data.data(), b.data.size());
  return 0;
}

// ---- cache
void* orc_cache_create(void* tree, int max_blocks) { return new Cache(static_cast<Tree*>(tree), max_blocks); }
void orc_cache_destroy(void* c) { delete static_cast<Cache*>(c); }
int orc_cache_initial_fill(void* c) {
  try { static_cast<Cache*>(c)->InitialFill(); } catch (...) { return -1; }
  return 0;
}
int orc_cache_root(void* c) { return static_cast<Cache*>(c)->root_node_index; }
int orc_cache_loaded(void* c) { return static_cast<Cache*>(c)->first_free; }
// which: 0 links, 1 far pointers, 2+k channel k+1 (bytes). Returns element/byte count copied.
int64_t orc_cache_copy(void* cp, int which, void* out, uint64_t cap_bytes) {
  Cache* c = static_cast<Cache*>(cp);
  const void* src;
  uint64_t n;
  if (which == 0) { src = c->links.data(); n = c->links.size() * 4; }
  else if (which == 1) { src = c->far_ptrs.data(); n = c->far_ptrs.size() * 4; }
  else {
    size_t k = size_t(which - 2);
    if (k >= c->chans.size()) return -1;
    src = c->chans[k].data(); n = c->chans[k].size();
  }
  if (cap_bytes < n) return -2;
  memcpy(out, src, n);
  return int64_t(n);
}
int orc_cache_block_slot(void* cp, int block) {
  Cache* c = static_cast<Cache*>(cp);
  auto it = c->block_to_cache.find(block);
  return it == c->block_to_cache.end() ? -1 : it->second;
}
int orc_cache_lru_order(void* cp, int* out, int cap) {
  Cache* c = static_cast<Cache*>(cp);
  int n = 0;
  for (auto& e : c->lru) {
    if (n < cap) out[n] = e.is_marker ? -1 : e.block;
    ++n;
  }
  return n;
}
void orc_cache_counters(void* cp, uint64_t out[4]) {
  Cache* c = static_cast<Cache*>(cp);
  out[0] = c->uploaded; out[1] = c->evicted; out[2] = c->patches; out[3] = c->far_writes;
}
// Session / transaction protocol, for hand-worked LRU sequences.
int orc_cache_op(void* cp, int op, int arg) {
  Cache* c = static_cast<Cache*>(cp);
  try {
    switch (op) {
      case 0: c->StartSession(); return 0;
      case 1: c->StartTransaction(); return 0;
      case 2: c->MarkBlockAsUsed(arg); return 0;
      case 3: c->MarkParentAsUsed(arg, false); return 0;
      case 4: c->RequestBlock(arg); return 0;
      case 5: return c->TransactionFilled() ? 1 : 0;
      case 6: return c->SessionFilled() ? 1 : 0;
      case 7: c->EndTransaction(); return 0;
      case 8: c->EndSession(); return 0;
    }
  } catch (...) {
    return -1;
  }
  return -2;
}

// ---- single-view renders over explicit pool arrays
// Fused InitTracingFrame + RaytracingPass + FinishTracingFrame for one pass.
// hit_t / hit_node / fetches_per_ray may be null. stats: {primary, shadow, fetches, hits}.
static int RenderPlainPool(const Pool& P, int nib, int W, int H, const float* vpi, float lod, const float* bg,
                           int flags, uint32_t* out_image, int32_t* codes, float* hit_t, int32_t* hit_node,
                           uint32_t* fetches_per_ray, uint64_t* stats, int threads);

int orc_render_plain(const uint32_t* links, const uint32_t* far_ptrs, int root, const uint32_t* normals,
                     const uint32_t* colors, int nib, int W, int H, const float* vpi, float lod, const float* bg,
                     int flags, uint32_t* out_image, int32_t* codes, float* hit_t, int32_t* hit_node,
                     uint32_t* fetches_per_ray, uint64_t* stats, int threads, uint64_t n_links, uint64_t n_far) {
  Pool P{links, far_ptrs, n_links, n_far, root, normals, colors};
  return RenderPlainPool(P, nib, W, H, vpi, lod, bg, flags, out_image, codes, hit_t, hit_node, fetches_per_ray,
                         stats, threads);
}

// The same render over the oracle cache's pool as it stands (no copy of the pool arrays).
int orc_cache_render_plain(void* cp, int W, int H, const float* vpi, float lod, const float* bg, int flags,
                           uint32_t* out_image, int32_t* codes, uint64_t* stats, int threads) {
  Cache* c = static_cast<Cache*>(cp);
  Pool P{c->links.data(), c->far_ptrs.data(), c->links.size(), c->far_ptrs.size(), c->root_node_index,
         c->Channel("normals"), c->Channel("colors")};
  return RenderPlainPool(P, c->nib, W, H, vpi, lod, bg, flags, out_image, codes, nullptr, nullptr, nullptr, stats,
                         threads);
}

static int RenderPlainPool(const Pool& P, int nib, int W, int H, const float* vpi, float lod, const float* bg,
                           int flags, uint32_t* out_image, int32_t* codes, float* hit_t, int32_t* hit_node,
                           uint32_t* fetches_per_ray, uint64_t* stats, int threads) {
  const Lights L = MakeLights();
  int nt = Threads(threads);
  std::vector<uint64_t> st(size_t(nt) * 4, 0);
  ParallelRows(H, nt, [&](int y, int t) {
    for (int x = 0; x < W; ++x) {
      size_t i = size_t(y) * W + x;
      Ray r = InitRay(vpi, x, y, W, H, lod);
      TracingState s{r.origin, r.direction, -1, 0u, 0xffffffffu, 0};
      uint64_t f = 0, sh = 0;
      RayCastResult res;
      int code = RaytracingPassPixel(P, s, nib, flags, L, &f, &sh, &res);
      out_image[i] = FinishPixel(P, s, bg, L);
      if (codes) codes[i] = code;
      if (hit_t) hit_t[i] = res.hit_t;
      if (hit_node) hit_node[i] = res.error_code ? -3 : res.hit_node;
      if (fetches_per_ray) fetches_per_ray[i] = uint32_t(f);
      st[t * 4 + 0] += 1;
      st[t * 4 + 1] += sh;
      st[t * 4 + 2] += f;
      st[t * 4 + 3] += (res.error_code == 0 && res.fault_block == -1 && res.hit_node >= 0) ? 1 : 0;
    }
  });
  if (stats) {
    for (int k = 0; k < 4; ++k) stats[k] = 0;
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 4; ++k) stats[k] += st[t * 4 + k];
  }
  return 0;
}

// Separate kernels: InitTracingFrame (state arrays are 12 floats/ints per pixel, AoS).
int orc_init_tracing_frame(float* states, int W, int H, const float* vpi, float lod) {
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      Ray r = InitRay(vpi, x, y, W, H, lod);
      TracingState s{r.origin, r.direction, -1, 0u, 0xffffffffu, 0};
      memcpy(&states[(size_t(y) * W + x) * 12], &s, 48);
    }
  return 0;
}
int orc_raytracing_pass(int32_t* codes, float* states, int W, int H, const uint32_t* links, const uint32_t* far_ptrs,
                        int root, const uint32_t* normals, const uint32_t* colors, int nib, int flags, int threads,
                        uint64_t n_links, uint64_t n_far) {
  Pool P{links, far_ptrs, n_links, n_far, root, normals, colors};
  const Lights L = MakeLights();
  ParallelRows(H, Threads(threads), [&](int y, int) {
    for (int x = 0; x < W; ++x) {
      size_t i = size_t(y) * W + x;
      TracingState s;
      memcpy(&s, &states[i * 12], 48);
      uint64_t f = 0, sh = 0;
      codes[i] = RaytracingPassPixel(P, s, nib, flags, L, &f, &sh, nullptr);
      memcpy(&states[i * 12], &s, 48);
    }
  });
  return 0;
}
int orc_finish_tracing_frame(uint32_t* out, const float* states, int W, int H, const float* bg,
                             const uint32_t* normals, const uint32_t* colors) {
  Pool P{nullptr, nullptr, 0, 0, 0, normals, colors};
  const Lights L = MakeLights();
  for (size_t i = 0; i < size_t(W) * H; ++i) {
    TracingState s;
    memcpy(&s, &states[i * 12], 48);
    out[i] = FinishPixel(P, s, bg, L);
  }
  return 0;
}

// LnFixed over an array (pinned by the tests against libm).
void orc_ln_fixed(const float* in, float* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = LnFixed(in[i]);
}

// RaytraceKernel of raytracer.cl (log-depth grey). hit_node may be null.
int orc_raytrace_depth(const uint32_t* links, const uint32_t* far_ptrs, int root, int W, int H, const float* vpi,
                       float lod, uint32_t* out, int32_t* hit_node, int threads, uint64_t n_links,
                       uint64_t n_far) {
  ParallelRows(H, Threads(threads), [&](int y, int) {
    for (int x = 0; x < W; ++x) {
      NearFar r = NearPlaneRay(vpi, x, y, W, H, lod);
      RayCastResult res;
      uint64_t f = 0;
      CastRay(links, n_links, far_ptrs, n_far, root, r.origin, r.dir, r.coef, r.bias, &res, &f);
      size_t i = size_t(y) * W + x;
      if (res.error_code || res.fault_block != -1)
        out[i] = rgba(255, 0, 0, 255);
      else if (res.hit_node == -1)
        out[i] = rgba(0, 0, 0, 255);
      else
        out[i] = DepthGrey(res.hit_t);
      if (hit_node) hit_node[i] = res.error_code ? -3 : (res.fault_block != -1 ? -2 : res.hit_node);
    }
  });
  return 0;
}

// RaytraceKernel of raytracer_volatile.cl over legacy links.
int orc_raytrace_restart(const uint32_t* legacy_links, int root, int W, int H, const float* vpi, float lod,
                         uint32_t* out, int32_t* hit_node, int threads) {
  ParallelRows(H, Threads(threads), [&](int y, int) {
    for (int x = 0; x < W; ++x) {
      NearFar r = NearPlaneRay(vpi, x, y, W, H, lod);
      float t = 0;
      int node = -1, fb = -1, err = 0;
      CastRayRestart(legacy_links, root, r.origin, r.dir, r.coef, r.bias, &t, &node, &fb, &err);
      size_t i = size_t(y) * W + x;
      if (err) {
        out[i] = rgba(255, 0, 0, 255);
      } else {
        float ht = cl_min(1.0f, t / 5.0f);
        float d = 255.0f - ht * 255.0f;
        uint32_t g = uint32_t(d);
        out[i] = rgba(g, g, g, 255);
      }
      if (hit_node) hit_node[i] = err ? -3 : node;
    }
  });
  return 0;
}

// Old link layout (raytracer_volatile.cl: ptr = link >> 9) from a cache pool.
int orc_legacy_links(const uint32_t* links, const uint32_t* far_ptrs, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t l = links[i], mask = l & 255, o = 0;
    if (mask) {
      uint64_t target;
      if (l & (1u << 8)) {
        target = l >> 11;
        o = 1u << 8;
      } else if (l & (1u << 10)) {
        target = far_ptrs[l >> 11];
      } else {
        target = uint64_t(int64_t(i) + int64_t(l >> 11) - (1 << 20));
      }
      if (target >= (1ull << 23)) return -1;
      o |= uint32_t(target << 9) | mask;
    }
    out[i] = o;
  }
  return 0;
}

// NodeToBlockKernel.
void orc_node_to_block(int n, int block_size, int32_t* data) {
  for (int i = 0; i < n; ++i) {
    int d = data[i];
    if (d > 0 && d != NO_HIT) data[i] = ((((d >> 1) - 1) / block_size + 1) << 1) | (d & 1);
  }
}

// Feedback extraction on block-level codes. lists: faults, hits, dup hits
// (caps per list); counts[3] receive the lengths, state[0..3] = {expected_faults,
// hits_to_extract, rng (lo), rng (hi)} in/out, state[4] = unique_count out.
int orc_feedback_extract(const int32_t* codes, int n, int64_t* state, int32_t* faults, int32_t* hits, int32_t* dups,
                         int cap, int32_t* counts) {
  Feedback fb;
  fb.expected_faults = int(state[0]);
  fb.hits_to_extract = int(state[1]);
  fb.rng = uint64_t(state[2]);
  std::vector<int> v(codes, codes + n);
  fb.Process(v);
  state[0] = fb.expected_faults;
  state[2] = int64_t(fb.rng);
  state[4] = fb.unique_count;
  counts[0] = int(fb.faults.size());
  counts[1] = int(fb.hits.size());
  counts[2] = int(fb.dup_hits.size());
  for (int i = 0; i < std::min(cap, counts[0]); ++i) faults[i] = fb.faults[i];
  for (int i = 0; i < std::min(cap, counts[1]); ++i) hits[i] = fb.hits[i];
  for (int i = 0; i < std::min(cap, counts[2]); ++i) dups[i] = fb.dup_hits[i];
  return 0;
}

// Full RenderFrame (gpu-ray-tracer.cpp:66-227) against the oracle cache.
// max_passes: 0 = no cap; otherwise break after pass i if i > 0 && i + 1 >= max_passes
// (the deterministic stand-in for the wall-clock limit of :124).
// pass_faults receives each pass's fault list concatenated (cap entries),
// pass_counts the per-pass lengths (cap_passes). Returns the number of passes, <0 on error.
int orc_render_frame(void* cp, int W, int H, const float* vpi, float lod, const float* bg, int flags,
                     int max_passes, int hits_to_extract, uint64_t* rng_state, int* expected_faults,
                     uint32_t* out_image, int32_t* last_codes, int32_t* pass_counts, int cap_passes,
                     int32_t* pass_faults, int cap_faults, uint64_t* stats, int threads) {
  Cache* c = static_cast<Cache*>(cp);
  const size_t P = size_t(W) * H;
  std::vector<TracingState> states(P);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      Ray r = InitRay(vpi, x, y, W, H, lod);
      states[size_t(y) * W + x] = TracingState{r.origin, r.direction, -1, 0u, 0xffffffffu, 0};
    }
  Feedback fb;
  fb.hits_to_extract = hits_to_extract;
  fb.rng = *rng_state;
  fb.expected_faults = *expected_faults;
  const Lights L = MakeLights();
  std::vector<int> codes(P);
  int nt = Threads(threads);
  int iterations, nfaults_total = 0;
  uint64_t rays = 0;
  try {
    c->StartSession();
    for (iterations = 0;; ++iterations) {
      Pool Pl{c->links.data(), c->far_ptrs.data(), c->links.size(), c->far_ptrs.size(), c->root_node_index,
              c->Channel("normals"), c->Channel("colors")};
      std::vector<uint64_t> cnt(nt, 0);
      ParallelRows(H, nt, [&](int y, int t) {
        for (int x = 0; x < W; ++x) {
          size_t i = size_t(y) * W + x;
          uint64_t f = 0, sh = 0;
          bool active = states[i].color_multiplier != 0;
          codes[i] = RaytracingPassPixel(Pl, states[i], c->nib, flags, L, &f, &sh, nullptr);
          cnt[t] += (active ? 1 : 0) + sh;
        }
      });
      for (auto v : cnt) rays += v;
      if (last_codes) memcpy(last_codes, codes.data(), P * 4);
      if (iterations > 0 && max_passes > 0 && iterations + 1 >= max_passes) break;
      std::vector<int> work(codes);
      fb.Process(work);
      c->StartTransaction();
      for (int h : fb.hits) c->MarkBlockAsUsed(h);
      for (int h : fb.dup_hits) c->MarkParentAsUsed(h, false);
      for (size_t i = 0; i < fb.faults.size(); ++i) {
        if (c->TransactionFilled()) break;
        c->RequestBlock(fb.faults[i]);
      }
      if (iterations < cap_passes) pass_counts[iterations] = int(fb.faults.size());
      for (int f : fb.faults)
        if (nfaults_total < cap_faults) pass_faults[nfaults_total++] = f;
      c->EndTransaction();
      if (fb.faults.empty() || c->SessionFilled()) break;
    }
    if (stats) {
      stats[0] = uint64_t(iterations + 1);
      stats[1] = c->uploaded;
      stats[2] = c->SessionFilled() ? 1 : 0;
      stats[3] = rays;
    }
    c->EndSession();
  } catch (...) {
    return -1;
  }
  Pool Pl{c->links.data(), c->far_ptrs.data(), c->links.size(), c->far_ptrs.size(), c->root_node_index,
          c->Channel("normals"), c->Channel("colors")};
  for (size_t i = 0; i < P; ++i) out_image[i] = FinishPixel(Pl, states[i], bg, L);
  *rng_state = fb.rng;
  *expected_faults = fb.expected_faults;
  return iterations + 1;
}

// Camera::ViewProjectionMatrix().Inverse().
void orc_camera_vpi(float yaw, float pitch, float roll, const float* pos, float fov, float nearv, float aspect,
                    float* out) {
  M4 t = Identity();
  t.m[3] = -pos[0];
  t.m[7] = -pos[1];
  t.m[11] = -pos[2];
  M4 rot = Mul(Mul(RotY(-yaw), RotX(pitch)), RotZ(-roll));
  M4 vp = Mul(Mul(Perspective(fov, aspect, nearv, 0), Transposed(rot)), t);
  M4 inv = Inverse(vp);
  memcpy(out, inv.m, 64);
}

// ---- sphere structure checks (stored-octree-test.cpp:39-233)
// StoredOctreeTraverser (CPU data/stored-octree-traverser.cpp:33-120): every
// block of a .tree file loaded, each block's roots attached to their parent
// node in the parent block (whose stored link must equal the root's
// parent_pointer_value), non-fault links followed inside a block, and every
// node reachable from the root required to have its children resolved
// (CheckTree). The reference crash()es and assert()s; the restatement
// records the first error instead:
//   1 a block fails to load, 2 a root's parent node is missing, 3 a parent
//   link differs from parent_pointer_value, 4 a node is reached twice,
//   5 a fault remains after loading every block, 6 no root block.
struct TravNode {
  bool resolved = false;               // children_ non-empty in the reference
  std::array<TravNode*, 8> children{};  // null where the mask has no child
  std::vector<uint8_t> data;            // the channels after node links
  uint32_t initial_link = 0;
};

struct Traverser {
  Tree* tree;
  std::map<std::pair<uint32_t, uint32_t>, std::unique_ptr<TravNode>> nodes;
  std::set<uint32_t> loaded;
  TravNode* root = nullptr;
  int error = 0;

  explicit Traverser(Tree* t) : tree(t) {
    for (uint32_t i = 0; i < t->blocks && !error; ++i) LoadIfNotLoaded(i);
    if (!error && !root) error = 6;
    if (!error) CheckTree(root);
  }
  void Fail(int e) {
    if (!error) error = e;
  }
  void LoadIfNotLoaded(uint32_t index) {
    if (error || loaded.count(index)) return;
    Block b;
    if (!LoadBlock(tree, index, &b)) return Fail(1);
    if (b.parent == ROOT_BLOCK_PARENT) {
      root = TraverseBlock(b, b.child_idx[0]);
    } else {
      LoadIfNotLoaded(b.parent);
      if (error) return;
      for (uint32_t ri = 0; ri < b.roots_count; ++ri) {
        auto it = nodes.find({b.parent, b.parent_ptr_idx[ri]});
        if (it == nodes.end()) return Fail(2);
        TravNode* n = it->second.get();
        if (n->initial_link != b.parent_ptr_value[ri]) return Fail(3);
        n->children.fill(nullptr);
        n->resolved = true;
        const uint32_t mask = b.parent_ptr_value[ri] & 255;
        int p = 0;
        for (int i = 0; i < 8; ++i)
          if (mask & (1u << i)) n->children[i] = TraverseBlock(b, b.child_idx[ri] + p++);
        if (error) return;
      }
    }
    loaded.insert(index);
  }
  TravNode* TraverseBlock(Block& b, uint32_t node_index) {
    if (error) return nullptr;
    if (node_index >= tree->nib) {
      Fail(2);
      return nullptr;
    }
    auto key = std::make_pair(b.index, node_index);
    if (nodes.count(key)) {
      Fail(4);
      return nullptr;
    }
    TravNode* n = (nodes[key] = std::unique_ptr<TravNode>(new TravNode())).get();
    size_t off = size_t(tree->nib) * 4;
    for (size_t c = 1; c < tree->channels.size(); ++c) {
      const int s = tree->channels[c].bytes;
      n->data.insert(n->data.end(), &b.data[off + size_t(s) * node_index], &b.data[off + size_t(s) * (node_index + 1)]);
      off += size_t(s) * tree->nib;
    }
    const uint32_t link = b.link(int(node_index));
    n->initial_link = link;
    const uint32_t mask = link & 255, ptr = link >> 11;
    const bool fault = (link >> 8) & 1;
    if (!mask) {
      n->resolved = true;
    } else if (!fault) {
      n->resolved = true;
      int p = 0;
      for (int i = 0; i < 8; ++i)
        if (mask & (1u << i)) n->children[i] = TraverseBlock(b, ptr + p++);
    }
    return n;
  }
  void CheckTree(TravNode* n) {
    if (error) return;
    if (!n->resolved) return Fail(5);
    for (TravNode* c : n->children)
      if (c) CheckTree(c);
  }
};

// DebugCheckConstantDepthTree's CheckSubtree (experimental/debug-check-tree.cpp:9-35):
// counts the nodes below `depth_left` 0 ("too deep") and the leaves above it
// ("not deep enough").
static void ConstantDepthRec(const TravNode* n, int depth_left, uint64_t* checked, uint64_t* errors) {
  ++*checked;
  if (depth_left < 0) ++*errors;
  bool has_children = false;
  for (const TravNode* c : n->children)
    if (c) {
      has_children = true;
      ConstantDepthRec(c, depth_left - 1, checked, errors);
    }
  if (!has_children && depth_left > 0) ++*errors;
}

static uint64_t CountNodes(const TravNode* n) {
  uint64_t k = 1;
  for (const TravNode* c : n->children)
    if (c) k += CountNodes(c);
  return k;
}

struct SphereCheck {
  float cx, cy, cz, r2;
  int max_depth;
  const uint32_t* links = nullptr;
  const uint32_t* far_ptrs = nullptr;
  const uint8_t* ch1 = nullptr;  // 4-byte node index
  const uint8_t* ch2 = nullptr;  // 1-byte node index % 211
  bool faultless = false;
  int error = 0;
};

static bool Straddle(const SphereCheck& s, const float* minp, const float* maxp) {
  bool inside = s.cx > minp[0] && s.cy > minp[1] && s.cz > minp[2] && s.cx < maxp[0] && s.cy < maxp[1] && s.cz < maxp[2];
  if (inside) return true;
  bool was_in = false, was_out = false;
  for (int i = 0; i < 8; ++i) {
    float v[3] = {(i & 1) ? maxp[0] : minp[0], (i & 2) ? maxp[1] : minp[1], (i & 4) ? maxp[2] : minp[2]};
    float dx = v[0] - s.cx, dy = v[1] - s.cy, dz = v[2] - s.cz;
    if (dx * dx + dy * dy + dz * dz < s.r2) {
      was_in = true;
      if (was_out) break;
    } else {
      was_out = true;
      if (was_in) break;
    }
  }
  return was_in && was_out;
}

// Walks the analytic sphere against a cache pool; cache_index -1 = no node, -2 = behind a fault.
static void CheckPoolRec(SphereCheck& s, int64_t cache_index, int depth, const float* minp, const float* maxp,
                         int32_t node_id) {
  if (s.error) return;
  bool present = Straddle(s, minp, maxp);
  if (!present) {
    if (cache_index >= 0) s.error = 1;  // unexpected node in GPU
    return;
  }
  if (cache_index == -1) {
    s.error = 2;  // node expected in GPU
    return;
  }
  if (depth < s.max_depth) {
    int64_t ncache = -100;
    uint32_t gmask = 0;
    if (cache_index == -2) {
      ncache = -2;
    } else {
      uint32_t link = s.links[cache_index];
      gmask = link & 255;
      if ((link >> 8) & 1) {
        if (s.faultless) { s.error = 3; return; }
        ncache = -2;
      } else if ((link >> 10) & 1) {
        ncache = int64_t(s.far_ptrs[link >> 11]) - 1;
      } else {
        ncache = cache_index + int64_t(link >> 11) - (1 << 20) - 1;
      }
    }
    float midp[3] = {(minp[0] + maxp[0]) * 0.5f, (minp[1] + maxp[1]) * 0.5f, (minp[2] + maxp[2]) * 0.5f};
    for (int i = 0; i < 8; ++i) {
      float nmin[3] = {minp[0], minp[1], minp[2]}, nmax[3] = {midp[0], midp[1], midp[2]};
      if (i & 1) nmin[0] = midp[0], nmax[0] = maxp[0];
      if (i & 2) nmin[1] = midp[1], nmax[1] = maxp[1];
      if (i & 4) nmin[2] = midp[2], nmax[2] = maxp[2];
      int64_t nci;
      if (ncache == -2) nci = -2;
      else if (gmask & (1u << i)) nci = ++ncache;
      else nci = -1;
      CheckPoolRec(s, nci, depth + 1, nmin, nmax, int32_t(uint32_t(node_id) * 8u + uint32_t(i)));
      if (s.error) return;
    }
  } else if (cache_index != -2) {
    if (s.links[cache_index] & 255) { s.error = 4; return; }
  }
  if (cache_index >= 0 && s.ch1) {
    uint32_t id;
    memcpy(&id, s.ch1 + 4 * cache_index, 4);
    if (id != uint32_t(node_id)) { s.error = 5; return; }
    if (int8_t(s.ch2[cache_index]) != int8_t(char(node_id % 211))) { s.error = 6; return; }
  }
}

// CheckTestOctreeSphereWithGPUData on host copies of the pool. 0 = pass.
int orc_check_sphere_pool(const uint32_t* links, const uint32_t* far_ptrs, const uint8_t* ch1, const uint8_t* ch2,
                          int root, float cx, float cy, float cz, float radius, int level, int faultless) {
  SphereCheck s;
  s.cx = cx; s.cy = cy; s.cz = cz; s.r2 = radius * radius; s.max_depth = level;
  s.links = links; s.far_ptrs = far_ptrs; s.ch1 = ch1; s.ch2 = ch2; s.faultless = faultless != 0;
  float mn[3] = {0, 0, 0}, mx[3] = {1, 1, 1};
  CheckPoolRec(s, root, 0, mn, mx, 1);
  return s.error;
}

// StoredOctreeTraverser construction + CheckTree. 0 = pass (error codes at
// Traverser); *nodes = nodes reachable from the root.
int orc_traverse_check(void* tp, uint64_t* nodes) {
  Traverser tr(static_cast<Tree*>(tp));
  if (nodes) *nodes = tr.error ? 0 : CountNodes(tr.root);
  return tr.error;
}

// DebugCheckConstantDepthTree (debug-check-tree.cpp:37-47): the traverser
// then CheckSubtree. Returns the traverser's error as a negative number, else
// the number of depth errors (0 = "checked ok").
int64_t orc_check_constant_depth(void* tp, int expected_depth, uint64_t* checked) {
  Traverser tr(static_cast<Tree*>(tp));
  if (tr.error) return -tr.error;
  uint64_t c = 0, errors = 0;
  ConstantDepthRec(tr.root, expected_depth, &c, &errors);
  if (checked) *checked = c;
  return int64_t(errors);
}

// CheckTestOctreeSphereWithLoader (scene sources/stored-octree-test.cpp:191-210):
// the file's channels must be node links(4), node index(4), node index %
// 211(1); then the analytic sphere is walked against the traverser: a node
// exactly where a cell straddles the surface, no children at the full depth,
// and the node index channels equal to the generator's ids. 0 = pass;
// 20 + the traverser's error; 11 wrong channels, 12 unexpected node,
// 13 missing node, 14 children at full depth, 15 / 16 channel data differs.
struct SphereTreeCheck {
  SphereCheck s;
  int error = 0;
};
static void CheckTreeRec(SphereTreeCheck& c, const TravNode* n, int depth, const float* minp, const float* maxp,
                         int32_t node_id) {
  if (c.error) return;
  if (!Straddle(c.s, minp, maxp)) {
    if (n) c.error = 12;
    return;
  }
  if (!n) {
    c.error = 13;
    return;
  }
  if (depth < c.s.max_depth) {
    float midp[3] = {(minp[0] + maxp[0]) * 0.5f, (minp[1] + maxp[1]) * 0.5f, (minp[2] + maxp[2]) * 0.5f};
    for (int i = 0; i < 8 && !c.error; ++i) {
      float nmin[3] = {minp[0], minp[1], minp[2]}, nmax[3] = {midp[0], midp[1], midp[2]};
      if (i & 1) nmin[0] = midp[0], nmax[0] = maxp[0];
      if (i & 2) nmin[1] = midp[1], nmax[1] = maxp[1];
      if (i & 4) nmin[2] = midp[2], nmax[2] = maxp[2];
      CheckTreeRec(c, n->children[i], depth + 1, nmin, nmax, int32_t(uint32_t(node_id) * 8u + uint32_t(i)));
    }
  } else {
    for (const TravNode* ch : n->children)
      if (ch) c.error = 14;
  }
  if (c.error) return;
  uint32_t id;
  memcpy(&id, n->data.data(), 4);
  if (id != uint32_t(node_id)) c.error = 15;
  else if (int8_t(n->data[4]) != int8_t(char(node_id % 211))) c.error = 16;
}

int orc_check_sphere_tree(void* tp, float cx, float cy, float cz, float radius, int level) {
  Tree* t = static_cast<Tree*>(tp);
  if (t->channels.size() != 3 || t->channels[0].name != "node links" || t->channels[0].bytes != 4 ||
      t->channels[1].name != "node index" || t->channels[1].bytes != 4 || t->channels[2].name != "node index % 211" ||
      t->channels[2].bytes != 1)
    return 11;
  Traverser tr(t);
  if (tr.error) return 20 + tr.error;
  SphereTreeCheck c;
  c.s.cx = cx; c.s.cy = cy; c.s.cz = cz; c.s.r2 = radius * radius; c.s.max_depth = level;
  float mn[3] = {0, 0, 0}, mx[3] = {1, 1, 1};
  CheckTreeRec(c, tr.root, 0, mn, mx, 1);
  return c.error;
}

}  // extern "C"
