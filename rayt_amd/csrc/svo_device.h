// Device-side sparse-voxel-octree traversal for gfx950.
//
// Semantics follow the reference kernels operation for operation so that hit
// node ids are bit-identical to the CPU restatement:
//   CastRay          GPU_renderer/kernels/raycast.cl:33-256
//   InitTracingFrame GPU_renderer/kernels/init_tracing_frame.cl:3-45
//   ProcessHit       GPU_renderer/kernels/shading.cl:4-33
//   FinishTracingFrame GPU_renderer/kernels/finish_tracing_frame.cl:3-24
// FP32 rules (fixed here because OpenCL leaves them to the implementation):
// no contraction (built with -ffp-contract=off), correctly rounded / and sqrt,
// dot = ((x*a + y*b) + z*c) + w*d, normalize(v) = v / sqrt(dot(v,v)),
// distance(a,b) = sqrt(dot(a-b,a-b)), fmax/fmin/max/min as the OpenCL
// definitions "y if x<y else x" / "y if y<x else x".
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtdev {

constexpr int kSMax = 23;                  // raycast.cl:10
constexpr float kEpsilon = 0.00000011920928955078125f;  // exp2(-S_MAX), raycast.cl:11
constexpr int kIterLimitSteps = 300;          // raycast.cl:95
constexpr int kNoHit = 0x7fffffff;         // common.cl:14
constexpr int kStackLevels = kSMax - 1;    // levels 0..21 (scale 22..1)

__device__ __forceinline__ float cl_fmax(float x, float y) {
  if (x != x) return y;
  return (x < y) ? y : x;
}
__device__ __forceinline__ float cl_fmin(float x, float y) {
  if (x != x) return y;
  return (y < x) ? y : x;
}
__device__ __forceinline__ float cl_max(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float cl_min(float x, float y) { return (y < x) ? y : x; }

struct CastResult {
  float hit_t;
  int hit_node;
  int fault_block;
  int hit_duplicate;
  int error_code;
  int node_depth;
};

// Traversal stack indexed by level = (S_MAX-1) - scale. LDS_LEVELS levels,
// starting at level `base`, live in LDS (one column per thread, lanes
// contiguous so a wave touches consecutive 8-byte words); the other levels
// live in a private array that is only touched for trees deeper than the LDS
// window. The private array is kept outside the struct (`spill` points to a
// StackSpill of the kernel) so that the struct itself stays in registers and
// the LDS accesses compile to ds_read/ds_write.
template <int LDS_LEVELS>
struct StackSpill {
  uint2 v[LDS_LEVELS >= kStackLevels ? 1 : kStackLevels];
};

//
// SPILL = false is the window-only stack, for trees the caller declares to
// fit the window (depth - 1 <= LDS_LEVELS, base 0): no private array and no
// branch per access. A level outside the window is never stored (the write
// is masked, the read clamped into the window) and `put` reports it, and the traversal then
// ends the ray with the iteration-limit error, so a tree deeper than declared
// shows error pixels instead of wrong ones.
template <int LDS_LEVELS, int BLOCK, bool SPILL = true>
struct SplitStack {
  uint2* lds;             // base of this workgroup's [LDS_LEVELS][BLOCK] array, offset by thread
  int top = kSMax - 1;    // (kSMax - 1) - first level held in LDS: a scale's LDS slot is top - scale
  uint2* spill = nullptr;  // StackSpill<LDS_LEVELS>::v of this thread (SPILL only)
  __device__ __forceinline__ void set_base(int first_lds_level) { top = (kSMax - 1) - first_lds_level; }
  // Returns false when a window-only stack could not store the level.
  __device__ __forceinline__ bool put(int scale, uint2 v) {
    if constexpr (LDS_LEVELS >= kStackLevels) {
      lds[((kSMax - 1) - scale) * BLOCK] = v;
      return true;
    } else if constexpr (!SPILL) {
      const unsigned l = unsigned(top - scale);
      const bool fits = l < unsigned(LDS_LEVELS);
      if (fits) lds[l * BLOCK] = v;
      return fits;
    } else {
      const int l = top - scale;
      if (unsigned(l) < unsigned(LDS_LEVELS)) {
        lds[l * BLOCK] = v;
      } else {
        spill[(kSMax - 1) - scale] = v;
        // keeps the two stores apart: merged, they become one flat store
        // through a selected pointer instead of a ds_write
        __asm__ volatile("" ::: "memory");
      }
      return true;
    }
  }
  __device__ __forceinline__ uint2 get(int scale) const {
    if constexpr (LDS_LEVELS >= kStackLevels) {
      return lds[((kSMax - 1) - scale) * BLOCK];
    } else if constexpr (!SPILL) {
      return lds[min(unsigned(top - scale), unsigned(LDS_LEVELS - 1)) * BLOCK];
    } else {
      const int l = top - scale;
      uint2 r;
      if (unsigned(l) < unsigned(LDS_LEVELS)) {
        r = lds[l * BLOCK];
      } else {
        r = spill[(kSMax - 1) - scale];
        __asm__ volatile("" ::: "memory");
      }
      return r;
    }
  }
};

// The block pool as seen by the traversal: node links and far pointers with
// their element counts. A link that points outside either array ends the ray
// with error code 2 (a corrupt pool never turns into a memory fault).
struct PoolView {
  const uint32_t* __restrict__ links;
  const uint32_t* __restrict__ far_ptrs;
  uint32_t nodes;
  uint32_t far_count;
  // Loads through global-address-space pointers: where the pool pointers are
  // read from memory (one pool per view) the compiler cannot tell their
  // address space and would emit FLAT loads, which also count against the
  // LDS wait counter and so stall the stack reads behind node fetches.
  __device__ __forceinline__ uint32_t link(uint32_t i) const {
    return ((const __attribute__((address_space(1))) uint32_t*)links)[i];
  }
  __device__ __forceinline__ uint32_t far(uint32_t i) const {
    return ((const __attribute__((address_space(1))) uint32_t*)far_ptrs)[i];
  }
};

// ESVO traversal over the cache link format, as a resumable state machine:
// TravInit is the set-up before the loop of raycast.cl:42-95 and TravStep is
// one iteration of the loop of raycast.cl:100-250 (including the loop test).
// CastRay runs the steps to the end; the batched kernel interleaves steps of
// different rays in one lane. `fetches` counts executed node-link and
// far-pointer loads (the F of the roofline accounting).
//
// The reference reloads the parent's descriptor at the top of an iteration
// whenever it is zero, which happens at the start and after every pop
// (raycast.cl:115). Here the load is issued where the zero would be set (at
// the start and at the end of a pop), so its latency overlaps the rest of the
// step; the loaded word is the same. For the fetch count, `pending_fetch`
// carries the load to the next iteration's top, after the iteration-limit
// test, where the reference performs it; a descriptor that is still zero is
// counted again at every iteration, as the reference reloads it.
struct Trav {
  float tx_coef, ty_coef, tz_coef, tx_bias, ty_bias, tz_bias;
  float ray_size_coef, ray_size_bias;
  float t_min, t_max, h;
  float posx, posy, posz, scale_exp2;
  uint32_t parent, child_descriptor, idx, octant_mask;
  int scale, iter;
  uint32_t pending_fetch;
};

__device__ __forceinline__ void TravInit(Trav& s, const PoolView& pool, uint32_t root_node, float px, float py,
                                         float pz, float dx, float dy, float dz, float ray_size_coef,
                                         float ray_size_bias) {
  px += 1.0f;
  py += 1.0f;
  pz += 1.0f;
  if (fabsf(dx) < kEpsilon) dx = copysignf(kEpsilon, dx);
  if (fabsf(dy) < kEpsilon) dy = copysignf(kEpsilon, dy);
  if (fabsf(dz) < kEpsilon) dz = copysignf(kEpsilon, dz);
  s.tx_coef = 1.0f / -fabsf(dx);
  s.ty_coef = 1.0f / -fabsf(dy);
  s.tz_coef = 1.0f / -fabsf(dz);
  s.tx_bias = s.tx_coef * px;
  s.ty_bias = s.ty_coef * py;
  s.tz_bias = s.tz_coef * pz;
  s.octant_mask = 0;
  if (dx > 0.0f) s.octant_mask ^= 1u, s.tx_bias = 3.0f * s.tx_coef - s.tx_bias;
  if (dy > 0.0f) s.octant_mask ^= 2u, s.ty_bias = 3.0f * s.ty_coef - s.ty_bias;
  if (dz > 0.0f) s.octant_mask ^= 4u, s.tz_bias = 3.0f * s.tz_coef - s.tz_bias;
  s.t_min = cl_fmax(cl_fmax(2.0f * s.tx_coef - s.tx_bias, 2.0f * s.ty_coef - s.ty_bias), 2.0f * s.tz_coef - s.tz_bias);
  s.t_max = cl_fmin(cl_fmin(s.tx_coef - s.tx_bias, s.ty_coef - s.ty_bias), s.tz_coef - s.tz_bias);
  s.h = s.t_max;
  s.t_min = cl_fmax(s.t_min, 0.0f);
  s.parent = root_node;
  s.child_descriptor = pool.link(root_node);
  s.pending_fetch = 1;
  s.idx = 0;
  s.posx = s.posy = s.posz = 1.0f;
  s.scale = kSMax - 1;
  s.scale_exp2 = 0.5f;
  if (1.5f * s.tx_coef - s.tx_bias > s.t_min) s.idx ^= 1u, s.posx = 1.5f;
  if (1.5f * s.ty_coef - s.ty_bias > s.t_min) s.idx ^= 2u, s.posy = 1.5f;
  if (1.5f * s.tz_coef - s.tz_bias > s.t_min) s.idx ^= 4u, s.posz = 1.5f;
  s.iter = 0;
  s.ray_size_coef = ray_size_coef;
  s.ray_size_bias = ray_size_bias;
}

// How a traversal ended (TravStep returns kRunning while it goes on).
enum TravStatus : int {
  kRunning = 0,
  kMiss = 1,       // left the octree (raycast.cl:254-255)
  kIterLimit = 2,  // iteration limit (raycast.cl:104-110), error 1
  kFault = 3,      // children not loaded (raycast.cl:134-141)
  kBadLink = 4,    // far or child pointer outside the pool, or below the finest scale: error 2
  kHit = 5,        // leaf or LOD stop (raycast.cl:159-165); `parent` then holds the hit node
};

// One loop iteration of raycast.cl:100-250. The conditions of the reference's
// branches are evaluated as predicates and the state is updated with selects;
// only the three node loads (the parent descriptor after a pop, a far
// pointer, the child descriptor) and the stack accesses are guarded. Every
// way out of the loop only sets the status, which TravResult decodes once, so
// the loop around the step has a single exit. The reference's loop test
// (scale < S_MAX) is evaluated where the scale grows, at the end of a pop,
// and ends the ray in that step.
//
// LOD = false drops the LOD stop test (raycast.cl:159-165) for rays whose
// LOD coefficient and bias come from a zero LOD voxel size (lod * x is 0 or
// NaN) and for shadow rays (coefficient and bias 0): tc_max * coef + bias is
// then 0 or NaN, never >= scale_exp2 (> 0), so the test is always false.
//
// LANE_ITER = false leaves the iteration limit (raycast.cl:104-110) to the
// caller: CastRay starts every ray of a wave on the same loop trip, so one
// wave-wide trip counter (a scalar register) stands for every lane's count.
template <class Stack, bool COUNT, bool LOD = true, bool LANE_ITER = true>
__device__ __forceinline__ int TravStep(Trav& s, const PoolView& pool, Stack& stack, uint32_t& fetches) {
  if (LANE_ITER && ++s.iter > kIterLimitSteps) return kIterLimit;
  if (COUNT) {
    fetches += (s.pending_fetch || s.child_descriptor == 0) ? 1u : 0u;
    s.pending_fetch = 0;
  }
  const uint32_t desc = s.child_descriptor;
  const float tx_corner = s.posx * s.tx_coef - s.tx_bias;
  const float ty_corner = s.posy * s.ty_coef - s.ty_bias;
  const float tz_corner = s.posz * s.tz_coef - s.tz_bias;
  // fmin of raycast.cl as one v_min_f32 (IEEE minNum). minNum returns the
  // other operand when one is NaN, as OpenCL fmin does; it may differ from
  // fmin only on (+0, -0), and a corner is never -0: posx * coef is nonzero
  // (posx >= 1, |coef| >= 1) and x - y rounds to +0 when x == y.
  const float tc_max = __builtin_fminf(__builtin_fminf(tx_corner, ty_corner), tz_corner);

  // The child slot's bit of the mask byte, and the mask bits below it (the
  // children stored before this one), as bit-field extracts: the same values
  // as (desc & (1 << slot)) and desc & ((1 << slot) - 1) & 255 for slot < 8.
  const uint32_t slot = s.idx ^ s.octant_mask;
  const bool want = __builtin_amdgcn_ubfe(desc, slot, 1) != 0 && s.t_min <= s.t_max;
  // the fault bit (8) and the far bit (10) in one test: a far link without
  // the fault bit reads 0x400, a near one 0, anything else is a fault
  const uint32_t kind = desc & 0x500u;
  const bool far_kind = kind == 0x400u, near_kind = kind == 0u;
  const bool fault = want & !(far_kind | near_kind);
  const uint32_t ptr = desc >> 11;
  const bool is_far = want & far_kind;
  const bool far_bad = is_far && ptr >= pool.far_count;
  uint32_t ofs = ptr + s.parent - (1u << 20);
  if (is_far && !far_bad) {
    ofs = pool.far(ptr);
    if (COUNT) ++fetches;
  }
  ofs += __popc(__builtin_amdgcn_ubfe(desc, 0, slot));
  // non-short-circuit forms: lane-mask ands and ors, no select through vector registers
  const bool bad = far_bad | (want & !fault & ((ofs >= pool.nodes) | (s.scale < 1)));
  const bool fetch = want && !fault && !bad;
  // the child word goes straight into the descriptor (the lanes that load it
  // do not pop, and the others keep theirs or reload it in the pop below)
  if (fetch) {
    s.child_descriptor = pool.link(ofs);
    if (COUNT) ++fetches;
  }
  int status = fault ? int(kFault) : bad ? int(kBadLink) : int(kRunning);

  // The lanes that advance do not read the child word: their update comes
  // first, while the child loads of the other lanes are in flight.
  if (!want) {
    const bool sx = tx_corner <= tc_max, sy = ty_corner <= tc_max, sz = tz_corner <= tc_max;
    const uint32_t step_mask = (sx ? 1u : 0u) | (sy ? 2u : 0u) | (sz ? 4u : 0u);
    const uint32_t ox = __float_as_uint(s.posx), oy = __float_as_uint(s.posy), oz = __float_as_uint(s.posz);
    s.posx = sx ? s.posx - s.scale_exp2 : s.posx;
    s.posy = sy ? s.posy - s.scale_exp2 : s.posy;
    s.posz = sz ? s.posz - s.scale_exp2 : s.posz;
    s.t_min = tc_max;
    s.idx ^= step_mask;
    if ((s.idx & step_mask) != 0) {
      // pop. Highest differing bit of the stepped coordinates: the reference
      // reads it from the exponent of (float)differing_bits; the two agree
      // whenever the result is below S_MAX (the value is then exact in fp32),
      // and both leave the loop otherwise. The reference differs each stepped
      // coordinate from itself plus the cell size, which is its value before
      // the step: the coordinates are multiples of the cell size (at least
      // 2^-23) in [0.5, 2), where the subtraction and the addition are
      // exact. A coordinate that did not step equals its old value and adds
      // no bits.
      const uint32_t diff = (__float_as_uint(s.posx) ^ ox) | (__float_as_uint(s.posy) ^ oy) |
                            (__float_as_uint(s.posz) ^ oz);
      // diff is nonzero here: a stepped coordinate (below 2) changes by the
      // cell size (at least 2^-23, scale >= 0)
      const int scale = 31 - __builtin_clz(diff);
      s.scale = scale;
      if (scale >= kSMax) {
        status = kMiss;  // the ray left the root cube: the loop test (raycast.cl:250, 254-255)
      } else {
        s.scale_exp2 = __uint_as_float(uint32_t(scale - kSMax + 127) << 23);
        const uint2 e = stack.get(scale);
        s.parent = e.x;
        s.t_max = __uint_as_float(e.y);
        // the coordinates cleared below the restored scale; the child index from their bits at that scale
        const uint32_t keep = ~0u << scale;
        const uint32_t px = __float_as_uint(s.posx) & keep, py = __float_as_uint(s.posy) & keep,
                       pz = __float_as_uint(s.posz) & keep;
        s.posx = __uint_as_float(px);
        s.posy = __uint_as_float(py);
        s.posz = __uint_as_float(pz);
        s.idx = __builtin_amdgcn_ubfe(px, scale, 1) | (__builtin_amdgcn_ubfe(py, scale, 1) << 1) |
                (__builtin_amdgcn_ubfe(pz, scale, 1) << 2);
        s.h = 0.0f;
        s.child_descriptor = pool.link(s.parent);
        s.pending_fetch = 1;
      }
    }
  }

  const uint32_t child = s.child_descriptor;  // the loaded child word on the lanes that fetched
  bool stop = (child & 255u) == 0;  // a leaf
  if constexpr (LOD) stop = stop || tc_max * s.ray_size_coef + s.ray_size_bias >= s.scale_exp2;
  const bool hit = fetch && stop;
  const bool push = fetch && !hit;
  if (hit) status = kHit;
  if (hit) s.parent = ofs;

  if (push) {
    const float half = s.scale_exp2 * 0.5f;
    const float tx_center = half * s.tx_coef + tx_corner;
    const float ty_center = half * s.ty_coef + ty_corner;
    const float tz_center = half * s.tz_coef + tz_corner;
    // deeper than a window-only stack holds: the ray ends as the iteration
    // limit does (error 1), with the state the next step would report it from
    if (tc_max < s.h && !stack.put(s.scale, make_uint2(s.parent, __float_as_uint(s.t_max)))) status = kIterLimit;
    s.h = tc_max;
    s.parent = ofs;
    s.scale--;
    s.scale_exp2 = half;
    const bool cx = tx_center > s.t_min, cy = ty_center > s.t_min, cz = tz_center > s.t_min;
    s.idx = (cx ? 1u : 0u) | (cy ? 2u : 0u) | (cz ? 4u : 0u);
    s.posx = cx ? s.posx + half : s.posx;
    s.posy = cy ? s.posy + half : s.posy;
    s.posz = cz ? s.posz + half : s.posz;
    s.t_max = tc_max;
  }
  return status;
}

// The RayCastResult of a finished traversal (raycast.cl:104-165, 254-255).
__device__ __forceinline__ void TravResult(const Trav& s, int status, CastResult& res) {
  res.fault_block = status == kFault ? int(s.child_descriptor >> 11) : -1;
  res.error_code = status == kIterLimit ? 1 : status == kBadLink ? 2 : 0;
  res.hit_duplicate = status == kHit ? int((s.child_descriptor >> 9) & 1u) : 0;
  res.hit_t = status == kMiss ? 1e30f : s.t_min;
  res.hit_node = status == kMiss ? -1 : int(s.parent);
  res.node_depth = status == kMiss ? 0 : status == kHit ? kSMax - s.scale : kSMax - s.scale - 1;
}

template <class Stack, bool COUNT, bool LOD = true>
__device__ __forceinline__ void CastRay(const PoolView& pool, uint32_t root_node, float px, float py, float pz,
                                        float dx, float dy, float dz, float ray_size_coef, float ray_size_bias,
                                        Stack& stack, CastResult& res, uint32_t& fetches) {
  Trav s;
  TravInit(s, pool, root_node, px, py, pz, dx, dy, dz, ray_size_coef, ray_size_bias);
  int status;
  int trips = 0;  // the same for every active lane: they all start here together
  do {
    if (++trips > kIterLimitSteps) {
      status = kIterLimit;
      break;
    }
    status = TravStep<Stack, COUNT, LOD, false>(s, pool, stack, fetches);
  } while (status == kRunning);
  TravResult(s, status, res);
}

// mul(float16, float3) of init_tracing_frame.cl:3-9: (x,y,z,1) against rows,
// divided by the fourth row's dot product; the w output is 0/d.
__device__ __forceinline__ float4 MulPoint(const float* m, float x, float y, float z) {
  const float d = ((x * m[12] + y * m[13]) + z * m[14]) + 1.0f * m[15];
  const float r0 = ((x * m[0] + y * m[1]) + z * m[2]) + 1.0f * m[3];
  const float r1 = ((x * m[4] + y * m[5]) + z * m[6]) + 1.0f * m[7];
  const float r2 = ((x * m[8] + y * m[9]) + z * m[10]) + 1.0f * m[11];
  return make_float4(r0 / d, r1 / d, r2 / d, 0.0f / d);
}

__device__ __forceinline__ float Dot4(float4 a, float4 b) {
  return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
}

struct PrimaryRay {
  float4 origin;     // xyz origin, w = ray size bias
  float4 direction;  // xyz direction, w = ray size coefficient
};

// InitTracingFrame for pixel (x, y) of a W x H view.
__device__ __forceinline__ PrimaryRay InitRay(const float* m, int x, int y, int W, int H, float lod_voxel_size) {
  const float fw = float(W), fh = float(H);
  const float u = (float(x) + 0.5f) / fw * 2.0f - 1.0f;
  const float v = (float(y) + 0.5f) / fh * 2.0f - 1.0f;
  const float4 origin = make_float4(m[2] / m[14], m[6] / m[14], m[10] / m[14], 0.0f / m[14]);
  const float4 p = MulPoint(m, u, v, 0.0f);
  float4 dir = make_float4(p.x - origin.x, p.y - origin.y, p.z - origin.z, p.w - origin.w);
  const float len = sqrtf(Dot4(dir, dir));
  dir = make_float4(dir.x / len, dir.y / len, dir.z / len, dir.w / len);

  const float4 p1 = MulPoint(m, float(x) / fw * 2.0f - 1.0f, float(y) / fh * 2.0f - 1.0f, 0.0f);
  const float4 p2 = MulPoint(m, float(x + 1) / fw * 2.0f - 1.0f, float(y + 1) / fh * 2.0f - 1.0f, 0.0f);
  const float4 d12 = make_float4(p1.x - p2.x, p1.y - p2.y, p1.z - p2.z, p1.w - p2.w);
  const float4 d01 = make_float4(origin.x - p1.x, origin.y - p1.y, origin.z - p1.z, origin.w - p1.w);
  const float coef = lod_voxel_size * sqrtf(Dot4(d12, d12)) / sqrtf(Dot4(d01, d01));

  PrimaryRay r;
  r.origin = make_float4(origin.x, origin.y, origin.z, 0.0f);
  r.direction = make_float4(dir.x, dir.y, dir.z, coef);
  return r;
}

// Light vectors of shading.cl:8-10, each normalize(v) * k with the pinned
// normalize; computed once per kernel argument set on the host and passed in.
struct Lights {
  float l1[3], l2[3], l3[3];
  float shadow_dir[3];  // unit direction towards light 1 (= -normalize(2,-2,3))
};

__device__ __forceinline__ uint32_t PackRGBA(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  return r | (g << 8) | (b << 16) | (a << 24);
}

// Light-1 term (before the max) for the char4 normal; used to decide whether a
// shadow ray is worth casting.
__device__ __forceinline__ float LightTerm(float nx, float ny, float nz, const float* L) {
  return (((-nx) * L[0] + (-ny) * L[1]) + (-nz) * L[2]) + (-0.0f) * 0.0f;
}

// ProcessHit colour. `occluded1` drops light 1 (shadow-ray extension; false
// reproduces the reference exactly).
__device__ __forceinline__ uint32_t ShadeColor(uint32_t cnormal, uint32_t ccolor, const Lights& lt, bool occluded1) {
  const float nx = float(int8_t(cnormal & 255u)) / 127.f;
  const float ny = float(int8_t((cnormal >> 8) & 255u)) / 127.f;
  const float nz = float(int8_t((cnormal >> 16) & 255u)) / 127.f;
  const float m1 = occluded1 ? 0.0f : cl_max(0.f, LightTerm(nx, ny, nz, lt.l1));
  const float m2 = cl_max(0.f, LightTerm(nx, ny, nz, lt.l2));
  const float m3 = cl_max(0.f, LightTerm(nx, ny, nz, lt.l3));
  const float light = ((0.1f + m1 * 1.0f) + m2 * 1.0f) + m3 * 1.0f;
  const float cr = cl_min(255.f, float(ccolor & 255u) * light);
  const float cg = cl_min(255.f, float((ccolor >> 8) & 255u) * light);
  const float cb = cl_min(255.f, float((ccolor >> 16) & 255u) * light);
  return PackRGBA(uint32_t(cr), uint32_t(cg), uint32_t(cb), 255u);
}

// FinishTracingFrame blend of colour (RGBA8) over the background.
__device__ __forceinline__ uint32_t Blend(uint32_t color, const float* bg) {
  const float w = float(color >> 24);
  const float k = 255.f - w;
  const float tr = (float(color & 255u) + bg[0] * k) + 0.5f;
  const float tg = (float((color >> 8) & 255u) + bg[1] * k) + 0.5f;
  const float tb = (float((color >> 16) & 255u) + bg[2] * k) + 0.5f;
  const float ta = (w + bg[3] * k) + 0.5f;
  return PackRGBA(uint32_t(cl_max(0.f, cl_min(255.f, tr))), uint32_t(cl_max(0.f, cl_min(255.f, tg))),
                  uint32_t(cl_max(0.f, cl_min(255.f, tb))), uint32_t(cl_max(0.f, cl_min(255.f, ta))));
}

}  // namespace rtdev
