/*
 * rayt_hip.h - C ABI of the MI355X (gfx950) sparse-voxel-octree renderer.
 *
 * This is the seam that replaces the reference's OpenCL wrapper layer
 * ("CL util": CLContext / CLBuffer / CLKernel / CLEvent) and exposes the
 * renderer, block cache and feedback pipeline that sit on it. Every entry
 * point returns 0 on success or a negative RT_ERR_* code; rt_last_error()
 * gives the message of the last failure on the calling thread. The reference
 * turns every such failure into crash() (util/common.h:43-53).
 *
 * Citations are reference paths (file:line) of the interface each entry
 * point replaces. Sizes and offsets are 64-bit (the reference's are int).
 */
#ifndef RAYT_HIP_H_
#define RAYT_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_OK 0
#define RT_ERR_INVALID (-1)
#define RT_ERR_HIP (-2)
#define RT_ERR_IO (-3)
#define RT_ERR_STATE (-4)
#define RT_ERR_NOMEM (-5)
#define RT_ERR_RANGE (-6)

typedef struct rt_ctx rt_ctx;       /* device + stream            (CLContext)  */
typedef struct rt_buf rt_buf;       /* device buffer               (CLBuffer)   */
typedef struct rt_evt rt_evt;       /* completion event            (CLEvent)    */
typedef struct rt_tree rt_tree;     /* block source                (StoredOctreeLoader) */
typedef struct rt_cache rt_cache;   /* device block pool + LRU     (GPUOctreeCacheManager) */
typedef struct rt_feedback rt_feedback; /* feedback extraction     (RayTracingFeedbackExtractor) */
typedef struct rt_tracer rt_tracer; /* frame renderer              (GPURayTracer) */
typedef struct rt_batch rt_batch;   /* many views in one launch (no reference counterpart) */
typedef struct rt_server rt_server; /* continuous-batching request loop (no reference counterpart) */

/* Message of the last failed call on this thread ("" if none). */
const char* rt_last_error(void);
/* Library build identification, e.g. "rayt_hip gfx950". */
const char* rt_version(void);

/* ------------------------------------------------------------------ device */
/* CLContext() + queue creation: CL util/cl-context.cpp:10-207. One HIP stream per context. */
int rt_ctx_create(int device, rt_ctx** out);
int rt_ctx_destroy(rt_ctx* ctx);
/* CLContext::WaitForAll (= clFinish): CL util/cl-context.h:16. */
int rt_ctx_finish(rt_ctx* ctx);
/* Number of compute units of the context's device. */
int rt_ctx_compute_units(rt_ctx* ctx, int* out);

/* CLBuffer(flags, size, ctx): CL util/cl-buffer.cpp:14-21 (64-bit size). */
int rt_buf_create(rt_ctx* ctx, uint64_t bytes, rt_buf** out);
int rt_buf_destroy(rt_buf* buf);
uint64_t rt_buf_size(const rt_buf* buf);
/* Raw device address (for callers that bring their own kernels). */
void* rt_buf_device_ptr(const rt_buf* buf);
/* CLBuffer::Write(start, length, data, blocking, wait_list, out_event): cl-buffer.h:22.
 * Non-blocking writes need `src` to stay valid until the next finish (as in the reference). */
int rt_buf_write(rt_buf* buf, uint64_t offset, uint64_t length, const void* src, int blocking,
                 rt_evt* const* wait, int nwait, rt_evt** out_event);
/* CLBuffer::Read(start, length, data, wait_list, out_event) (non-blocking): cl-buffer.h:24;
 * blocking != 0 gives the blocking debug Read of cl-buffer.h:27. */
int rt_buf_read(const rt_buf* buf, uint64_t offset, uint64_t length, void* dst, int blocking,
                rt_evt* const* wait, int nwait, rt_evt** out_event);
/* CLBuffer::CopyFrom(buf, src_start, dest_start, length, ...): cl-buffer.h:19. */
int rt_buf_copy(rt_buf* dst, uint64_t dst_offset, const rt_buf* src, uint64_t src_offset, uint64_t length,
                rt_evt* const* wait, int nwait, rt_evt** out_event);
/* Fill with a 32-bit pattern (no reference counterpart; used by tests). */
int rt_buf_fill_u32(rt_buf* buf, uint64_t offset_words, uint64_t count, uint32_t value);

/* CLEvent::WaitFor / CLEventList::WaitFor: CL util/cl-event.h:37,55. */
int rt_evt_wait(int n, rt_evt* const* evts);
/* CLEvent copy / destructor reference counting: cl-event.cpp:8-60. */
int rt_evt_retain(rt_evt* e);
int rt_evt_release(rt_evt* e);
/* Profiling: elapsed ns between two completed events (PROFILE_CL_EVENT, cl-event.cpp:61-90). */
int rt_evt_elapsed_ns(rt_evt* start, rt_evt* end, double* out_ns);
/* Records an event on the context's stream (marker). */
int rt_evt_record(rt_ctx* ctx, rt_evt** out);

/* ----------------------------------------------------- renderer kernels
 * Typed launches of the reference kernels (CLKernel::SetArg + Run2D/Run1D,
 * CL util/cl-kernel.cpp:64-164). view_proj_inv is the row-major fmat4.m of
 * Camera::ViewProjectionMatrix().Inverse(). Tracing state is held in
 * `states` (48 bytes per pixel, structure of arrays: float4 origin[P],
 * float4 direction[P], int4 misc[P]).
 */
#define RT_STATE_BYTES_PER_PIXEL 48
/* InitTracingFrame: GPU_renderer/kernels/init_tracing_frame.cl:11-45. */
int rt_init_tracing_frame(rt_ctx* ctx, rt_buf* states, int width, int height, const float view_proj_inv[16],
                          float lod_voxel_size);
/* RaytracingPass: GPU_renderer/kernels/raytracing_pass.cl:4-61. normals/colors may be NULL only when
 * no ray can hit. nodes_in_block == 0 writes the reference's node-level hit codes ((node+1)<<1|dup);
 * nodes_in_block > 0 writes block-level codes ((node/nodes_in_block+1)<<1|dup), i.e. NodeToBlockKernel fused.
 * flags: RT_FLAG_SHADOWS casts a shadow ray per lit hit. The traversal stack follows rt_plain_config. */
#define RT_FLAG_SHADOWS 1
int rt_raytracing_pass(rt_ctx* ctx, rt_buf* faults_and_hits, rt_buf* states, int width, int height,
                       const rt_buf* node_links, const rt_buf* far_pointers, int root_node_index,
                       const rt_buf* normals, const rt_buf* colors, int nodes_in_block, int flags);
/* FinishTracingFrame: GPU_renderer/kernels/finish_tracing_frame.cl:3-24. */
int rt_finish_tracing_frame(rt_ctx* ctx, rt_buf* out_image, const rt_buf* states, int width, int height,
                            const float background[4], const rt_buf* normals, const rt_buf* colors);
/* Single-pass primary ray with log-depth grey output: GPU_renderer/kernels/raytracer.cl:16-72. */
int rt_raytrace_depth(rt_ctx* ctx, rt_buf* out_image, int width, int height, const rt_buf* node_links,
                      const rt_buf* far_pointers, int root_node_index, const float view_proj_inv[16],
                      float lod_voxel_size);
/* Stackless restart traversal over the old link layout (ptr = link>>9, no far/dup bits):
 * raytracer_volatile.cl:25-177. legacy_links from rt_legacy_links_from_cache. */
int rt_raytrace_restart(rt_ctx* ctx, rt_buf* out_image, int width, int height, const rt_buf* legacy_links,
                        int root_node_index, const float view_proj_inv[16], float lod_voxel_size);
/* Converts a fault-free resident cache pool into the old link layout; fails with RT_ERR_RANGE if a
 * child index needs more than 23 bits. */
int rt_legacy_links_from_cache(rt_ctx* ctx, rt_buf* legacy_links, const rt_buf* node_links,
                               const rt_buf* far_pointers, uint64_t node_count);

/* Plain single-view render, fused InitTracingFrame + RaytracingPass + FinishTracingFrame for one pass
 * (no reference counterpart as one kernel; output identical to the three kernels for one pass).
 * codes (may be NULL) receives block-level feedback codes when nodes_in_block > 0.
 * stats (may be NULL, host, 4 x uint64): {primary rays, shadow rays, node fetches, hits} - filled
 * only by the counting variant (flags | RT_FLAG_COUNT), which is slower and not meant for timing. */
#define RT_FLAG_COUNT 2
int rt_render_plain(rt_ctx* ctx, rt_buf* out_image, rt_buf* codes, int width, int height,
                    const rt_buf* node_links, const rt_buf* far_pointers, int root_node_index,
                    const rt_buf* normals, const rt_buf* colors, int nodes_in_block,
                    const float view_proj_inv[16], float lod_voxel_size, const float background[4], int flags,
                    uint64_t* stats);
/* Launch shape of the plain kernel (and of the tracer's resident single-pass frames) on this context, which
 * rt_raytracing_pass also uses (its workgroups: 16x16 pixels for wg_waves 4, 8x8 for 1):
 * lds_stack_levels (8, 10, 12, 14, 16, 22) traversal-stack levels held in LDS, wg_waves 1 (one wave per
 * workgroup, 8x8 pixels) or 4 (16x16 pixels), tree_depth of the scene (0 = unknown) to keep the deepest
 * levels in LDS. When tree_depth - 1 <= lds_stack_levels the whole stack lives in LDS (no private spill
 * array). The declared depth must not be smaller than the tree's: a primary ray that would go deeper ends
 * with the iteration-limit error (red pixel, HIT_CODE_NO_HIT code), a shadow ray as unoccluded; no access
 * leaves the stack. Default: 16 levels, 4 waves, depth unknown. Results do not depend on the shape. */
int rt_plain_config(rt_ctx* ctx, int lds_stack_levels, int wg_waves, int tree_depth);

/* ------------------------------------------------------ GPU algorithms
 * NodeToBlockKernel: GPU_renderer/kernels/feedback_compression.cl:4-13. */
int rt_node_to_block(rt_ctx* ctx, int n, int block_size, rt_buf* data);
/* DifferenceFlagKernel: feedback_compression.cl:15-21. */
int rt_difference_flag(rt_ctx* ctx, int n, const rt_buf* in, rt_buf* out);
/* CompressKernel: feedback_compression.cl:23-34. */
int rt_compress(rt_ctx* ctx, int n, const rt_buf* index, const rt_buf* in, rt_buf* out);
/* GPUSort::Sort - ascending sort of signed int32 in place: GPU algorithms/gpu-sort.cpp:16-49.
 * temp must hold n int32. */
int rt_sort_i32(rt_ctx* ctx, rt_buf* data, int n, rt_buf* temp);
/* GPUScan::Scan - inclusive prefix sum of int32 in place: GPU algorithms/gpu-scan.cpp:15-43. */
int rt_scan_i32(rt_ctx* ctx, rt_buf* data, int n, rt_buf* temp);
/* GlobalMergeSortPassKernel(N, length, in, out): GPU algorithms/kernels/merge_sort.cl:47-73. Merges
 * the sorted runs of `length` (a power of two) of `in` pairwise into `out` (stable). For the CLKernel
 * shim, which runs the reference's GPUSort pass by pass. */
int rt_merge_pass_i32(rt_ctx* ctx, int n, int length, const rt_buf* in, rt_buf* out);
/* ScanPassKernel(N, length, in, out): GPU algorithms/kernels/scan.cl:2-12. One Hillis-Steele pass. */
int rt_scan_pass_i32(rt_ctx* ctx, int n, int length, const rt_buf* in, rt_buf* out);
/* No counterpart in the reference. rt_sort_i32 that allocates nothing, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same kernels and
 * result, with the ping-pong buffer and tile histograms in a static device
 * array of each GPU (about 36 MiB, held until the process ends). n above
 * RT_SORT_NOALLOC_MAX_N is refused with RT_ERR_RANGE. Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (other streams,
 * other contexts, other threads) race; calls queued on one context do not. */
#define RT_SORT_NOALLOC_MAX_N (1 << 23)
int rt_sort_i32_noalloc(rt_ctx* ctx, rt_buf* data, int n);

/* --------------------------------------------------------- block sources */
typedef struct rt_tree_info {
  uint32_t nodes_in_block;
  uint32_t blocks_count;
  uint32_t root_block_index;
  uint32_t channels_count;
  uint64_t bytes_in_block_content;
  uint64_t source_nodes; /* nodes of the generated tree (0 for files) */
} rt_tree_info;

typedef struct rt_terrain_params {
  int depth;
  uint64_t seed;
  int octaves;
  float height_lo, height_hi;
  int spheres;
  float sphere_rmin, sphere_rmax;
  int boxes;
  float box_emin, box_emax;
  int threads;
} rt_terrain_params;

/* StoredOctreeLoader(file_name): CPU data/stored-octree-loader.cpp:9-19. */
int rt_tree_open(const char* path, rt_tree** out);
/* WriteTestOctreeSphere + BuildOctree, in memory: scene sources/stored-octree-test.cpp:317-323. */
int rt_tree_build_sphere(float cx, float cy, float cz, float radius, int level, int nodes_in_block, rt_tree** out);
/* Seeded fBm terrain + spheres + boxes, blocked as StoredOctreeBuilder blocks a tree
 * (CPU data/stored-octree-builder-impl.h:446-459, 348-411), into a tree held in host RAM. The subtrees
 * below the root block are generated and blocked by params->threads workers in parallel and numbered
 * as the sequential builder numbers them, so the node tree is never held whole (depth 15: ~5 G nodes). */
int rt_tree_build_terrain(const rt_terrain_params* params, int nodes_in_block, rt_tree** out);
/* The same blocks, written to a .tree file as they are numbered (StoredOctreeRawWriter,
 * CPU data/stored-octree-raw-writer.cpp:10-99) instead of held in RAM; *out reads that file back
 * (rt_tree_open). */
int rt_tree_build_terrain_file(const rt_terrain_params* params, int nodes_in_block, const char* path,
                               rt_tree** out);
/* The same blocks from the whole node tree generated level by level in host RAM first (depth <= 14);
 * no reference counterpart: the byte-for-byte check of the parallel builder. */
int rt_tree_build_terrain_levelwise(const rt_terrain_params* params, int nodes_in_block, rt_tree** out);
/* ImportObjScene (scene sources/import-obj-scene.cpp:640-652) in memory: reads a Wavefront OBJ with
 * its MTL libraries and PNM textures, voxelises it to depth max_level with channels "normals" (char4)
 * and "colors" (uchar4) (BuilderDataSource, :529-638) and blocks it like BuildOctree; rt_tree_write
 * then stores the file ImportObjScene writes. threads 0 = all cores. Fails with RT_ERR_IO where the
 * reference crashes (bad numbers or faces, unknown material names, unreadable textures) and for
 * texture formats other than PNM (.ppm/.pgm/.pnm). */
int rt_tree_import_obj(const char* obj_path, int max_level, int nodes_in_block, int threads, rt_tree** out);
/* StoredOctreeRawWriter output: CPU data/stored-octree-raw-writer.cpp:35-99. */
int rt_tree_write(const rt_tree* tree, const char* path);
int rt_tree_info_get(const rt_tree* tree, rt_tree_info* out);
/* Channel i: bytes per node and name (name_cap bytes, NUL-terminated). */
int rt_tree_channel(const rt_tree* tree, int index, int* bytes_in_node, char* name, int name_cap);
/* StoredOctreeLoader::LoadBlock: stored-octree-loader.cpp:31-58. header76 receives the 76-byte
 * block header record, payload the channel data (bytes_in_block_content). */
int rt_tree_load_block(rt_tree* tree, uint32_t index, uint8_t* header76, uint8_t* payload, uint64_t payload_cap);
int rt_tree_destroy(rt_tree* tree);

/* ------------------------------------------------------------ block cache */
/* GPUOctreeCacheManager(max_blocks_count, loader, ctx): GPU data/gpu-octree-cache-manager.cpp:11-31. */
int rt_cache_create(rt_ctx* ctx, rt_tree* tree, int max_blocks_count, rt_cache** out);
int rt_cache_destroy(rt_cache* cache);
/* InitialFillCache: gpu-octree-cache-manager.cpp:60-108. */
int rt_cache_initial_fill(rt_cache* cache);
/* root_node_index(): gpu-octree-cache-manager.cpp:41-43. */
int rt_cache_root_node_index(const rt_cache* cache);
int rt_cache_nodes_in_block(const rt_cache* cache);
int rt_cache_max_blocks(const rt_cache* cache);
int rt_cache_loaded_blocks(const rt_cache* cache);
/* GPUOctreeData::ChannelByIndex / ChannelByName / far_pointers_buffer: gpu-octree-data.h:24-27.
 * Returns NULL when absent. Borrowed: owned by the cache. */
rt_buf* rt_cache_channel(rt_cache* cache, int index);
rt_buf* rt_cache_channel_by_name(rt_cache* cache, const char* name);
rt_buf* rt_cache_far_pointers(rt_cache* cache);
/* GPUOctreeCache::BlockIndexInCache: gpu-octree-cache.cpp:51-56 (-1 if absent). */
int rt_cache_block_slot(const rt_cache* cache, int tree_block_index);
/* Tree block held by a slot (-1 if free); device page-table mirror. */
int rt_cache_slot_block(const rt_cache* cache, int slot);
/* Session / transaction protocol: gpu-octree-cache-manager.cpp:110-212. */
int rt_cache_start_session(rt_cache* cache);
int rt_cache_start_transaction(rt_cache* cache);
int rt_cache_mark_block_used(rt_cache* cache, int slot);
int rt_cache_mark_parent_used(rt_cache* cache, int slot);
int rt_cache_request_block(rt_cache* cache, int tree_block_index);
int rt_cache_transaction_filled(const rt_cache* cache); /* 1/0, <0 on error */
int rt_cache_session_filled(const rt_cache* cache);
int rt_cache_end_transaction(rt_cache* cache);
int rt_cache_end_session(rt_cache* cache);
/* LRU order, least recently used first (blocks as slots, markers as -1). Writes up to cap entries;
 * returns the count. */
int rt_cache_lru_order(const rt_cache* cache, int* out, int cap);
/* Counters since creation: {uploaded blocks, evicted blocks, pointer patches, far pointer writes}. */
int rt_cache_counters(const rt_cache* cache, uint64_t out[4]);

/* ---------------------------------------------------- feedback extraction */
/* RayTracingFeedbackExtractor(w, h, ctx): GPU_renderer/ray-tracing-feedback-extractor.cpp:11-34. */
int rt_feedback_create(rt_ctx* ctx, int count, rt_feedback** out);
int rt_feedback_destroy(rt_feedback* fb);
/* set_hits_to_extract: ray-tracing-feedback-extractor.h:20-21 (default 4000). */
int rt_feedback_set_hits_to_extract(rt_feedback* fb, int value);
/* Seeds the hit-window generator that replaces rand() (ray-tracing-feedback-extractor.cpp:147). */
int rt_feedback_set_seed(rt_feedback* fb, uint64_t seed);
/* Process(faults_and_hits, nodes_in_block): ray-tracing-feedback-extractor.cpp:38-42.
 * codes_are_block_level != 0 skips NodeToBlock (codes written block-level by the renderer). */
int rt_feedback_process(rt_feedback* fb, rt_buf* faults_and_hits, int nodes_in_block, int codes_are_block_level);
/* fault_blocks / hit_blocks / duplicate_hit_blocks: ray-tracing-feedback-extractor.h:30-32.
 * which: 0 faults, 1 hits, 2 duplicate hits. Copies up to cap entries, returns the list length. */
int rt_feedback_list(const rt_feedback* fb, int which, int* out, int cap);
/* Unique count and expected_faults after the last Process (ray-tracing-feedback-extractor.cpp:97,180). */
int rt_feedback_stats(const rt_feedback* fb, int* unique_count, int* expected_faults);

/* ----------------------------------------------------------------- tracer */
/* GPURayTracer(cache_manager, w, h, ctx): GPU_renderer/gpu-ray-tracer.cpp:20-44. */
int rt_tracer_create(rt_cache* cache, int width, int height, rt_tracer** out);
int rt_tracer_destroy(rt_tracer* tr);
/* set_frame_time_limit (seconds; gpu-ray-tracer.h:17-18) and set_lod_voxel_size (gpu-ray-tracer.h:20-21). */
int rt_tracer_set_frame_time_limit(rt_tracer* tr, double seconds);
int rt_tracer_set_lod_voxel_size(rt_tracer* tr, float size);
/* Deterministic pass cap in place of the wall-clock cut (0 = no cap). */
int rt_tracer_set_max_passes(rt_tracer* tr, int passes);
int rt_tracer_set_flags(rt_tracer* tr, int flags);
int rt_tracer_set_seed(rt_tracer* tr, uint64_t seed);
/* RenderFrame(camera): gpu-ray-tracer.cpp:66-227, camera given as its inverse view-projection. */
int rt_tracer_render_frame(rt_tracer* tr, const float view_proj_inv[16]);
/* out_image(): gpu-ray-tracer.cpp:229-231 (uchar4 RGBA, borrowed). */
rt_buf* rt_tracer_out_image(rt_tracer* tr);
/* faults_and_hits buffer of the last pass (int32 per pixel, borrowed). */
rt_buf* rt_tracer_codes(rt_tracer* tr);
rt_feedback* rt_tracer_feedback(rt_tracer* tr);
/* Per-frame statistics: {passes, total faults, uploaded blocks, session filled, rays traced}. */
int rt_tracer_frame_stats(const rt_tracer* tr, uint64_t out[5]);
/* Per-pass fault lists of the last frame, concatenated: counts[pass] entries each. */
int rt_tracer_pass_faults(const rt_tracer* tr, int* counts, int counts_cap, int* faults, int faults_cap);
/* Last frame: {device ms inside tracing-pass launches (HIP events), host wall ms of RenderFrame}
 * (the PROFILE_TIMER "GPU tracing passes" / "frame" of gpu-ray-tracer.cpp:124-263). */
int rt_tracer_timings(const rt_tracer* tr, double out_ms[2]);
/* No reference counterpart. Pipelining of frames over a fully resident pool (default on): the call
 * queues the frame's trace and feedback walk and returns; the frame's host half (LRU marks of its hit
 * lists, statistics, timings) runs when the next frame's trace has been queued behind it, or when any
 * entry point reads or changes the tracer's, its cache's or its feedback extractor's state. Results,
 * lists, statistics and LRU order are those of the unpipelined frame; the device idles less between
 * frames. enabled = 0 finishes a pending frame and renders every later frame to the end in the call. */
int rt_tracer_set_pipelining(rt_tracer* tr, int enabled);
/* No reference counterpart: device ms inside tracing-pass launches summed over every finished frame
 * since creation, and the number of those frames (a pending frame is finished first). */
int rt_tracer_trace_totals(rt_tracer* tr, double* trace_ms, uint64_t* frames);

/* ---------------------------------------------------------------- batched
 * Many views traced by one kernel launch (no reference counterpart: the
 * reference renders one view per RenderFrame). Each view's pixels and
 * feedback codes are those of rt_render_plain for that view, written at
 * pixel_offset into shared output buffers. */
typedef struct rt_view {
  float view_proj_inv[16];
  int width, height;
  float lod_voxel_size;
  int flags;             /* RT_FLAG_SHADOWS */
  int scene;             /* 0: the cache given to rt_batch_create; else an index from rt_batch_add_scene */
  uint64_t pixel_offset; /* first pixel of this view in the output buffers */
} rt_view;

/* The batch renders scene 0 = `cache`; every scene must stay resident (no faults) while batches run. */
int rt_batch_create(rt_cache* cache, int max_views, rt_batch** out);
int rt_batch_destroy(rt_batch* batch);
/* Registers another resident scene of the same context; returns its scene index (>= 1) or a negative error. */
int rt_batch_add_scene(rt_batch* batch, rt_cache* cache);
/* Launch shape: tile_shape 0 = 8x8, 1 = 16x4, 2 = 32x2 pixels per wave; lds_levels (8, 10, 12, 16) deepest
 * traversal-stack levels kept in LDS; wgs_per_cu 0 = one wave per tile, dispatched by the hardware, or 1..32
 * persistent one-wave workgroups per compute unit that take tiles from a shared queue (the LDS window caps how
 * many are resident: 32 at 10 levels);
 * tree_depth of the batch's scenes (0..22; 0 = unknown: the
 * stack spills to a private array beyond the LDS window). With tree_depth - 1 <= lds_levels the whole stack
 * lives in LDS, under the same rule on the declared depth as rt_plain_config. Default: 8x8, 12 levels, one wave
 * per tile, depth unknown. */
int rt_batch_config(rt_batch* batch, int tile_shape, int lds_levels, int wgs_per_cu, int tree_depth);
/* Enqueues one batched launch on the cache's context stream. stats (may be NULL) makes it the counting variant
 * and receives {primary rays, shadow rays, node fetches, hits}. */
int rt_batch_render(rt_batch* batch, const rt_view* views, int nviews, rt_buf* out_image, rt_buf* codes,
                    uint64_t* stats);
/* Device time of the last rt_batch_render launch (HIP events; waits for it). */
int rt_batch_last_ms(rt_batch* batch, double* ms);
/* Autotuner: times every feasible (tile shape, LDS stack levels, workgroups per CU) on this batch (median of
 * `repeats` launches after one warm-up), keeps the fastest as the batch configuration and returns it in
 * best = {tile_shape, lds_levels, wgs_per_cu} with its time. */
int rt_batch_autotune(rt_batch* batch, const rt_view* views, int nviews, rt_buf* out_image, rt_buf* codes,
                      int repeats, int best[3], double* best_ms);

/* ---------------------------------------------------------------- serving
 * Continuous-batching request loop (no reference counterpart: the reference
 * renders one view per RenderFrame on the calling thread). A server thread
 * takes submitted requests in FIFO order and admits each when
 *  - its output fits in the device framebuffer arena (8 bytes per pixel:
 *    RGBA8 + feedback code), and
 *  - its scene is resident: scenes are loaded whole into their own block
 *    pool on first use, within a device pool budget, evicting idle scenes
 *    (no request in flight) least recently used first.
 * Admitted requests form a round: one batched launch
 * (rt_batch_render semantics; each result equals rt_render_plain of that
 * view) on one of `rounds_in_flight` streams, followed by the copy of each
 * result to the host. Requests that arrive while rounds run are admitted into
 * the next round as soon as a round slot, arena space and pool space free up;
 * rounds on different streams overlap on the device. */
typedef struct rt_server_config {
  int max_views_per_round;    /* 1..65536 */
  int rounds_in_flight;       /* 1..4 */
  uint64_t framebuffer_bytes; /* device arena for request outputs */
  uint64_t pool_bytes;        /* device budget for resident scenes' block pools */
  int tile_shape, lds_levels, wgs_per_cu; /* launch shape of every round (rt_batch_config) */
  int max_block_loads; /* blocks loaded per scene between rounds, out of core (0: until the pool is full); a
                        * bound on the time a load holds the queue, as the reference's frame time limit */
} rt_server_config;

#define RT_OUT_IMAGE 1      /* bring the RGBA8 image to host memory */
#define RT_OUT_CODES 2      /* bring the feedback codes to host memory */
#define RT_DELIVER_COPY 0   /* results are copied into out_image / out_codes before completion */
#define RT_DELIVER_PINNED 1 /* the completion points into the server's pinned memory until rt_server_release */

typedef struct rt_request {
  float view_proj_inv[16];
  int width, height;
  float lod_voxel_size;
  int flags;           /* RT_FLAG_SHADOWS */
  int scene;           /* index returned by rt_server_add_scene */
  int outputs;         /* RT_OUT_IMAGE | RT_OUT_CODES (0: render only) */
  int delivery;        /* RT_DELIVER_COPY or RT_DELIVER_PINNED */
  int pad;
  uint32_t* out_image; /* RT_DELIVER_COPY + RT_OUT_IMAGE: width*height words, valid until completion */
  int32_t* out_codes;  /* RT_DELIVER_COPY + RT_OUT_CODES: width*height codes, valid until completion */
} rt_request;

typedef struct rt_completion {
  uint64_t id;
  int status;              /* RT_OK, or the error that failed the request's round */
  int round_views;         /* views in the round that rendered it */
  double submit_ms;        /* times on the server's clock (ms since rt_server_create) */
  double admit_ms;
  double done_ms;          /* result in host memory */
  const uint32_t* image;   /* RT_DELIVER_PINNED: the image in pinned memory (NULL otherwise) */
  const int32_t* codes;    /* RT_DELIVER_PINNED: the codes in pinned memory (NULL otherwise) */
  int fault_blocks;        /* distinct tree blocks the rays found missing from the scene's pool (their pixels
                            * show the coarser parent, as FinishTracingFrame does); 0 = full detail. Always 0
                            * on a scene whose whole tree is resident */
  int pad2;
} rt_completion;

typedef struct rt_server_stats {
  uint64_t submitted, completed, failed, rounds;
  uint64_t pixels, primary_rays, shadow_rays;
  uint64_t scene_loads, scene_evictions;
  int max_rounds_in_flight; /* largest number of rounds seen on the device together */
  int pad;
  double p50_ms, p99_ms, mean_ms, max_ms; /* request latency, submit -> result in host memory */
  double span_ms;       /* first admission -> last completion since the last reset */
  double mrays_per_s;   /* (primary + shadow rays) / span */
  double mean_views_per_round;
  uint64_t block_uploads;     /* blocks loaded into partial scene pools from feedback */
  uint64_t load_transactions; /* cache transactions run between rounds for partial scenes */
} rt_server_stats;

/* Starts the server thread. */
int rt_server_create(rt_ctx* ctx, const rt_server_config* cfg, rt_server** out);
/* Stops the thread after the requests in flight are done; pending requests complete as failed. */
int rt_server_destroy(rt_server* server);
/* Registers a scene (the tree must outlive the server); tree_depth places the LDS stack window (0 = 11).
 * Returns the scene index or a negative error (RT_ERR_RANGE when its block pool exceeds pool_bytes). */
int rt_server_add_scene(rt_server* server, rt_tree* tree, int tree_depth);
/* The same with a block pool of max_blocks slots (<= 0: the whole tree). A pool smaller than the tree is
 * rendered out of core: it starts with the breadth-first fill of InitialFillCache
 * (gpu-octree-cache-manager.cpp:60-108); each round's feedback codes of the scene go through the feedback
 * extractor (ray-tracing-feedback-extractor.cpp:38-185) and, once no round of the scene is in flight, one
 * session / transaction (gpu-octree-cache-manager.cpp:110-212) marks the hit blocks used and loads the
 * faulted ones, so later requests see more detail (one RenderFrame pass per request, the reference's frame
 * after its time limit). Requests of a scene with loads due wait for its rounds in flight to finish. */
int rt_server_add_scene_pool(rt_server* server, rt_tree* tree, int tree_depth, int max_blocks);
/* Queues a request (thread-safe, does not wait); rejects requests that can never be admitted. */
int rt_server_submit(rt_server* server, const rt_request* req, uint64_t* id);
/* Queues n requests at once (all or none are accepted); ids receives their ids in order. */
int rt_server_submit_many(rt_server* server, const rt_request* reqs, int n, uint64_t* ids);
/* Returns up to cap completions, waiting up to timeout_ms for the first one; the count or a negative error. */
int rt_server_poll(rt_server* server, rt_completion* out, int cap, int timeout_ms);
/* Returns the pinned result of a RT_DELIVER_PINNED request to the server: its framebuffer range is free
 * for later requests, and the completion's pointers are no longer valid. */
int rt_server_release(rt_server* server, uint64_t id);
/* Waits until every submitted request has completed (or timeout_ms passes: RT_ERR_STATE). */
int rt_server_drain(rt_server* server, int timeout_ms);
int rt_server_stats_get(rt_server* server, rt_server_stats* out);
/* Clears latencies and counters (for measuring after a warm-up). */
int rt_server_reset_stats(rt_server* server);

/* ----------------------------------------------------------------- camera */
/* Camera::ViewProjectionMatrix().Inverse(): scene/camera.cpp:23-28, util/mat.cpp:6-194. */
int rt_camera_view_proj_inv(float yaw, float pitch, float roll, const float position[3], float fov_deg,
                            float near_clip, float aspect, float out16[16]);

#ifdef __cplusplus
}
#endif

#endif /* RAYT_HIP_H_ */
