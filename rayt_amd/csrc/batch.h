// Internal interface of the batched renderer, used by the request server.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "rt_internal.h"

namespace rayt {

// Scenes one batch can address (views pick one by index).
constexpr int kMaxBatchScenes = 64;

// A batch with no scene yet (BatchSetScenes gives it its scenes).
int BatchCreate(rt_ctx* ctx, int max_views, rt_batch** out);
// Scene list of the next launches: view.scene indexes `caches`; depths[i] is
// the tree depth of scene i (places the LDS stack window; 0 = the batch's
// configured depth).
int BatchSetScenes(rt_batch* b, const std::vector<rt_cache*>& caches, const std::vector<int>& depths);
// Launches of `b` go to stream `s` (nullptr: back to the context stream).
void BatchSetStream(rt_batch* b, hipStream_t s);
// Whether launches record the start event rt_batch_last_ms reads (default on).
void BatchSetTiming(rt_batch* b, bool timing);
// Counting launches count the rays and hits but not the node fetches (which
// then read 0), so no traversal step does counting work (default off).
void BatchSetRayCounts(rt_batch* b, bool rays_only);
// Enqueues one launch without waiting; a launch with `count` leaves
// {primary rays, shadow rays, node fetches, hits} for BatchStats. Waits for
// the previous launch of `b` first (its descriptors are staged in place).
int BatchLaunch(rt_batch* b, const rt_view* views, int nviews, rt_buf* out_image, rt_buf* codes, bool count);
// Counters of the last (counting) launch; waits for it.
int BatchStats(rt_batch* b, uint64_t out[4]);
// Event recorded after the last launch of `b` (and its counter read-back).
hipEvent_t BatchDoneEvent(rt_batch* b);

}  // namespace rayt
