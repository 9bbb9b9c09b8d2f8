// Device, buffer and event layer of the C ABI: the replacement of the
// reference's CL util (CLContext / CLBuffer / CLEvent). One HIP stream per
// context executes everything in order; events mark points on it.
#include "rt_internal.h"

#include <cstring>
#include <vector>

namespace rayt {

static thread_local std::string g_last_error;

void SetError(const std::string& msg) { g_last_error = msg; }
int Fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
const char* LastError() { return g_last_error.c_str(); }

rt_ctx* CtxRetain(rt_ctx* ctx) {
  if (ctx) ctx->refs.fetch_add(1, std::memory_order_relaxed);
  return ctx;
}

void CtxRelease(rt_ctx* ctx) {
  if (!ctx || ctx->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

}  // namespace rayt

namespace {

__global__ void FillU32Kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

int WaitList(rt_ctx* ctx, rt_evt* const* wait, int nwait) {
  for (int i = 0; i < nwait; ++i) {
    RT_REQUIRE(wait && wait[i], RT_ERR_INVALID, "null event in wait list");
    RT_HIP_CHECK(hipStreamWaitEvent(ctx->stream, wait[i]->ev, 0));
  }
  return RT_OK;
}

int RecordOut(rt_ctx* ctx, rt_evt** out) {
  if (!out) return RT_OK;
  rt_evt* e = new rt_evt();
  hipError_t err = hipEventCreate(&e->ev);
  if (err != hipSuccess) {
    delete e;
    return rayt::Fail(RT_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(err));
  }
  err = hipEventRecord(e->ev, ctx->stream);
  if (err != hipSuccess) {
    (void)hipEventDestroy(e->ev);
    delete e;
    return rayt::Fail(RT_ERR_HIP, std::string("hipEventRecord: ") + hipGetErrorString(err));
  }
  *out = e;
  return RT_OK;
}

}  // namespace

extern "C" {

const char* rt_last_error(void) { return rayt::LastError(); }
const char* rt_version(void) { return "rayt_hip gfx950"; }

int rt_ctx_create(int device, rt_ctx** out) {
  RT_REQUIRE(out, RT_ERR_INVALID, "null out");
  int n = 0;
  RT_HIP_CHECK(hipGetDeviceCount(&n));
  RT_REQUIRE(device >= 0 && device < n, RT_ERR_RANGE, "device index out of range");
  RT_HIP_CHECK(hipSetDevice(device));
  rt_ctx* c = new rt_ctx();
  c->device = device;
  hipDeviceProp_t props;
  hipError_t e = hipGetDeviceProperties(&props, device);
  if (e != hipSuccess) {
    delete c;
    return rayt::Fail(RT_ERR_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
  }
  c->compute_units = props.multiProcessorCount;
  const int scratch_rc = rayt::NoAllocScratchInit(device);
  if (scratch_rc != RT_OK) {
    delete c;
    return scratch_rc;
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return rayt::Fail(RT_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = c;
  return RT_OK;
}

int rt_ctx_destroy(rt_ctx* ctx) {
  rayt::CtxRelease(ctx);
  return RT_OK;
}

int rt_ctx_finish(rt_ctx* ctx) {
  RT_REQUIRE(ctx, RT_ERR_INVALID, "null context");
  RT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return RT_OK;
}

int rt_ctx_compute_units(rt_ctx* ctx, int* out) {
  RT_REQUIRE(ctx && out, RT_ERR_INVALID, "null argument");
  *out = ctx->compute_units;
  return RT_OK;
}

int rt_buf_create(rt_ctx* ctx, uint64_t bytes, rt_buf** out) {
  RT_REQUIRE(ctx && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(bytes > 0, RT_ERR_INVALID, "buffer size must be positive");
  RT_HIP_CHECK(hipSetDevice(ctx->device));
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return rayt::Fail(RT_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  rt_buf* b = new rt_buf();
  b->ctx = rayt::CtxRetain(ctx);
  b->ptr = p;
  b->size = bytes;
  *out = b;
  return RT_OK;
}

int rt_buf_destroy(rt_buf* buf) {
  if (!buf) return RT_OK;
  if (buf->owned && buf->ptr) {
    // best effort, as a destructor: the buffer is gone whatever they report
    (void)hipStreamSynchronize(buf->ctx->stream);
    (void)hipFree(buf->ptr);
  }
  rayt::CtxRelease(buf->ctx);
  delete buf;
  return RT_OK;
}

uint64_t rt_buf_size(const rt_buf* buf) { return buf ? buf->size : 0; }
void* rt_buf_device_ptr(const rt_buf* buf) { return buf ? buf->ptr : nullptr; }

int rt_buf_write(rt_buf* buf, uint64_t offset, uint64_t length, const void* src, int blocking,
                 rt_evt* const* wait, int nwait, rt_evt** out_event) {
  RT_REQUIRE(buf && src, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(length > 0 && offset + length <= buf->size && offset + length >= offset, RT_ERR_RANGE,
             "write outside buffer");
  int rc = WaitList(buf->ctx, wait, nwait);
  if (rc) return rc;
  RT_HIP_CHECK(hipMemcpyAsync(static_cast<char*>(buf->ptr) + offset, src, length, hipMemcpyHostToDevice,
                              buf->ctx->stream));
  rc = RecordOut(buf->ctx, out_event);
  if (rc) return rc;
  if (blocking) RT_HIP_CHECK(hipStreamSynchronize(buf->ctx->stream));
  return RT_OK;
}

int rt_buf_read(const rt_buf* buf, uint64_t offset, uint64_t length, void* dst, int blocking,
                rt_evt* const* wait, int nwait, rt_evt** out_event) {
  RT_REQUIRE(buf && dst, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(length > 0 && offset + length <= buf->size && offset + length >= offset, RT_ERR_RANGE,
             "read outside buffer");
  int rc = WaitList(buf->ctx, wait, nwait);
  if (rc) return rc;
  RT_HIP_CHECK(hipMemcpyAsync(dst, static_cast<const char*>(buf->ptr) + offset, length, hipMemcpyDeviceToHost,
                              buf->ctx->stream));
  rc = RecordOut(buf->ctx, out_event);
  if (rc) return rc;
  if (blocking) RT_HIP_CHECK(hipStreamSynchronize(buf->ctx->stream));
  return RT_OK;
}

int rt_buf_copy(rt_buf* dst, uint64_t dst_offset, const rt_buf* src, uint64_t src_offset, uint64_t length,
                rt_evt* const* wait, int nwait, rt_evt** out_event) {
  RT_REQUIRE(dst && src, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(length > 0 && dst_offset <= dst->size && length <= dst->size - dst_offset && src_offset <= src->size &&
                 length <= src->size - src_offset,
             RT_ERR_RANGE, "copy outside buffer");
  int rc = WaitList(dst->ctx, wait, nwait);
  if (rc) return rc;
  RT_HIP_CHECK(hipMemcpyAsync(static_cast<char*>(dst->ptr) + dst_offset,
                              static_cast<const char*>(src->ptr) + src_offset, length, hipMemcpyDeviceToDevice,
                              dst->ctx->stream));
  return RecordOut(dst->ctx, out_event);
}

int rt_buf_fill_u32(rt_buf* buf, uint64_t offset_words, uint64_t count, uint32_t value) {
  RT_REQUIRE(buf, RT_ERR_INVALID, "null buffer");
  RT_REQUIRE(offset_words <= buf->size / 4 && count <= buf->size / 4 - offset_words, RT_ERR_RANGE,
             "fill outside buffer");
  if (!count) return RT_OK;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  FillU32Kernel<<<dim3(unsigned(blocks)), dim3(256), 0, buf->ctx->stream>>>(
      static_cast<uint32_t*>(buf->ptr) + offset_words, count, value);
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_evt_wait(int n, rt_evt* const* evts) {
  for (int i = 0; i < n; ++i) {
    RT_REQUIRE(evts && evts[i], RT_ERR_INVALID, "null event");
    RT_HIP_CHECK(hipEventSynchronize(evts[i]->ev));
  }
  return RT_OK;
}

int rt_evt_retain(rt_evt* e) {
  RT_REQUIRE(e, RT_ERR_INVALID, "null event");
  e->refs.fetch_add(1);
  return RT_OK;
}

int rt_evt_release(rt_evt* e) {
  if (!e) return RT_OK;
  if (e->refs.fetch_sub(1) == 1) {
    (void)hipEventDestroy(e->ev);
    delete e;
  }
  return RT_OK;
}

int rt_evt_elapsed_ns(rt_evt* start, rt_evt* end, double* out_ns) {
  RT_REQUIRE(start && end && out_ns, RT_ERR_INVALID, "null argument");
  float ms = 0;
  RT_HIP_CHECK(hipEventElapsedTime(&ms, start->ev, end->ev));
  *out_ns = double(ms) * 1e6;
  return RT_OK;
}

int rt_evt_record(rt_ctx* ctx, rt_evt** out) {
  RT_REQUIRE(ctx && out, RT_ERR_INVALID, "null argument");
  return RecordOut(ctx, out);
}

}  // extern "C"
