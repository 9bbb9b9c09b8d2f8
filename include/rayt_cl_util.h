/* C++ replacements of the reference's OpenCL wrappers (rayt/CL util): the
 * classes CLContext, CLEvent, CLEventList, CLBuffer and CLKernel, with the
 * names and member signatures their callers use (GPURayTracer,
 * RayTracingFeedbackExtractor, GPUSort, GPUScan, GPUOctreeData,
 * GPUOctreeChannel, GPUOctreeCache, GPUOctreeCacheManager), implemented over
 * the C ABI of rayt_hip.h. Header-only; link with -lrayt_hip.
 *
 *   CLContext    CL util/cl-context.h:7-23
 *   CLEvent      CL util/cl-event.h:21-40
 *   CLEventList  CL util/cl-event.h:41-66
 *   CLBuffer     CL util/cl-buffer.h:9-35
 *   CLKernel     CL util/cl-kernel.h:9-39
 *
 * Differences a caller can see:
 *  - Every call that fails calls rayt_cl_crash() with the library's message;
 *    by default it prints the message and aborts, as the reference's crash()
 *    exits (util/common.h:43-53). Define RAYT_CL_CRASH before including this
 *    header to route it elsewhere.
 *  - The context pointer taken by CLBuffer and CLKernel may be any shared
 *    pointer type with get() (boost::shared_ptr in the reference).
 *  - CLKernel compiles nothing: the kernel is looked up by name among the
 *    library's gfx950 kernels when it runs, with the arguments set by index
 *    as the reference kernels declare them; the -D WIDTH / HEIGHT
 *    parameters of the frame kernels come from Run2D's sizes. A name the
 *    library does not have is refused at construction.
 *  - All work of one context runs on one in-order stream, so a wait list of
 *    events of the same context is always satisfied by the time a later call
 *    runs (the reference's out-of-order queue needed the lists).
 */
#ifndef RAYT_CL_UTIL_H_
#define RAYT_CL_UTIL_H_

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "rayt_hip.h"

#ifndef RAYT_CL_CRASH
#define RAYT_CL_CRASH(msg)                                   \
  do {                                                       \
    std::fprintf(stderr, "rayt CL util: %s\n", (msg));       \
    std::fflush(stderr);                                     \
    std::abort();                                            \
  } while (0)
#endif

/* OpenCL's buffer flags (cl.h), accepted and ignored: device memory is one kind here. */
#ifndef CL_MEM_READ_WRITE
typedef uint64_t cl_mem_flags;
#define CL_MEM_READ_WRITE (1 << 0)
#define CL_MEM_WRITE_ONLY (1 << 1)
#define CL_MEM_READ_ONLY (1 << 2)
#endif

namespace rayt {

inline void rayt_cl_check(int rc) {
  if (rc < 0) RAYT_CL_CRASH(rt_last_error());
}

class CLContext {
 public:
  CLContext() { rayt_cl_check(rt_ctx_create(0, &ctx_)); }
  ~CLContext() { rt_ctx_destroy(ctx_); }
  CLContext(const CLContext&) = delete;
  CLContext& operator=(const CLContext&) = delete;
  void WaitForAll() { rayt_cl_check(rt_ctx_finish(ctx_)); }  // clFinish
  rt_ctx* handle() const { return ctx_; }

 private:
  rt_ctx* ctx_ = nullptr;
};

class CLEvent {
 public:
  CLEvent() = default;
  CLEvent(const CLEvent& e) : e_(e.e_) {
    if (e_) rt_evt_retain(e_);
  }
  ~CLEvent() { reset(nullptr); }
  CLEvent& operator=(const CLEvent& e) {
    if (e.e_) rt_evt_retain(e.e_);
    reset(e.e_);
    return *this;
  }
  // Takes over one reference of e.
  void reset(rt_evt* e) {
    if (e_) rt_evt_release(e_);
    e_ = e;
  }
  rt_evt* event() const { return e_; }  // may be null
  void WaitFor() const {
    if (e_) rayt_cl_check(rt_evt_wait(1, &e_));
  }

 private:
  rt_evt* e_ = nullptr;
};

class CLEventList {
 public:
  CLEventList() = default;
  CLEventList(const CLEventList& l) { *this = l; }
  ~CLEventList() { Clear(); }
  CLEventList& operator=(const CLEventList& l) {
    if (this == &l) return *this;
    Clear();
    for (rt_evt* e : l.v_) {
      rt_evt_retain(e);
      v_.push_back(e);
    }
    return *this;
  }
  void Clear() {
    for (rt_evt* e : v_) rt_evt_release(e);
    v_.clear();
  }
  void Add(const CLEvent& e) {
    if (!e.event()) return;
    rt_evt_retain(e.event());
    v_.push_back(e.event());
  }
  int size() const { return int(v_.size()); }
  void WaitFor() const {
    if (!v_.empty()) rayt_cl_check(rt_evt_wait(int(v_.size()), v_.data()));
  }
  rt_evt* const* events() const { return v_.data(); }

 private:
  std::vector<rt_evt*> v_;
};

namespace cl_util_detail {
// Keeps a caller's shared pointer (of any type) alive.
template <class Ptr>
std::shared_ptr<const void> Keep(Ptr p) {
  auto h = std::make_shared<Ptr>(std::move(p));
  return std::shared_ptr<const void>(h, h.get());
}
inline rt_evt* const* Events(const CLEventList* l) { return l && l->size() ? l->events() : nullptr; }
inline int Count(const CLEventList* l) { return l ? l->size() : 0; }
}  // namespace cl_util_detail

class CLBuffer {
 public:
  template <class ContextPtr>
  CLBuffer(cl_mem_flags /*flags*/, int size, ContextPtr context)
      : keep_(cl_util_detail::Keep(context)), ctx_(context.get()), size_(size) {
    rayt_cl_check(rt_buf_create(ctx_->handle(), uint64_t(size), &buf_));
  }
  // A view of a buffer owned elsewhere (a channel of a block pool, a tracer's output), not freed here.
  explicit CLBuffer(rt_buf* borrowed)
      : buf_(borrowed), size_(borrowed ? int(rt_buf_size(borrowed)) : 0), owned_(false) {}
  ~CLBuffer() {
    if (owned_) rt_buf_destroy(buf_);
  }
  CLBuffer(const CLBuffer&) = delete;
  CLBuffer& operator=(const CLBuffer&) = delete;

  rt_buf* buffer() const { return buf_; }
  int size() const { return size_; }

  void CopyFrom(CLBuffer& buf, int src_start, int dest_start, int length, const CLEventList* wait_list,
                CLEvent* out_event) {
    rt_evt* e = nullptr;
    rayt_cl_check(rt_buf_copy(buf_, uint64_t(dest_start), buf.buf_, uint64_t(src_start), uint64_t(length),
                              cl_util_detail::Events(wait_list), cl_util_detail::Count(wait_list),
                              out_event ? &e : nullptr));
    if (out_event) out_event->reset(e);
  }
  void CopyFrom(CLBuffer& buf, const CLEventList* wait_list, CLEvent* out_event) {
    CopyFrom(buf, 0, 0, size_, wait_list, out_event);
  }
  void Write(int start, int length, const void* data, bool blocking, const CLEventList* wait_list,
             CLEvent* out_event) {
    rt_evt* e = nullptr;
    rayt_cl_check(rt_buf_write(buf_, uint64_t(start), uint64_t(length), data, blocking ? 1 : 0,
                               cl_util_detail::Events(wait_list), cl_util_detail::Count(wait_list),
                               out_event ? &e : nullptr));
    if (out_event) out_event->reset(e);
  }
  // Non-blocking, as in the reference (cl-buffer.cpp:112).
  void Read(int start, int length, void* data, const CLEventList* wait_list, CLEvent* out_event) {
    rt_evt* e = nullptr;
    rayt_cl_check(rt_buf_read(buf_, uint64_t(start), uint64_t(length), data, 0, cl_util_detail::Events(wait_list),
                              cl_util_detail::Count(wait_list), out_event ? &e : nullptr));
    if (out_event) out_event->reset(e);
  }
  // Finishes all queued work, then reads (the reference's debug read, cl-buffer.cpp:122-129).
  void Read(int start, int length, void* data) const {
    rayt_cl_check(rt_buf_read(buf_, uint64_t(start), uint64_t(length), data, 1, nullptr, 0, nullptr));
  }

 private:
  std::shared_ptr<const void> keep_;
  CLContext* ctx_ = nullptr;
  rt_buf* buf_ = nullptr;
  int size_ = 0;
  bool owned_ = true;
};

class CLKernel {
 public:
  template <class ContextPtr>
  CLKernel(std::string file_name, std::string params, std::string kernel_name, ContextPtr context)
      : keep_(cl_util_detail::Keep(context)), ctx_(context.get()), file_(std::move(file_name)),
        name_(std::move(kernel_name)) {
    Init(params);
  }
  template <class ContextPtr>
  CLKernel(std::string include_dir, std::string file_name, std::string params, std::string kernel_name,
           ContextPtr context)
      : CLKernel(include_dir + "/" + file_name, std::move(params), std::move(kernel_name), context) {}
  ~CLKernel() {
    if (temp_) rt_buf_destroy(temp_);
  }
  CLKernel(const CLKernel&) = delete;
  CLKernel& operator=(const CLKernel&) = delete;

  // The reference sizes local memory and work ranges by this; the kernels here pick their own.
  int work_group_size() const { return 256; }

  void SetArg(int index, int size, void* data) {
    Arg& a = At(index);
    a.raw.assign(static_cast<const uint8_t*>(data), static_cast<const uint8_t*>(data) + (data ? size : 0));
  }
  void SetLocalBufferArg(int index, int size) { At(index).local = size; }
  void SetIntArg(int index, int val) { At(index).i = val; }
  void SetFloatArg(int index, float val) { At(index).f[0] = val; }
  template <class Vec3>
  void SetFloat4Arg(int index, Vec3 xyz, float w) {
    Arg& a = At(index);
    a.f[0] = xyz.x, a.f[1] = xyz.y, a.f[2] = xyz.z, a.f[3] = w;
  }
  void SetBufferArg(int index, const CLBuffer* buf) { At(index).buf = buf ? buf->buffer() : nullptr; }
  template <class Mat4>
  void SetFloat16Arg(int index, Mat4 mat) {
    memcpy(At(index).f, mat.m, sizeof(float) * 16);
  }

  void Run1D(int size, const CLEventList* wait_list, CLEvent* out_event) { Run(size, 1, wait_list, out_event); }
  void Run2D(int size0, int size1, const CLEventList* wait_list, CLEvent* out_event) {
    Run(size0, size1, wait_list, out_event);
  }

 private:
  struct Arg {
    int i = 0;
    float f[16] = {};
    rt_buf* buf = nullptr;
    int local = 0;
    std::vector<uint8_t> raw;
  };
  enum Kind {
    kInitTracingFrame,
    kRaytracingPass,
    kFinishTracingFrame,
    kRaytrace,
    kRaytraceVolatile,
    kNodeToBlock,
    kDifferenceFlag,
    kCompress,
    kLocalMergeSort,
    kGlobalMergeSortPass,
    kScanPass,
  };

  std::shared_ptr<const void> keep_;
  CLContext* ctx_ = nullptr;
  std::string file_, name_;
  Kind kind_ = kScanPass;
  std::map<std::string, std::string> defines_;
  std::vector<Arg> args_;
  rt_buf* temp_ = nullptr;  // LocalMergeSortKernel's scratch (the whole sort runs in that one launch)
  uint64_t temp_bytes_ = 0;

  Arg& At(int index) {
    if (index < 0) RAYT_CL_CRASH("negative kernel argument index");
    if (size_t(index) >= args_.size()) args_.resize(size_t(index) + 1);
    return args_[size_t(index)];
  }
  const Arg& Get(int index) const {
    if (size_t(index) >= args_.size()) RAYT_CL_CRASH(("kernel argument not set: " + name_).c_str());
    return args_[size_t(index)];
  }

  void Init(const std::string& params) {
    // "-D NAME=value" parameters (the reference passes WIDTH, HEIGHT, MAX_ALLOWED_TRACING_STATE_SIZE)
    size_t p = 0;
    while ((p = params.find("-D", p)) != std::string::npos) {
      p += 2;
      while (p < params.size() && params[p] == ' ') ++p;
      size_t e = params.find(' ', p);
      const std::string def = params.substr(p, e == std::string::npos ? std::string::npos : e - p);
      const size_t eq = def.find('=');
      defines_[def.substr(0, eq)] = eq == std::string::npos ? "1" : def.substr(eq + 1);
      p = e == std::string::npos ? params.size() : e;
    }
    static const std::pair<const char*, Kind> kNames[] = {
        {"InitTracingFrame", kInitTracingFrame},   {"RaytracingPass", kRaytracingPass},
        {"FinishTracingFrame", kFinishTracingFrame}, {"NodeToBlockKernel", kNodeToBlock},
        {"DifferenceFlagKernel", kDifferenceFlag}, {"CompressKernel", kCompress},
        {"LocalMergeSortKernel", kLocalMergeSort}, {"GlobalMergeSortPassKernel", kGlobalMergeSortPass},
        {"ScanPassKernel", kScanPass}};
    for (auto& n : kNames)
      if (name_ == n.first) {
        kind_ = n.second;
        return;
      }
    if (name_ == "RaytraceKernel") {
      kind_ = file_.find("volatile") != std::string::npos ? kRaytraceVolatile : kRaytrace;
      return;
    }
    RAYT_CL_CRASH(("no gfx950 kernel named " + name_ + " (" + file_ + ")").c_str());
  }

  void Run(int size0, int size1, const CLEventList* wait_list, CLEvent* out_event) {
    (void)wait_list;  // one in-order stream per context: the listed events precede this launch
    rt_ctx* c = ctx_->handle();
    const int W = size0, H = size1;
    switch (kind_) {
      case kInitTracingFrame:
        rayt_cl_check(rt_init_tracing_frame(c, Get(0).buf, W, H, Get(1).f, Get(2).f[0]));
        break;
      case kRaytracingPass:  // node-level codes (nodes_in_block 0), no shadow rays: the reference kernel
        rayt_cl_check(rt_raytracing_pass(c, Get(0).buf, Get(1).buf, W, H, Get(2).buf, Get(3).buf, Get(4).i,
                                         Get(5).buf, Get(6).buf, 0, 0));
        break;
      case kFinishTracingFrame:
        rayt_cl_check(rt_finish_tracing_frame(c, Get(0).buf, Get(1).buf, W, H, Get(2).f, Get(3).buf, Get(4).buf));
        break;
      case kRaytrace:
        rayt_cl_check(rt_raytrace_depth(c, Get(0).buf, W, H, Get(1).buf, Get(2).buf, Get(3).i, Get(4).f, Get(5).f[0]));
        break;
      case kRaytraceVolatile:
        rayt_cl_check(rt_raytrace_restart(c, Get(0).buf, W, H, Get(1).buf, Get(2).i, Get(3).f, Get(4).f[0]));
        break;
      case kNodeToBlock:
        rayt_cl_check(rt_node_to_block(c, Get(0).i, Get(1).i, Get(2).buf));
        break;
      case kDifferenceFlag:
        rayt_cl_check(rt_difference_flag(c, Get(0).i, Get(1).buf, Get(2).buf));
        break;
      case kCompress:
        rayt_cl_check(rt_compress(c, Get(0).i, Get(1).buf, Get(2).buf, Get(3).buf));
        break;
      case kLocalMergeSort: {
        // The reference sorts each work-group-sized run here and merges the runs in later passes; the whole
        // array is sorted at once instead, and the later merge passes then leave it as it is.
        const int n = Get(0).i;
        if (uint64_t(n) * 4 > temp_bytes_) {
          if (temp_) rt_buf_destroy(temp_);
          temp_ = nullptr;
          temp_bytes_ = uint64_t(n > 0 ? n : 1) * 4;
          rayt_cl_check(rt_buf_create(c, temp_bytes_, &temp_));
        }
        rayt_cl_check(rt_sort_i32(c, Get(1).buf, n, temp_));
        break;
      }
      case kGlobalMergeSortPass:
        rayt_cl_check(rt_merge_pass_i32(c, Get(0).i, Get(1).i, Get(2).buf, Get(3).buf));
        break;
      case kScanPass:
        rayt_cl_check(rt_scan_pass_i32(c, Get(0).i, Get(1).i, Get(2).buf, Get(3).buf));
        break;
    }
    if (out_event) {
      rt_evt* e = nullptr;
      rayt_cl_check(rt_evt_record(c, &e));
      out_event->reset(e);
    }
  }
};

}  // namespace rayt

#endif  // RAYT_CL_UTIL_H_
