// Drives the CL util shims (include/rayt_cl_util.h) the way the reference's
// callers drive CL util: GPUSort and GPUScan pass by pass with event lists
// (GPU algorithms/gpu-sort.cpp:16-49, gpu-scan.cpp:15-43), their checks
// TestGPUSort / TestGPUScan (application/test-application.cpp:150-240), and
// one frame of GPURayTracer (gpu-ray-tracer.cpp:66-227: InitTracingFrame,
// one RaytracingPass, NodeToBlock over the codes, FinishTracingFrame) over a
// block pool that holds the whole tree.
//
// usage: cl_util_shim_test SORT_N TREE VPI_FILE W H IMAGE_OUT CODES_OUT
// Prints "sort passed", "scan passed" and "frame done"; exits non-zero on a
// mismatch. The frame's outputs are compared with the oracle by the caller.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <vector>

#include "rayt_cl_util.h"

using namespace rayt;

struct Mat4 {
  float m[16];
};
struct Vec3 {
  float x, y, z;
};

// GPUSort over the shims: a local sort of work-group runs, then merge passes of doubling run length.
class ShimSort {
 public:
  ShimSort(int length, std::shared_ptr<CLContext> ctx)
      : length_(length), temp_(0, length * 4, ctx), local_("GPU algorithms/kernels/merge_sort.cl", "",
                                                           "LocalMergeSortKernel", ctx),
        pass_("GPU algorithms/kernels/merge_sort.cl", "", "GlobalMergeSortPassKernel", ctx) {}
  void Sort(CLBuffer& data) {
    CLEventList wait;
    CLEvent ev;
    local_.SetIntArg(0, length_);
    local_.SetBufferArg(1, &data);
    local_.SetLocalBufferArg(2, local_.work_group_size() * 4);
    local_.Run1D(length_, nullptr, &ev);
    CLBuffer* in = &data;
    CLBuffer* out = &temp_;
    for (int len = local_.work_group_size(); len < length_; len *= 2) {
      wait.Clear();
      wait.Add(ev);
      pass_.SetIntArg(0, length_);
      pass_.SetIntArg(1, len);
      pass_.SetBufferArg(2, in);
      pass_.SetBufferArg(3, out);
      pass_.Run1D(length_, &wait, &ev);
      std::swap(in, out);
    }
    if (in != &data) {
      wait.Clear();
      wait.Add(ev);
      data.CopyFrom(*in, &wait, &ev);
    }
    ev.WaitFor();
  }

 private:
  int length_;
  CLBuffer temp_;
  CLKernel local_, pass_;
};

// GPUScan over the shims: Hillis-Steele passes ping-ponging between the data and a temporary.
class ShimScan {
 public:
  ShimScan(int length, std::shared_ptr<CLContext> ctx)
      : length_(length), temp_(0, length * 4, ctx), pass_("GPU algorithms/kernels/scan.cl", "", "ScanPassKernel", ctx) {}
  void Scan(CLBuffer& data) {
    CLEventList wait;
    CLEvent ev;
    CLBuffer* in = &data;
    CLBuffer* out = &temp_;
    for (int len = 1; len < length_; len *= 2) {
      pass_.SetIntArg(0, length_);
      pass_.SetIntArg(1, len);
      pass_.SetBufferArg(2, in);
      pass_.SetBufferArg(3, out);
      pass_.Run1D(length_, &wait, &ev);
      wait.Clear();
      wait.Add(ev);
      std::swap(in, out);
    }
    if (in != &data) {
      data.CopyFrom(*in, &wait, &ev);
      wait.Clear();
      wait.Add(ev);
    }
    wait.WaitFor();
  }

 private:
  int length_;
  CLBuffer temp_;
  CLKernel pass_;
};

static bool TestSortAndScan(int n, std::shared_ptr<CLContext> ctx, std::mt19937& rng) {
  std::uniform_int_distribution<int> small(-100, 100);
  std::vector<int> src(n), expected(n), found(n);
  ShimSort srt(n, ctx);
  ShimScan scan(n, ctx);
  CLBuffer buf(0, n * 4, ctx);
  for (int test = 0; test < 10; ++test) {
    for (int& v : src) v = small(rng);
    expected = src;
    std::sort(expected.begin(), expected.end());
    buf.Write(0, n * 4, src.data(), true, nullptr, nullptr);
    ctx->WaitForAll();
    srt.Sort(buf);
    buf.Read(0, n * 4, found.data());
    if (found != expected) {
      std::printf("sort failed: n %d test %d\n", n, test);
      return false;
    }
    expected = src;
    for (int i = 1; i < n; ++i) expected[i] = int(uint32_t(expected[i]) + uint32_t(expected[i - 1]));
    buf.Write(0, n * 4, src.data(), true, nullptr, nullptr);
    ctx->WaitForAll();
    scan.Scan(buf);
    buf.Read(0, n * 4, found.data());
    if (found != expected) {
      std::printf("scan failed: n %d test %d\n", n, test);
      return false;
    }
  }
  return true;
}

int main(int argc, char** argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s SORT_N TREE VPI_FILE W H IMAGE_OUT CODES_OUT\n", argv[0]);
    return 2;
  }
  auto ctx = std::make_shared<CLContext>();
  std::mt19937 rng(2013);
  // the reference's size (2013, not a multiple of the work group), then the caller's
  for (int n : {2013, std::atoi(argv[1])})
    if (!TestSortAndScan(n, ctx, rng)) return 1;
  std::printf("sort passed\nscan passed\n");

  const int W = std::atoi(argv[4]), H = std::atoi(argv[5]);
  Mat4 vpi;
  FILE* f = std::fopen(argv[3], "rb");
  if (!f || std::fread(vpi.m, 4, 16, f) != 16) return 2;
  std::fclose(f);
  rt_tree* tree = nullptr;
  rayt_cl_check(rt_tree_open(argv[2], &tree));
  rt_tree_info info;
  rayt_cl_check(rt_tree_info_get(tree, &info));
  rt_cache* cache = nullptr;
  rayt_cl_check(rt_cache_create(ctx->handle(), tree, int(info.blocks_count), &cache));
  rayt_cl_check(rt_cache_initial_fill(cache));
  const int P = W * H;
  const int nib = rt_cache_nodes_in_block(cache);
  {
    CLBuffer links(rt_cache_channel(cache, 0)), far_ptrs(rt_cache_far_pointers(cache));
    CLBuffer normals(rt_cache_channel_by_name(cache, "normals")), colors(rt_cache_channel_by_name(cache, "colors"));
    CLBuffer out_image(CL_MEM_WRITE_ONLY, 4 * P, ctx), faults_and_hits(0, 4 * P, ctx), states(0, 48 * P, ctx);
    char params[300];
    std::snprintf(params, sizeof(params), "-D WIDTH=%d -D HEIGHT=%d -D MAX_ALLOWED_TRACING_STATE_SIZE=%d", W, H, 48);
    CLKernel init("GPU_renderer/kernels", "init_tracing_frame.cl", params, "InitTracingFrame", ctx);
    CLKernel pass("GPU_renderer/kernels", "raytracing_pass.cl", params, "RaytracingPass", ctx);
    CLKernel finish("GPU_renderer/kernels", "finish_tracing_frame.cl", params, "FinishTracingFrame", ctx);
    CLKernel to_block("GPU_renderer/kernels", "feedback_compression.cl", "", "NodeToBlockKernel", ctx);
    CLEvent ev;
    CLEventList wait;
    init.SetBufferArg(0, &states);
    init.SetFloat16Arg(1, vpi);
    init.SetFloatArg(2, 0.0f);
    init.Run2D(W, H, nullptr, &ev);
    wait.Add(ev);
    pass.SetBufferArg(0, &faults_and_hits);
    pass.SetBufferArg(1, &states);
    pass.SetBufferArg(2, &links);
    pass.SetBufferArg(3, &far_ptrs);
    pass.SetIntArg(4, rt_cache_root_node_index(cache));
    pass.SetBufferArg(5, &normals);
    pass.SetBufferArg(6, &colors);
    pass.Run2D(W, H, &wait, &ev);
    wait.Clear();
    wait.Add(ev);
    to_block.SetIntArg(0, P);
    to_block.SetIntArg(1, nib);
    to_block.SetBufferArg(2, &faults_and_hits);
    to_block.Run1D(P, &wait, &ev);
    wait.Clear();
    wait.Add(ev);
    finish.SetBufferArg(0, &out_image);
    finish.SetBufferArg(1, &states);
    finish.SetFloat4Arg(2, Vec3{0.0f, float(216 / 255.), 1.0f}, 1.0f);
    finish.SetBufferArg(3, &normals);
    finish.SetBufferArg(4, &colors);
    finish.Run2D(W, H, &wait, &ev);
    ev.WaitFor();
    std::vector<uint32_t> img(P);
    std::vector<int> codes(P);
    out_image.Read(0, 4 * P, img.data());
    faults_and_hits.Read(0, 4 * P, codes.data());
    FILE* fi = std::fopen(argv[6], "wb");
    FILE* fc = std::fopen(argv[7], "wb");
    if (!fi || !fc) return 2;
    std::fwrite(img.data(), 4, P, fi);
    std::fwrite(codes.data(), 4, P, fc);
    std::fclose(fi);
    std::fclose(fc);
  }
  rt_cache_destroy(cache);
  rt_tree_destroy(tree);
  std::printf("frame done\n");
  return 0;
}
