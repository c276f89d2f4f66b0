// libcatn_hip: gfx950 executor of the contraction engine.
//
// One persistent kernel launch per planning segment: workgroup b runs the op
// program of network b (a segment ends only where the host planner needs a
// Dmax-cut kept count back from the device). Site tensors of each network live
// in one device arena; descriptors (offset, dl, d, dr) are device-resident
// because virtual bond dimensions are produced by the device SVDs.
#include <hip/hip_runtime.h>

#include <string.h>

#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "capi_impl.h"

// Fixed device region of the no-allocation entry points (catn_tsvd_fixed):
// 64 MiB per GPU, held until the process ends. In host code `scratch` is a
// host-side stand-in for the device array: the device address comes from
// hipGetSymbolAddress, once per device, and only that address is passed to
// copies and kernels.
constexpr size_t kFixedDoubles = size_t(8) << 20;
static __device__ __align__(256) double scratch[kFixedDoubles];

namespace catn {

constexpr int kThreads = 256;
constexpr int kWideThreads = 1024;
constexpr int kMaxDevices = 64;
// device address of `scratch` per device (null until the device's first executor)
static void* g_fixed_ptr[kMaxDevices] = {};
// static LDS of the kernels: red[32] doubles + flag, padded
constexpr int kStaticLdsBytes = 320;

// One network per workgroup: workgroup b runs program range b of `ops` on
// instance b. `red` holds 32 reduction doubles (two halves of one slot per
// wavefront, up to 16 wavefronts).
__device__ __forceinline__ void programs_wg(InstState* states, const Op* ops, const int* ranges,
                                            int lds_cap, int qb, int jac_lanes, double* red,
                                            int* flag) {
  WG g;
  g.tid = threadIdx.x;
  g.nthr = blockDim.x;
  g.red = red;
  g.flag = flag;
  g.lds_cap = lds_cap;
  g.qb = qb;
  g.jac_lanes = jac_lanes;
  const int b = blockIdx.x;
  InstState* st = states + b;
  g.sub = st->sub_ticks;
  run_program(g, st, ops, ranges[2 * b], ranges[2 * b + 1]);
}

// Reverse sweep: workgroup b walks the tape of instance b backwards.
__device__ __forceinline__ void backward_wg(InstState* states, int lds_cap, int qb, int jac_lanes,
                                            double* red, int* flag) {
  WG g;
  g.tid = threadIdx.x;
  g.nthr = blockDim.x;
  g.red = red;
  g.flag = flag;
  g.lds_cap = lds_cap;
  g.qb = qb;
  g.jac_lanes = jac_lanes;
  InstState* st = states + blockIdx.x;
  g.sub = st->sub_ticks;
  run_backward(g, st);
}

// Up to 256 lanes per workgroup, several workgroups per CU.
__global__ __launch_bounds__(kThreads, 4) void exec_programs(InstState* states, const Op* ops,
                                                             const int* ranges, int lds_cap,
                                                             int qb, int jac_lanes) {
  __shared__ double red[32];
  __shared__ int flag;
  programs_wg(states, ops, ranges, lds_cap, qb, jac_lanes, red, &flag);
}

__global__ __launch_bounds__(kThreads, 4) void exec_backward(InstState* states, int lds_cap, int qb,
                                                             int jac_lanes) {
  __shared__ double red[32];
  __shared__ int flag;
  backward_wg(states, lds_cap, qb, jac_lanes, red, &flag);
}

// Wide workgroups (512 or 1024 lanes, up to 16 wavefronts, 4 per SIMD at
// 1024): for batches of fewer networks than CUs, where each network's ops get
// a whole CU.
__global__ __launch_bounds__(kWideThreads, 1) void exec_programs_wide(InstState* states,
                                                                      const Op* ops,
                                                                      const int* ranges,
                                                                      int lds_cap, int qb,
                                                                      int jac_lanes) {
  __shared__ double red[32];
  __shared__ int flag;
  programs_wg(states, ops, ranges, lds_cap, qb, jac_lanes, red, &flag);
}

__global__ __launch_bounds__(kWideThreads, 1) void exec_backward_wide(InstState* states,
                                                                      int lds_cap, int qb,
                                                                      int jac_lanes) {
  __shared__ double red[32];
  __shared__ int flag;
  backward_wg(states, lds_cap, qb, jac_lanes, red, &flag);
}

// Device builder of ising_network sites: block (x, y) handles sites
// x * blockDim.x + threadIdx.x of network y's instance.
__global__ __launch_bounds__(256) void build_ising_sites(const IsingNetRec* nets,
                                                         const IsingInstRec* insts,
                                                         const IsingSiteRec* sites,
                                                         const int32_t* edges, const double* w,
                                                         const double* h) {
  const IsingNetRec nr = nets[blockIdx.y];
  const IsingInstRec ir = insts[nr.inst];
  const int s = ir.site_begin + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= ir.site_end) return;
  ising_build_site(nr, ir, sites[s], edges, w, h);
}

static void hipcheck(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(-100, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

class HipExecutor : public Executor {
 public:
  explicit HipExecutor(int device) : device_(device) {
    int count = 0;
    hipcheck(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (count <= 0) throw Error(-100, "no HIP device available");
    hipcheck(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    hipcheck(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    name_ = std::string("hip-") + prop.gcnArchName;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
      tick_hz_ = 1000.0 * khz;
    (void)hipGetLastError();   // an unsupported attribute must not surface at the first launch
    // dynamic LDS a workgroup may take: the device limit less the kernels'
    // static LDS (reduction slots and flag), at most kLdsMaxDoubles
    int max_shared = 0;
    hipcheck(hipDeviceGetAttribute(&max_shared, hipDeviceAttributeMaxSharedMemoryPerBlock, device),
             "hipDeviceGetAttribute(MaxSharedMemoryPerBlock)");
    lds_limit_ = (max_shared - kStaticLdsBytes) / (int)sizeof(double);
    if (lds_limit_ > kLdsMaxDoubles) lds_limit_ = kLdsMaxDoubles;
    if (lds_limit_ < kLdsMinDoubles)
      throw Error(-100, "device LDS per workgroup is below the interpreter's minimum");
    if (device < 0 || device >= kMaxDevices) throw Error(-100, "device index out of range");
    if (!g_fixed_ptr[device]) {   // once per device, with the device current
      void* ptr = nullptr;
      hipcheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
      g_fixed_ptr[device] = ptr;
    }
    hipcheck(hipStreamCreate(&stream_), "hipStreamCreate");
    hipcheck(hipEventCreate(&ev0_), "hipEventCreate");
    hipcheck(hipEventCreate(&ev1_), "hipEventCreate");
    hipcheck(hipFuncSetAttribute((const void*)exec_programs,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(double) * lds_limit_)),
             "hipFuncSetAttribute");
    hipcheck(hipFuncSetAttribute((const void*)exec_backward,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(double) * lds_limit_)),
             "hipFuncSetAttribute");
    hipcheck(hipFuncSetAttribute((const void*)exec_programs_wide,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(double) * lds_limit_)),
             "hipFuncSetAttribute");
    hipcheck(hipFuncSetAttribute((const void*)exec_backward_wide,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(double) * lds_limit_)),
             "hipFuncSetAttribute");
  }
  ~HipExecutor() override {
    (void)hipStreamSynchronize(stream_);
    if (d_ops_) (void)hipFree(d_ops_);
    if (d_ranges_) (void)hipFree(d_ranges_);
    if (d_states_) (void)hipFree(d_states_);
    (void)hipEventDestroy(ev0_);
    (void)hipEventDestroy(ev1_);
    (void)hipStreamDestroy(stream_);
  }
  const char* name() const override { return name_.c_str(); }
  bool cut_reverse_enabled() const override { return false; }
  double tick_hz() const override { return tick_hz_; }
  int lds_limit() const override { return lds_limit_; }
  void* dmalloc(size_t bytes) override {
    void* p = nullptr;
    hipcheck(hipMalloc(&p, bytes < 8 ? 8 : bytes), "hipMalloc");
    return p;
  }
  void dfree(void* p) override {
    if (!p) return;
    (void)hipStreamSynchronize(stream_);   // nothing in flight may still use it
    (void)hipFree(p);
    (void)hipGetLastError();               // a failed free must not surface at a later launch
  }
  void h2d(void* dst, const void* src, size_t bytes) override {
    hipcheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    hipcheck(hipStreamSynchronize(stream_), "h2d sync");
  }
  void d2h(void* dst, const void* src, size_t bytes) override {
    hipcheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hipcheck(hipStreamSynchronize(stream_), "d2h sync");
  }
  void d2d(void* dst, const void* src, size_t bytes) override {
    hipcheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_), "d2d");
    hipcheck(hipStreamSynchronize(stream_), "d2d sync");
  }
  void build_ising(const IsingBuild& b) override {
    if (b.nets.empty() || b.max_sites == 0) return;
    if (b.nets.size() > 65535) throw Error(-2, "ising_network_batch: more than 65535 networks");
    // one upload of the packed records and arrays, one launch
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_n = 0, o_i = al(o_n + sizeof(IsingNetRec) * b.nets.size());
    const size_t o_s = al(o_i + sizeof(IsingInstRec) * b.insts.size());
    const size_t o_e = al(o_s + sizeof(IsingSiteRec) * b.sites.size());
    const size_t o_w = al(o_e + sizeof(int32_t) * 2 * (size_t)b.nedges);
    const size_t o_h = al(o_w + sizeof(double) * (size_t)b.nedges);
    const size_t total = o_h + sizeof(double) * (size_t)b.nfields;
    std::vector<char> host(total);
    memcpy(host.data() + o_n, b.nets.data(), sizeof(IsingNetRec) * b.nets.size());
    memcpy(host.data() + o_i, b.insts.data(), sizeof(IsingInstRec) * b.insts.size());
    memcpy(host.data() + o_s, b.sites.data(), sizeof(IsingSiteRec) * b.sites.size());
    if (b.nedges) {
      memcpy(host.data() + o_e, b.edges, sizeof(int32_t) * 2 * (size_t)b.nedges);
      memcpy(host.data() + o_w, b.w, sizeof(double) * (size_t)b.nedges);
    }
    memcpy(host.data() + o_h, b.h, sizeof(double) * (size_t)b.nfields);
    char* d = (char*)dmalloc(total);
    try {
      h2d(d, host.data(), total);
      const dim3 grid((unsigned)((b.max_sites + 255) / 256), (unsigned)b.nets.size());
      hipcheck(hipGetLastError(), "before launch build_ising_sites");
      hipLaunchKernelGGL(build_ising_sites, grid, dim3(256), 0, stream_, (const IsingNetRec*)(d + o_n),
                         (const IsingInstRec*)(d + o_i), (const IsingSiteRec*)(d + o_s),
                         (const int32_t*)(d + o_e), (const double*)(d + o_w),
                         (const double*)(d + o_h));
      finish_launch("build_ising_sites");
    } catch (...) {
      dfree(d);
      throw;
    }
    dfree(d);
  }
  // a second executor on the same device: its own stream, events and
  // staging buffers (the networks' memory stays with this one)
  Executor* sibling() override { return new HipExecutor(device_); }
  void bind_thread() override { hipcheck(hipSetDevice(device_), "hipSetDevice"); }
  void* fixed_region(size_t* bytes) override {
    if (bytes) *bytes = sizeof(double) * kFixedDoubles;
    return g_fixed_ptr[device_];
  }
  // one instance in caller-provided device memory: no allocation (the
  // interpreter kernel, its launch shape for a batch of one, the copies)
  void run_fixed(InstState* hs, void* d_state, const Op& op, void* d_op,
                 void* d_ranges) override {
    const int lds_cap = lds_for(1);
    if (lds_cap > lds_limit_) throw Error(-2, "launch LDS budget above the device limit");
    const int32_t rng[2] = {0, 1};
    hipcheck(hipMemcpyAsync(d_op, &op, sizeof(Op), hipMemcpyHostToDevice, stream_), "op");
    hipcheck(hipMemcpyAsync(d_ranges, rng, sizeof(rng), hipMemcpyHostToDevice, stream_), "range");
    hipcheck(hipMemcpyAsync(d_state, hs, sizeof(InstState), hipMemcpyHostToDevice, stream_), "state");
    hipcheck(hipEventRecord(ev0_, stream_), "event record");
    hipcheck(hipGetLastError(), "before launch exec_programs");
    hipLaunchKernelGGL(threads > kThreads ? exec_programs_wide : exec_programs, dim3(1),
                       dim3(threads), sizeof(double) * lds_cap, stream_, (InstState*)d_state, (const Op*)d_op, (const int*)d_ranges, lds_cap, qb,
                       jac_lanes);
    finish_launch("exec_programs (fixed region)");
    hipcheck(hipMemcpyAsync(hs, d_state, sizeof(InstState), hipMemcpyDeviceToHost, stream_),
             "state back");
    hipcheck(hipStreamSynchronize(stream_), "state back sync");
  }
  void run(const std::vector<InstState*>& states,
           const std::vector<const std::vector<Op>*>& programs) override {
    const int nb = (int)states.size();
    if (nb == 0) return;
    // concatenated programs; a program shared by several networks is uploaded once
    std::vector<Op> all;
    std::vector<int> ranges(2 * nb);
    std::unordered_map<const std::vector<Op>*, int> at;
    for (int b = 0; b < nb; ++b) {
      auto it = at.find(programs[b]);
      if (it != at.end()) {
        ranges[2 * b] = ranges[2 * it->second];
        ranges[2 * b + 1] = ranges[2 * it->second + 1];
        continue;
      }
      at.emplace(programs[b], b);
      ranges[2 * b] = (int)all.size();
      all.insert(all.end(), programs[b]->begin(), programs[b]->end());
      ranges[2 * b + 1] = (int)all.size();
    }
    if (all.empty()) return;
    const int lds_cap = lds_for(nb);
    if (lds_cap > lds_limit_) throw Error(-2, "launch LDS budget above the device limit");
    reserve(all.size(), nb);
    // stage every instance state into one contiguous array: one copy each way
    std::vector<InstState> packed(nb);
    for (int b = 0; b < nb; ++b) packed[b] = *states[b];
    hipcheck(hipMemcpyAsync(d_ops_, all.data(), sizeof(Op) * all.size(), hipMemcpyHostToDevice, stream_), "ops");
    hipcheck(hipMemcpyAsync(d_ranges_, ranges.data(), sizeof(int) * ranges.size(), hipMemcpyHostToDevice, stream_), "ranges");
    hipcheck(hipMemcpyAsync(d_states_, packed.data(), sizeof(InstState) * nb, hipMemcpyHostToDevice, stream_), "states");
    hipcheck(hipEventRecord(ev0_, stream_), "event record");
    hipcheck(hipGetLastError(), "before launch exec_programs");
    hipLaunchKernelGGL(threads > kThreads ? exec_programs_wide : exec_programs, dim3(nb),
                       dim3(threads), sizeof(double) * lds_cap, stream_, d_states_, d_ops_,
                       d_ranges_, lds_cap, qb, jac_lanes);
    finish_launch("exec_programs");
    float ms = 0.0f;
    hipcheck(hipEventElapsedTime(&ms, ev0_, ev1_), "event elapsed");
    launches += 1;
    kernel_seconds += 1e-3 * ms;
    // on this executor's stream: a null-stream copy would wait for the
    // kernels of sibling executors' streams too
    hipcheck(hipMemcpyAsync(packed.data(), d_states_, sizeof(InstState) * nb,
                            hipMemcpyDeviceToHost, stream_), "states back");
    hipcheck(hipStreamSynchronize(stream_), "states back sync");
    for (int b = 0; b < nb; ++b) *states[b] = packed[b];
  }

  void run_backward(const std::vector<InstState*>& states, int lds_cap) override {
    const int nb = (int)states.size();
    if (nb == 0) return;
    if (lds_cap <= 0 || lds_cap > lds_limit_) lds_cap = lds_limit_;
    reserve(0, nb);
    std::vector<InstState> packed(nb);
    for (int b = 0; b < nb; ++b) packed[b] = *states[b];
    hipcheck(hipMemcpyAsync(d_states_, packed.data(), sizeof(InstState) * nb, hipMemcpyHostToDevice, stream_), "states");
    hipcheck(hipEventRecord(ev0_, stream_), "event record");
    hipcheck(hipGetLastError(), "before launch exec_backward");
    hipLaunchKernelGGL(threads > kThreads ? exec_backward_wide : exec_backward, dim3(nb),
                       dim3(threads), sizeof(double) * lds_cap, stream_, d_states_, lds_cap, qb,
                       jac_lanes);
    finish_launch("exec_backward");
    float ms = 0.0f;
    hipcheck(hipEventElapsedTime(&ms, ev0_, ev1_), "event elapsed");
    bwd_launches += 1;
    bwd_seconds += 1e-3 * ms;
    // on this executor's stream: a null-stream copy would wait for the
    // kernels of sibling executors' streams too
    hipcheck(hipMemcpyAsync(packed.data(), d_states_, sizeof(InstState) * nb,
                            hipMemcpyDeviceToHost, stream_), "states back");
    hipcheck(hipStreamSynchronize(stream_), "states back sync");
    for (int b = 0; b < nb; ++b) *states[b] = packed[b];
  }

 private:
  // After a launch: the stream is always drained before any error leaves the
  // executor, so no kernel can outlive the device memory the caller frees
  // while it unwinds.
  void finish_launch(const char* what) {
    const hipError_t launch = hipGetLastError();
    const hipError_t rec = hipEventRecord(ev1_, stream_);
    const hipError_t sync = hipStreamSynchronize(stream_);
    hipcheck(launch, what);
    hipcheck(sync, what);
    hipcheck(rec, "event record");
  }
  void reserve(size_t nops, int nb) {
    if (nops > ops_cap_) {
      if (d_ops_) (void)hipFree(d_ops_);
      ops_cap_ = nops * 2;
      hipcheck(hipMalloc(&d_ops_, sizeof(Op) * ops_cap_), "hipMalloc ops");
    }
    if ((size_t)nb > st_cap_) {
      if (d_ranges_) (void)hipFree(d_ranges_);
      if (d_states_) (void)hipFree(d_states_);
      st_cap_ = (size_t)nb * 2;
      hipcheck(hipMalloc(&d_ranges_, sizeof(int) * 2 * st_cap_), "hipMalloc ranges");
      hipcheck(hipMalloc(&d_states_, sizeof(InstState) * st_cap_), "hipMalloc states");
    }
  }
  int device_;
  std::string name_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  Op* d_ops_ = nullptr;
  int* d_ranges_ = nullptr;
  InstState* d_states_ = nullptr;
  size_t ops_cap_ = 0, st_cap_ = 0;
  double tick_hz_ = 1e8;
  int lds_limit_ = 0;
};

Executor* make_executor(int device) { return new HipExecutor(device); }

}  // namespace catn
