// CPU emulator of libcatn_hip — TEST BUILD ONLY.
//
// Compiles the product's engine (engine.h), planner and C ABI with g++ and runs
// each "workgroup" as a single CPU thread (barriers become no-ops). It lets the
// CPU test suite exercise the exact device algorithm and host bookkeeping
// without a GPU. It is never loaded by the product path (catn_amd), which
// requires the HIP library.
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <vector>

#include "../../catn.jl_amd/csrc/engine.h"
#include "../../catn.jl_amd/csrc/capi_impl.h"

namespace catn {

// launches run by sibling executors (the concurrent groups of a batch)
static std::atomic<long> g_sibling_runs{0};

class CpuExecutor : public Executor {
 public:
  const char* name() const override { return "cpu-emulator"; }
  // concurrent batch groups run on host threads (the LDS stand-in is per thread)
  Executor* sibling() override { return new CpuExecutor(); }
  // CATN_EMU_POISON=1 fills new device blocks with 0xFF bytes (NaN doubles,
  // -1 integers), as hipMalloc gives no zeroed memory: a read of memory the
  // engine never wrote then shows up in the results.
  void* dmalloc(size_t bytes) override {
    void* p = calloc(1, bytes < 8 ? 8 : bytes);
    if (p && getenv("CATN_EMU_POISON")) memset(p, 0xFF, bytes < 8 ? 8 : bytes);
    return p;
  }
  void dfree(void* p) override { free(p); }
  // CATN_EMU_REFUSE_CUT_REVERSE=1 behaves like the gfx950 executor, which
  // refuses gradient runs across a Dmax cut
  bool cut_reverse_enabled() const override { return !getenv("CATN_EMU_REFUSE_CUT_REVERSE"); }
  void h2d(void* dst, const void* src, size_t bytes) override { memcpy(dst, src, bytes); }
  void d2h(void* dst, const void* src, size_t bytes) override { memcpy(dst, src, bytes); }
  void d2d(void* dst, const void* src, size_t bytes) override { memcpy(dst, src, bytes); }
  void build_ising(const IsingBuild& b) override {
    for (const IsingNetRec& nr : b.nets) {
      const IsingInstRec& ir = b.insts[nr.inst];
      for (int s = ir.site_begin; s < ir.site_end; ++s)
        ising_build_site(nr, ir, b.sites[s], b.edges, b.w, b.h);
    }
  }
  void* fixed_region(size_t* bytes) override {
    static double region[size_t(8) << 20];   // the device library's 64 MiB fixed region
    if (bytes) *bytes = sizeof(region);
    return region;
  }
  void run_fixed(InstState* hs, void* d_state, const Op& op, void* d_op, void*) override {
    memcpy(d_state, hs, sizeof(InstState));
    memcpy(d_op, &op, sizeof(Op));
    std::vector<InstState*> st = {(InstState*)d_state};
    std::vector<Op> prog = {op};
    std::vector<const std::vector<Op>*> pr = {&prog};
    run(st, pr);
    memcpy(hs, d_state, sizeof(InstState));
  }
  void run(const std::vector<InstState*>& states,
           const std::vector<const std::vector<Op>*>& programs) override {
    if (parent) g_sibling_runs.fetch_add(1);
    double red[32];
    int flag = 0;
    WG g;
    g.tid = 0;
    g.nthr = 1;
    g.red = red;
    g.flag = &flag;
    g.lds_cap = lds_for((int)states.size());
    g.qb = qb;
    g.jac_lanes = jac_lanes;
    if (getenv("CATN_EMU_LDS")) g.lds_cap = atoi(getenv("CATN_EMU_LDS"));   // a device batch's budget
    for (size_t b = 0; b < states.size(); ++b) {
      const std::vector<Op>& p = *programs[b];
      g.sub = states[b]->sub_ticks;
      run_program(g, states[b], p.data(), 0, (int)p.size());
    }
  }
  void run_backward(const std::vector<InstState*>& states, int lds_cap) override {
    double red[32];
    int flag = 0;
    WG g;
    g.tid = 0;
    g.nthr = 1;
    g.red = red;
    g.flag = &flag;
    g.lds_cap = lds_cap > 0 ? lds_cap : kLdsMaxDoubles;
    g.qb = qb;
    g.jac_lanes = jac_lanes;
    for (InstState* st : states) {
      g.sub = st->sub_ticks;
      catn::run_backward(g, st);
    }
  }
};

Executor* make_executor(int) { return new CpuExecutor(); }

}  // namespace catn

extern "C" long catn_emu_sibling_runs() { return catn::g_sibling_runs.load(); }
