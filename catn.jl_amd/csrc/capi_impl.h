// Host side of libcatn_hip: network construction (ising_network,
// TensorNetwork), the contraction! driver loop around the planner, the
// MPSNode primitive entry points and tsvd. Shared by every executor; the
// executor only moves bytes and runs programs on the device.
#pragma once
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <chrono>
#include <cmath>
#include <exception>
#include <map>
#include <mutex>
#include <thread>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/catn_hip.h"
#include "ops.h"
#include <unordered_map>

#include "planner.h"

namespace catn {

// Packed description of a batch of ising networks whose site values are
// computed on the device (catn_ising_network_batch): per instance its sites'
// records (arena offset, kind, edge, side, node) and its slices of the
// concatenated edge / coupling / field arrays; per network its instance, arena,
// storage type, beta and pinning.
struct IsingSiteRec { int64_t off; int32_t kind, edge, side, node; };
struct IsingInstRec { int32_t site_begin, site_end; int64_t edge_off, field_off; };
struct IsingNetRec {
  double* arena;
  double beta;
  int32_t inst, f32, pos1, val1, pos2, val2;
};
struct IsingBuild {
  std::vector<IsingSiteRec> sites;
  std::vector<IsingInstRec> insts;
  std::vector<IsingNetRec> nets;
  const int32_t* edges = nullptr;
  const double* w = nullptr;
  const double* h = nullptr;
  int64_t nedges = 0, nfields = 0;   // lengths of the concatenated arrays
  int32_t max_sites = 0;             // largest site count of an instance
};

// one site of a packed ising build (device kernel thread or host loop)
CATN_HD void ising_build_site(const IsingNetRec& nr, const IsingInstRec& ir, const IsingSiteRec& sr,
                             const int32_t* edges, const double* w, const double* h) {
  double sn[2] = {1.0, -1.0}, so[2] = {1.0, -1.0}, v[8];
  double wv = 0.0;
  const double hv = h[ir.field_off + sr.node - 1];
  if (sr.kind != 4) {
    const int i = edges[2 * (ir.edge_off + sr.edge)], j = edges[2 * (ir.edge_off + sr.edge) + 1];
    ising_spins(sr.side == 0 ? i : j, nr.pos1, nr.val1, nr.pos2, nr.val2, sn);
    ising_spins(sr.side == 0 ? j : i, nr.pos1, nr.val1, nr.pos2, nr.val2, so);
    wv = w[ir.edge_off + sr.edge];
  }
  const int cnt = ising_site_entries(sr.kind, sr.side, sn, so, wv, hv, nr.beta, v);
  if (nr.f32) {
    float* dst = (float*)nr.arena + sr.off;
    for (int q = 0; q < cnt; ++q) dst[q] = (float)v[q];
  } else {
    double* dst = nr.arena + sr.off;
    for (int q = 0; q < cnt; ++q) dst[q] = v[q];
  }
}

class Executor {
 public:
  virtual ~Executor() = default;
  virtual const char* name() const = 0;
  virtual void* dmalloc(size_t bytes) = 0;
  virtual void dfree(void* p) = 0;
  virtual void h2d(void* dst, const void* src, size_t bytes) = 0;
  virtual void d2h(void* dst, const void* src, size_t bytes) = 0;
  // Execute programs[k] on instance k whose state (host mirror) is states[k]:
  // the executor moves the states to the device, runs one workgroup per
  // instance and writes the updated states back; returns after completion.
  virtual void run(const std::vector<InstState*>& states,
                   const std::vector<const std::vector<Op>*>& programs) = 0;
  // Reverse sweep over the recorded tape of every instance (one workgroup
  // each, one launch) with an LDS budget of lds_cap doubles.
  virtual void run_backward(const std::vector<InstState*>& states, int lds_cap) = 0;
  // frequency of the device wall clock read by the op timers (Hz)
  virtual double tick_hz() const { return 1.0; }
  // A sibling on the same device with its own stream and staging buffers,
  // used by one host thread while others use theirs (concurrent groups of a
  // batch); null when this executor has none. The caller owns it and copies
  // the launch shape into it.
  virtual Executor* sibling() { return nullptr; }
  // Make the calling host thread use this executor's device.
  virtual void bind_thread() {}
  // The executor that owns the networks' memory when this one is a sibling.
  Executor* parent = nullptr;
  // streams (concurrent groups) of a forward batch of distinct networks
  // (1..kGroupStreams; 4 = the HIP runtime's default hardware queues)
  int batch_streams = 4;
  // Whether the reverse sweep may run over a Dmax cut (cut_bondim_opt!) on
  // this executor. The gfx950 executor refuses it: gradient runs of
  // networks with a Dmax cut faulted the GPU and the cause is not yet known,
  // so such a run is rejected on the host before any of its launches.
  virtual bool cut_reverse_enabled() const { return true; }
  // launches of the interpreter kernel and their summed duration (seconds,
  // measured with events on the stream the kernel runs on)
  int64_t launches = 0;
  double kernel_seconds = 0.0;
  // the same for the reverse-sweep kernel
  int64_t bwd_launches = 0;
  double bwd_seconds = 0.0;
  // host-side time: planning, host<->device copies + launch overhead
  double plan_seconds = 0.0, run_seconds = 0.0;
  // workgroup size of the interpreter (64, 128 or 256 lanes; 512 or 1024 run
  // the wide kernels, one workgroup per CU)
  int threads = 256;
  // LDS budget per workgroup in doubles; 0 = lds_doubles_for_batch(batch)
  int lds_override = 0;
  // tile parameters: Householder panel width, global Jacobi lanes per pair.
  // The reverse sweep replays the forward's factorisations with the same ones.
  int qb = kQBDefault, jac_lanes = kJacLanesDefault;
  int lds_for(int nb) const {
    if (lds_override > 0) return lds_override;
    const int a = lds_doubles_for_batch(nb), lim = lds_limit();
    return a < lim ? a : lim;
  }
  // the largest LDS budget (doubles) a launch of this executor may take
  virtual int lds_limit() const { return kLdsMaxDoubles; }
  // device-to-device copy on the executor's stream (returns after completion)
  virtual void d2d(void* dst, const void* src, size_t bytes) = 0;
  // Fill the device-built sites of ising networks (catn_ising_network_batch):
  // for every network and every site of its instance, ising_site_entries
  // written into the network's arena. Returns after completion.
  virtual void build_ising(const IsingBuild& b) = 0;
  // The library's fixed device region of the no-allocation entry points: its
  // device address on this executor's device and its size in bytes (null if
  // there is none). The library owns it for the life of the process.
  virtual void* fixed_region(size_t* bytes) = 0;
  // Run `op` on one instance whose state lives in caller-provided device
  // memory: hs (host mirror) is copied to d_state, the op to d_op and its
  // range to d_ranges (8 bytes), one workgroup runs it, and hs is read back.
  // Allocates nothing.
  virtual void run_fixed(InstState* hs, void* d_state, const Op& op, void* d_op,
                         void* d_ranges) = 0;
};

Executor* make_executor(int device);   // provided by the executor TU

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& s) : std::runtime_error(s), code(c) {}
};

}  // namespace catn

struct catn_net;

struct catn_ctx {
  std::unique_ptr<catn::Executor> ex;
  // live networks of this context: destroying the context releases their
  // device memory first (host languages may finalise the two in any order)
  std::set<catn_net*> nets;
};

struct catn_net {
  catn_ctx* ctx = nullptr;
  catn::Planner pl;
  int n = 0;             // tn.n
  double beta = 1.0;     // tn.beta
  int num_isolated = 0;
  bool consumed = false;
  // host staging of site tensors not yet on the device
  // dev: allocated in the arena and filled on the device (no host data)
  struct Staged { int slot; int dl, d, dr; std::vector<double> data; bool dev = false; };
  std::vector<Staged> staged;
  int nslots = 0;
  // device state
  catn::InstState hs{};          // authoritative between launches (the executor
                                 // stages it to the device for each launch)
  bool allocated = false;
  int slot_cap = 0, res_cap = 0;
  int64_t arena_cap = 0, scratch_cap = 0;
  int64_t arena_hwm = 0;
  // initial site tensors: arena offset of every slot uploaded with data
  std::vector<int64_t> off0;
  // ising_network provenance of the initial sites (chain rule of the adjoint)
  struct IsingSite { int slot, edge, side, node, kind; };   // kind: 0 leaf, 1 first,
                                                           // 2 last, 3 interior, 4 isolated
  bool is_ising = false;
  std::vector<IsingSite> isites;
  std::vector<int32_t> iedges;
  std::vector<double> iw, ih;
  int ipos1 = 0, ival1 = 0, ipos2 = 0, ival2 = 0;
  bool have_adjoint = false;
  // per tensor added by catn_net_add_tensor_node / catn_tensor_network, in
  // order: the slot whose initial version holds its data, and its size
  std::vector<std::pair<int, int64_t>> tensor_slot;
  // final planner after contraction!, shared by the networks that followed one
  // plan (read-only); plan() reads it, own_plan() takes a private copy
  std::shared_ptr<const catn::Planner> done_pl;
  const catn::Planner& plan() const { return done_pl ? *done_pl : pl; }
  catn::Planner& own_plan() {
    if (done_pl) {
      pl = *done_pl;
      done_pl.reset();
    }
    return pl;
  }
  void finish_plan() { done_pl = std::make_shared<const catn::Planner>(std::move(pl)); }
  bool f32 = false;        // FP32 site storage (catn_params.precision = 1 or 2)
  int bf16_terms = 2;      // BF16 terms per FP32 operand of the theta GEMM
  bool cplx = false;       // complex site tensors (catn_params.eltype = 1)
  size_t elem_bytes() const { return (cplx ? 2 : 1) * (f32 ? sizeof(float) : sizeof(double)); }
  // doubles per element of host data (interleaved re, im when complex)
  int ed() const { return cplx ? 2 : 1; }

  ~catn_net() {
    release();
    if (ctx) ctx->nets.erase(this);
  }

  void attach(catn_ctx* c) {
    ctx = c;
    c->nets.insert(this);
  }

  void release() {
    if (!ctx || !allocated) return;
    catn::Executor* ex = ctx->ex.get();
    ex->dfree(hs.arena); ex->dfree(hs.scratch); ex->dfree(hs.sites);
    ex->dfree(hs.results); ex->dfree(hs.dresults);
    if (hs.tape) ex->dfree(hs.tape);
    if (hs.adj) ex->dfree(hs.adj);
    hs.tape = nullptr; hs.adj = nullptr;
    allocated = false;
  }

  // switch on the tape for the coming contraction (before any op has run)
  // the reverse sweep's device memory: a tape of `cap` records and the
  // adjoint arena (allocated once, kept for a later gradient run)
  // `cap` records for the contraction behind the build_tape kept ones
  void alloc_record(int cap) {
    ensure_device();
    alloc_tape(cap + build_tape);
    if (!hs.adj) hs.adj = (double*)ctx->ex->dmalloc(sizeof(double) * (size_t)arena_cap);
  }

  // records kept from the construction run (the raw2mps splits of a network
  // built from tensors) at the start of the tape: the reverse sweep ends with
  // them and so reaches the input tensors
  int build_tape = 0;
  bool tensors_taped = false;   // built by catn_tensor_network, splits on the tape

  // a device tape of at least `total` records that keeps its first build_tape
  void alloc_tape(int total) {
    if (hs.tape && hs.tape_cap >= total) return;
    catn::Executor* ex = ctx->ex.get();
    auto* t = (catn::TapeRec*)ex->dmalloc(sizeof(catn::TapeRec) * (size_t)total);
    if (hs.tape) {
      if (build_tape > 0) ex->d2d(t, hs.tape, sizeof(catn::TapeRec) * (size_t)build_tape);
      ex->dfree(hs.tape);
    }
    hs.tape = t;
    hs.tape_cap = total;
  }

  void enable_record(int cap) {
    ensure_device();
    if (hs.tape_len != build_tape || !pl.prog.empty())
      throw catn::Error(-5, "gradient: the network has already executed device ops");
    alloc_record(cap);
    hs.record = 1;
  }

  int new_device_slot(int dl, int d, int dr) {
    Staged s;
    s.slot = nslots++;
    s.dl = dl; s.d = d; s.dr = dr;
    s.dev = true;
    staged.push_back(std::move(s));
    return staged.back().slot;
  }

  int new_slot(int dl, int d, int dr, const double* data) {
    Staged s;
    s.slot = nslots++;
    s.dl = dl; s.d = d; s.dr = dr;
    const size_t cnt = (size_t)dl * d * dr * ed();
    if (data) s.data.assign(data, data + cnt);
    else s.data.clear();   // placeholder slot, descriptor written by a device op
    staged.push_back(std::move(s));
    return staged.back().slot;
  }

  void ensure_device() {
    if (!ctx) throw catn::Error(-5, "the network's context has been destroyed");
    catn::Executor* ex = ctx->ex.get();
    if (!allocated) {
      int64_t staged_elems = 0;
      for (auto& s : staged)
        staged_elems += ((s.dev ? (int64_t)s.dl * s.d * s.dr : (int64_t)s.data.size() / ed()) + 31) / 32 * 32;
      if (arena_cap <= 0) arena_cap = std::max<int64_t>(int64_t(1) << 24, 4 * staged_elems);
      if (scratch_cap <= 0) scratch_cap = int64_t(1) << 22;
      arena_cap = std::max<int64_t>(arena_cap, 2 * staged_elems + 1024);
      slot_cap = std::max(1024, 2 * nslots);
      res_cap = 4096;
      hs = catn::InstState{};
      hs.psi = 1.0; hs.sign = 1.0;
      hs.seed = pl.p.seed;
      hs.f32 = f32 ? (bf16_terms == 3 ? 2 : 1) : 0;
      hs.cplx = cplx ? 1 : 0;
      hs.arena_cap = arena_cap;
      hs.scratch_cap = scratch_cap;
      hs.arena = (double*)ex->dmalloc(elem_bytes() * arena_cap);   // floats when f32, pairs when complex
      hs.scratch = (double*)ex->dmalloc(sizeof(double) * scratch_cap);
      hs.sites = (catn::SiteDesc*)ex->dmalloc(sizeof(catn::SiteDesc) * slot_cap);
      hs.results = (int32_t*)ex->dmalloc(sizeof(int32_t) * res_cap);
      hs.dresults = (double*)ex->dmalloc(sizeof(double) * res_cap);
      hs.nsites = slot_cap;
      hs.nresults = res_cap;
      allocated = true;
      std::vector<catn::SiteDesc> zero(slot_cap);
      for (auto& z : zero) { z.off = 0; z.dl = z.d = z.dr = 0; z.live = 0; }
      ex->h2d(hs.sites, zero.data(), sizeof(catn::SiteDesc) * slot_cap);
    }
    if (nslots > slot_cap) throw catn::Error(-3, "slot table capacity exceeded");
    if (pl.nresults > res_cap) throw catn::Error(-3, "result table capacity exceeded");
    if (!staged.empty()) {
      // pack staged data behind the current arena top
      std::vector<double> buf;
      std::vector<std::pair<int, catn::SiteDesc>> descs;
      int64_t top = (hs.arena_top + 31) & ~int64_t(31);
      const int64_t base = top;
      for (auto& s : staged) {
        catn::SiteDesc d{};
        d.dl = s.dl; d.d = s.d; d.dr = s.dr; d.live = (s.dev || !s.data.empty()) ? 1 : 0;
        if (s.dev) {                // filled on the device after the upload
          d.off = top;
          if ((int)off0.size() <= s.slot) off0.resize(s.slot + 1, -1);
          off0[s.slot] = top;
          top = (top + (int64_t)s.dl * s.d * s.dr + 31) & ~int64_t(31);
        } else if (!s.data.empty()) {
          d.off = top;
          if ((int)off0.size() <= s.slot) off0.resize(s.slot + 1, -1);
          off0[s.slot] = top;
          buf.resize((size_t)(top - base) * ed(), 0.0);
          buf.insert(buf.end(), s.data.begin(), s.data.end());
          top = (top + (int64_t)s.data.size() / ed() + 31) & ~int64_t(31);
        }
        descs.push_back({s.slot, d});
      }
      if (top > arena_cap) throw catn::Error(-4, "arena capacity exceeded while uploading sites");
      if (!buf.empty()) {
        if (f32) {
          std::vector<float> fb(buf.begin(), buf.end());
          ex->h2d((float*)hs.arena + base * ed(), fb.data(), sizeof(float) * fb.size());
        } else {
          ex->h2d(hs.arena + base * ed(), buf.data(), sizeof(double) * buf.size());
        }
      }
      // staged slots are always the contiguous id range [first, nslots)
      const int first = descs.front().first;
      std::vector<catn::SiteDesc> block(descs.size());
      for (size_t k = 0; k < descs.size(); ++k) {
        if (descs[k].first != first + (int)k) throw catn::Error(-3, "staged slots not contiguous");
        block[k] = descs[k].second;
      }
      ex->h2d(hs.sites + first, block.data(), sizeof(catn::SiteDesc) * block.size());
      hs.arena_top = top;
      staged.clear();
    }
  }

  void pull_state() { arena_hwm = std::max<int64_t>(arena_hwm, hs.arena_top); }

  void check_status() {
    if (hs.status != catn::STATUS_OK) {
      static const char* names[] = {"ok", "arena full", "scratch full", "shape mismatch",
                                    "SVD did not converge", "unsupported (randomized SVD branch "
                                    "or an op without a reverse step)", "tape full",
                                    "reverse sweep could not replay a forward factorisation"};
      const int s = hs.status;
      std::string msg = "device op failed: ";
      msg += (s >= 0 && s <= 7) ? names[s] : "unknown";
      msg += " at op " + std::to_string(hs.fail_op);
      throw catn::Error(-10 - s, msg);
    }
  }

  void reset_acc() {
    hs.lnZ = 0.0; hs.err = 0.0; hs.psi = 1.0; hs.sign = 1.0; hs.lntail = 0.0;
    hs.psi_i = 0.0; hs.sign_i = 0.0;
  }

  int read_result(int idx) {
    int32_t v = 0;
    ctx->ex->d2h(&v, hs.results + idx, sizeof(int32_t));
    return v;
  }

  catn::SiteDesc read_desc(int slot) {
    catn::SiteDesc d{};
    ctx->ex->d2h(&d, hs.sites + slot, sizeof(d));
    return d;
  }
};

namespace catn {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// run the pending programs of several networks in one launch
inline double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// share[k] >= 0: network k runs the program planned for network share[k]
// (equal planner states give equal programs), so it is uploaded once
inline void run_nets(Executor* ex, const std::vector<catn_net*>& nets,
                     const std::vector<int>* share = nullptr) {
  const double t0 = wall();
  std::vector<InstState*> st;
  std::vector<const std::vector<Op>*> pr;
  for (size_t k = 0; k < nets.size(); ++k) {
    catn_net* n = nets[k];
    n->ensure_device();
    st.push_back(&n->hs);
    const int l = share ? (*share)[k] : -1;
    pr.push_back(l >= 0 ? &nets[l]->pl.prog : &n->pl.prog);
  }
  ex->run(st, pr);
  for (catn_net* n : nets) {
    n->pull_state();
    n->pl.prog.clear();
    n->check_status();
  }
  ex->run_seconds += wall() - t0;
}

inline Executor* exec_of(const catn_net* n) {
  if (!n->ctx) throw Error(-5, "the network's context has been destroyed");
  return n->ctx->ex.get();
}

inline void run_one(catn_net* n) { run_nets(exec_of(n), {n}); }

// host threads parallel_for may use on the calling thread (0: up to 16); the
// concurrent groups of a batch split them
inline thread_local int g_plan_threads = 0;

// f(i) for i in [0, n) on up to 16 host threads (planners of different
// networks share no state); the first exception is rethrown.
template <class F>
void parallel_for(int n, F f) {
  int T = (int)std::thread::hardware_concurrency();
  T = std::max(1, std::min(T, g_plan_threads > 0 ? g_plan_threads : 16));
  if (n < 16 || T == 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  T = std::min(T, n);
  std::atomic<int> next(0);
  std::exception_ptr err;
  std::mutex mu;
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&] {
      for (int i; (i = next++) < n;) {
        try {
          f(i);
        } catch (...) {
          std::lock_guard<std::mutex> lk(mu);
          if (!err) err = std::current_exception();
        }
      }
    });
  for (auto& x : th) x.join();
  if (err) std::rethrow_exception(err);
}

// ising_network site construction (src/ising.jl:20-227). The structure (per
// node: one site per incident edge in edge-list order, its kind and shape) is
// built here; the values come from ising_site_entries, on the host or, with
// device_values, by the device builder (catn_ising_network_batch): the sites
// are then allocated in the arena and filled on the device.
inline void build_ising(catn_net* net, int n, int nedges, const int32_t* edges, const double* w,
                        const double* h, double beta, int pos1, int val1, int pos2, int val2,
                        bool device_values = false) {
  if (net->cplx) throw Error(-2, "ising_network: the Ising site tensors are real (eltype must be 0)");
  std::vector<int> degree(n + 1, 0);
  for (int e = 0; e < nedges; ++e) {
    const int i = edges[2 * e], j = edges[2 * e + 1];
    if (i < 1 || j < 1 || i > n || j > n) throw Error(-2, "ising_network: edge endpoint out of range");
    degree[i]++;
    degree[j]++;
  }
  for (int k = 1; k <= n; ++k) net->pl.nodes[k] = PNode{};
  auto add_site = [&](int id, int dl, int d, int dr, int kind, int e, int side, const double* v) {
    const int slot = device_values ? net->new_device_slot(dl, d, dr) : net->new_slot(dl, d, dr, v);
    net->isites.push_back({slot, e, side, id, kind});
    return slot;
  };
  for (int e = 0; e < nedges; ++e) {
    const int i = edges[2 * e], j = edges[2 * e + 1];
    double si[2], sj[2];
    ising_spins(i, pos1, val1, pos2, val2, si);
    ising_spins(j, pos1, val1, pos2, val2, sj);
    for (int side = 0; side < 2; ++side) {
      const int id = side == 0 ? i : j;
      const int other = side == 0 ? j : i;
      PNode& nd = net->pl.nodes[id];
      int dl, d = 2, dr, kind;
      if (degree[id] == 1) { dl = 1; dr = 1; kind = 0; }                           // leaf
      else if (nd.nb.empty()) { dl = 1; dr = 2; kind = 1; }                       // first
      else if ((int)nd.nb.size() == degree[id] - 1) { dl = 2; dr = 1; kind = 2; } // last
      else { dl = 2; dr = 2; kind = 3; }                                          // interior
      double site[8];
      if (!device_values)
        ising_site_entries(kind, side, side == 0 ? si : sj, side == 0 ? sj : si, w[e], h[id - 1],
                           beta, site);
      const int slot = add_site(id, dl, d, dr, kind, e, side, site);
      nd.slot.push_back(slot);
      nd.phys.push_back(d);
      nd.nb.push_back(other);
    }
  }
  for (int k = 1; k <= n; ++k) {
    if (degree[k] == 0) {
      double v = 0.0;
      if (!device_values) ising_site_entries(4, 0, nullptr, nullptr, 0.0, h[k - 1], beta, &v);
      const int slot = add_site(k, 1, 1, 1, 4, -1, 0, &v);
      net->pl.nodes[k].slot.push_back(slot);
      net->pl.nodes[k].phys.push_back(1);
    }
  }
  int iso = 0;
  for (auto& kv : net->pl.nodes) {
    kv.second.cano = kv.second.order();
    if (kv.second.nb.empty()) ++iso;
  }
  net->num_isolated = iso;
  net->n = n;
  net->beta = beta;
  net->is_ising = true;
  net->iedges.assign(edges, edges + 2 * nedges);
  net->iw.assign(w, w + nedges);
  net->ih.assign(h, h + n);
  net->ipos1 = pos1; net->ival1 = val1; net->ipos2 = pos2; net->ival2 = val2;
}

// Chain rule of the reverse sweep through the ising_network site formulas
// (src/ising.jl:20-227): site entries are products of B = exp(beta w s_i s_j')
// (side i of an edge; side j carries the identity) and the field vector
// f = exp(beta h s) (leaf and first sites only); adj holds d lnZ / d site.
// dbeta = (sum_e w_e dw_e + sum_i h_i dh_i) / beta.
inline void ising_chain_rule(const catn_net* net, const std::vector<double>& adj0, int64_t base,
                             double* dw, double* dh, double* dbeta) {
  const int n = net->n, ne = (int)net->iw.size();
  const double beta = net->beta;
  std::vector<double> gw(ne, 0.0), gh(n, 0.0);
  auto spins = [&](int id, double s[2]) {
    s[0] = 1.0; s[1] = -1.0;
    if (id == net->ipos1) s[net->ival1 - 1] = 0.0;
    else if (id == net->ipos2) s[net->ival2 - 1] = 0.0;
  };
  for (const auto& is : net->isites) {
    const double* a = adj0.data() + (net->off0[is.slot] - base);
    const int id = is.node;
    const double hv = net->ih[id - 1];
    if (is.kind == 4) {
      gh[id - 1] += a[0] * beta * (std::exp(beta * hv) - std::exp(-beta * hv));
      continue;
    }
    const int i = net->iedges[2 * is.edge], j = net->iedges[2 * is.edge + 1];
    double si[2], sj[2];
    spins(i, si);
    spins(j, sj);
    const double we = net->iw[is.edge];
    double mat[2][2], dmat[2][2];   // mat and d mat / d w_e
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2; ++y) {
        if (is.side == 0) {
          mat[x][y] = std::exp(beta * we * (si[x] * sj[y]));
          dmat[x][y] = beta * si[x] * sj[y] * mat[x][y];
        } else {
          mat[x][y] = x == y ? 1.0 : 0.0;
          dmat[x][y] = 0.0;
        }
      }
    double s[2];
    spins(id, s);
    const double f[2] = {std::exp(beta * hv * s[0]), std::exp(beta * hv * s[1])};
    const double df[2] = {beta * s[0] * f[0], beta * s[1] * f[1]};
    double ddw = 0.0, ddh = 0.0;
    switch (is.kind) {
      case 0:   // site[c] = sum_x f[x] mat[x][c]
        for (int c = 0; c < 2; ++c)
          for (int x = 0; x < 2; ++x) {
            ddw += a[c] * f[x] * dmat[x][c];
            ddh += a[c] * df[x] * mat[x][c];
          }
        break;
      case 1:   // site[x + 2y] = f[y] mat[y][x]
        for (int x = 0; x < 2; ++x)
          for (int y = 0; y < 2; ++y) {
            ddw += a[x + 2 * y] * f[y] * dmat[y][x];
            ddh += a[x + 2 * y] * df[y] * mat[y][x];
          }
        break;
      case 2:   // site[x + 2y] = mat[x][y]
        for (int x = 0; x < 2; ++x)
          for (int y = 0; y < 2; ++y) ddw += a[x + 2 * y] * dmat[x][y];
        break;
      default:  // site[x + 2 sv + 4 x] = mat[x][sv]
        for (int x = 0; x < 2; ++x)
          for (int sv = 0; sv < 2; ++sv) ddw += a[x + 2 * sv + 4 * x] * dmat[x][sv];
        break;
    }
    gw[is.edge] += ddw;
    gh[id - 1] += ddh;
  }
  double db = 0.0;
  for (int e = 0; e < ne; ++e) db += net->iw[e] * gw[e];
  for (int k = 0; k < n; ++k) db += net->ih[k] * gh[k];
  if (dw) for (int e = 0; e < ne; ++e) dw[e] = gw[e];
  if (dh) for (int k = 0; k < n; ++k) dh[k] = gh[k];
  if (dbeta) *dbeta = beta != 0.0 ? db / beta : 0.0;
}

// raw2mps emitted as device split ops (src/mps_node.jl:31-50)
inline void add_tensor_node(catn_net* net, int id, int nd, const int64_t* dims, const double* data,
                            const std::vector<int>& nb) {
  PNode node;
  node.nb = nb;
  if (nd == 0) {
    // empty MPS (src/mps_node.jl:34)
    net->tensor_slot.push_back({-1, 0});
  } else if (nd == 1) {
    const int slot = net->new_slot(1, (int)dims[0], 1, data);
    net->tensor_slot.push_back({slot, dims[0]});
    node.slot.push_back(slot);
    node.phys.push_back((int)dims[0]);
  } else {
    int64_t total = 1;
    for (int k = 0; k < nd; ++k) total *= dims[k];
    const int rem = net->new_slot(1, (int)total, 1, data);
    net->tensor_slot.push_back({rem, total});
    for (int k = 0; k < nd - 1; ++k) {
      const int s = net->new_slot(0, 0, 0, nullptr);
      net->pl.emit(OP_RAW_SPLIT, rem, s, (int)dims[k], net->pl.p.chi, 0, 0, net->pl.p.cutoff);
      node.slot.push_back(s);
      node.phys.push_back((int)dims[k]);
    }
    node.slot.push_back(rem);
    node.phys.push_back((int)dims[nd - 1]);
  }
  node.cano = node.order();
  net->pl.nodes[id] = node;
}

inline void contract_nets(Executor* ex, const std::vector<catn_net*>& nets) {
  for (catn_net* n : nets) {
    if (n->consumed) throw Error(-5, "contraction!: network already contracted");
    if (exec_of(n) != (ex->parent ? ex->parent : ex))
      throw Error(-2, "contraction: networks of another context");
    n->ensure_device();
    n->reset_acc();
  }
  // Planning is a deterministic function of the planner state and of the cut
  // results applied between segments. Networks whose states are equal before
  // the start (pinned networks of one ising instance differ only in site
  // values) follow one owner: only the owner starts and plans, the followers
  // run its uploaded program, a follower takes a copy of the owner's planner
  // only when its cut result parts from the owner's, and at the end every
  // follower shares its owner's final planner (read-only, copied on a later
  // mutation).
  std::vector<catn_net*> active = nets;
  const int n0 = (int)active.size();
  std::vector<int> att(n0, -1);   // index in `active` of the owner followed, or -1
  {
    const double t0 = wall();
    std::vector<uint64_t> fp(n0);
    parallel_for(n0, [&](int k) { fp[k] = active[k]->pl.fingerprint(); });
    std::unordered_map<uint64_t, int> first;
    for (int k = 0; k < n0; ++k) {
      auto it = first.find(fp[k]);
      if (it == first.end()) first.emplace(fp[k], k);
      else att[k] = it->second;
    }
    parallel_for(n0, [&](int k) {   // equal fingerprints must be equal states
      if (att[k] >= 0 && !active[k]->pl.same_state(active[att[k]]->pl)) att[k] = -1;
    });
    std::vector<int> owners;
    for (int k = 0; k < n0; ++k)
      if (att[k] < 0) owners.push_back(k);
    parallel_for((int)owners.size(), [&](int q) { active[owners[q]]->pl.start(); });
    ex->plan_seconds += wall() - t0;
  }
  while (!active.empty()) {
    const int na = (int)active.size();
    std::vector<int> cut(na);
    double t0 = wall();
    // work lists, so that few owners are planned without starting threads
    std::vector<int> owners;
    for (int k = 0; k < na; ++k)
      if (att[k] < 0) owners.push_back(k);
    parallel_for((int)owners.size(), [&](int q) {
      const int k = owners[q];
      cut[k] = active[k]->pl.step() == Planner::NEED_CUT ? 1 : 0;
    });
    for (int k = 0; k < na; ++k)
      if (att[k] >= 0) cut[k] = cut[att[k]];
    ex->plan_seconds += wall() - t0;
    if (!ex->cut_reverse_enabled())
      for (int k = 0; k < na; ++k)
        if (cut[k] && active[k]->hs.record)
          throw Error(-15, "gradient across a Dmax cut (cut_bondim_opt!) is disabled on this "
                           "device; use the pinning route (magnetization / correlation) or a "
                           "Dmax that the network never exceeds");
    run_nets(ex, active, &att);
    t0 = wall();
    // regroup by cut result: a follower whose kept count differs from its
    // owner's joins (or founds) the subgroup of its count; a founder copies
    // the owner's planner before the owner applies its own count
    std::vector<int> own(na, -1);   // new owner index (in `active`) per network
    std::vector<int> copy_from(na, -1);
    {
      std::vector<std::vector<std::pair<int, int>>> sub(na);   // per owner: (myd, founder)
      for (int k = 0; k < na; ++k) {
        if (att[k] < 0) { own[k] = k; continue; }
        const int l = att[k];
        if (!cut[k] || active[k]->hs.last_cut == active[l]->hs.last_cut) { own[k] = l; continue; }
        int f = -1;
        for (const auto& pr : sub[l])
          if (pr.first == active[k]->hs.last_cut) f = pr.second;
        if (f < 0) {
          sub[l].push_back({active[k]->hs.last_cut, k});
          copy_from[k] = l;
          f = k;
        }
        own[k] = f;
      }
    }
    // the new founders' planners, and those finished followers need (edge log,
    // maxdim); programs were cleared by run_nets, so nothing large is copied
    std::vector<int> copies, applies;
    for (int k = 0; k < na; ++k) {
      if (copy_from[k] >= 0) copies.push_back(k);
      if (cut[k] && own[k] == k) applies.push_back(k);
    }
    parallel_for((int)copies.size(), [&](int q) {
      const int k = copies[q];
      active[k]->pl = active[copy_from[k]]->pl;
    });
    // finished: an owner's final planner becomes shared and read-only; its
    // finished followers refer to it
    for (int k = 0; k < na; ++k)
      if (!cut[k] && att[k] < 0) active[k]->finish_plan();
    for (int k = 0; k < na; ++k)
      if (!cut[k] && att[k] >= 0) active[k]->done_pl = active[att[k]]->done_pl;
    parallel_for((int)applies.size(), [&](int q) {
      const int k = applies[q];
      active[k]->pl.apply_cut(active[k]->hs.last_cut);
    });
    std::vector<catn_net*> still;
    std::vector<int> pos(na, -1), next_att;
    for (int k = 0; k < na; ++k) {
      if (cut[k]) {
        pos[k] = (int)still.size();
        still.push_back(active[k]);
      } else {
        active[k]->consumed = true;
      }
    }
    for (int k = 0; k < na; ++k)
      if (cut[k]) next_att.push_back(own[k] == k ? -1 : pos[own[k]]);
    ex->plan_seconds += wall() - t0;
    active.swap(still);
    att.swap(next_att);
  }
}

// Forward contraction of a batch of distinct networks in concurrent groups.
// contract_nets advances a batch one planning segment at a time: every launch
// waits for the slowest network of the batch before any network starts its
// next segment, so a batch whose networks differ in size takes the sum over
// segments of the slowest segment. Here the networks are dealt into up to
// kGroupStreams groups, each contracted by its own host thread on a sibling
// executor (its own stream), so a group waits only for its own networks.
// They are dealt in decreasing order of node count, round robin, so the
// largest networks land in different groups; within a group, networks of one
// structure still share plans. A batch of fewer than 3 structures (such as
// the pinned networks of one or two instances) keeps the one-stream path,
// where all of them share plans. Returns false (nothing run) when the batch
// does not qualify or the executor has no siblings.
constexpr int kGroupStreams = 16;     // streams of a batch at most
constexpr int kGroupMinNets = 4;      // networks per group at least

inline bool contract_nets_grouped(Executor* ex, const std::vector<catn_net*>& nets) {
  const int nn = (int)nets.size();
  const int G = std::min(std::min(kGroupStreams, ex->batch_streams), nn / kGroupMinNets);
  if (G < 2 || ex->parent) return false;
  for (catn_net* n : nets) {
    if (n->consumed) throw Error(-5, "contraction!: network already contracted");
    if (exec_of(n) != ex) throw Error(-2, "contraction: networks of another context");
    if (n->hs.record) return false;
  }
  {
    // planner-state classes (networks of one structure: the pinned networks of
    // an instance, lattices of one size); a batch of few classes shares its
    // plans best on one stream
    std::vector<uint64_t> fp(nn);
    parallel_for(nn, [&](int k) { fp[k] = nets[k]->pl.fingerprint(); });
    std::unordered_map<uint64_t, int> classes;
    for (int k = 0; k < nn; ++k) classes.emplace(fp[k], k);
    if ((int)classes.size() < 3) return false;
  }
  // group 0 runs on this executor's own stream, groups 1.. on siblings: G
  // streams in all, one per hardware queue (a fifth stream would share a
  // queue and serialise two groups)
  std::vector<std::unique_ptr<Executor>> sib;
  for (int g = 1; g < G; ++g) {
    try {
      sib.emplace_back(ex->sibling());
    } catch (const Error&) {
      return false;   // no second stream on this device: the one-stream path
    }
    if (!sib.back()) return false;
    Executor* s = sib.back().get();
    s->parent = ex;
    s->threads = ex->threads;
    s->lds_override = ex->lds_override;
    s->qb = ex->qb;
    s->jac_lanes = ex->jac_lanes;
  }
  // uploads through the owning executor, before any group starts
  for (catn_net* n : nets) n->ensure_device();
  std::vector<int> order(nn);
  for (int k = 0; k < nn; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return nets[a]->pl.nodes.size() > nets[b]->pl.nodes.size();
  });
  std::vector<std::vector<catn_net*>> groups(G);
  for (int k = 0; k < nn; ++k) groups[k % G].push_back(nets[order[k]]);
  std::vector<std::exception_ptr> errs(G);
  std::vector<std::thread> th;
  for (int g = 0; g < G; ++g)
    th.emplace_back([&, g] {
      try {
        Executor* e = g == 0 ? ex : sib[g - 1].get();
        e->bind_thread();
        g_plan_threads = std::max(1, 16 / G);
        contract_nets(e, groups[g]);
      } catch (...) {
        errs[g] = std::current_exception();
      }
    });
  for (auto& t : th) t.join();
  for (auto& s : sib) {   // groups 1..: launches and times (kernel time summed over streams)
    ex->launches += s->launches;
    ex->kernel_seconds += s->kernel_seconds;
    ex->plan_seconds += s->plan_seconds;
    ex->run_seconds += s->run_seconds;
  }
  for (auto& e : errs)
    if (e) std::rethrow_exception(e);
  return true;
}

// tape records per network when the caller gives none
constexpr int kDefaultTapeCap = 1 << 16;

// contraction! with the tape on, then one reverse-sweep launch for all
inline void gradient_nets(Executor* ex, const std::vector<catn_net*>& nets, int tape_cap) {
  for (catn_net* n : nets) {
    if (n->consumed) throw Error(-5, "contraction!: network already contracted");
    if (n->cplx)
      throw Error(-2, "gradient: the reverse sweep is built for real networks; complex networks "
                      "run the plain forward (catn_contraction)");
    n->enable_record(tape_cap);
  }
  contract_nets(ex, nets);
  int lds = 0;
  std::vector<InstState*> st;
  for (catn_net* n : nets) {
    lds = std::max(lds, (int)n->hs.lds_max);
    st.push_back(&n->hs);
  }
  const double t0 = wall();
  ex->run_backward(st, lds);
  ex->run_seconds += wall() - t0;
  for (catn_net* n : nets) {
    n->check_status();
    n->have_adjoint = true;
  }
}

}  // namespace catn

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
#define CATN_TRY try {
#define CATN_CATCH                                                      \
  }                                                                     \
  catch (const catn::Error& e) { return catn::fail(e.code, e.what()); } \
  catch (const catn::PlanError& e) { return catn::fail(-2, e.what()); } \
  catch (const std::exception& e) { return catn::fail(-1, e.what()); }

extern "C" {

void catn_params_default(catn_params* p) {
  p->Dmax = 32; p->chi = 32; p->cutoff = 1e-15; p->norm_method = 1; p->select = 1;
  p->reverse = 1; p->svdopt = 1; p->swapopt = 1; p->compress = 0; p->cut_bond = 0;
  p->seed = 1; p->arena_elems = 0; p->scratch_elems = 0;
  p->rsvd_both = 7000; p->rsvd_any = 20000; p->precision = 0; p->loops = 0;
  p->braket_dmax = -1; p->eltype = 0;
}

const char* catn_last_error(void) { return catn::g_last_error.c_str(); }

int catn_create(catn_ctx** ctx, int device) {
  CATN_TRY
  auto c = new catn_ctx;
  c->ex.reset(catn::make_executor(device));
  *ctx = c;
  return 0;
  CATN_CATCH
}

int catn_destroy(catn_ctx* ctx) {
  for (catn_net* n : ctx->nets) {
    n->release();
    n->ctx = nullptr;
  }
  ctx->nets.clear();
  delete ctx;
  return 0;
}

const char* catn_backend(catn_ctx* ctx) { return ctx->ex->name(); }

static void apply_params(catn_net* net, const catn_params* p) {
  catn_params d;
  catn_params_default(&d);
  if (!p) p = &d;
  catn::Params& q = net->pl.p;
  q.Dmax = p->Dmax; q.chi = p->chi; q.cutoff = p->cutoff; q.norm_method = p->norm_method;
  q.select = p->select; q.reverse = p->reverse != 0; q.svdopt = p->svdopt != 0;
  q.swapopt = p->swapopt != 0; q.compress = p->compress != 0; q.cut_bond = p->cut_bond != 0;
  q.seed = p->seed;
  q.rsvd_both = p->rsvd_both;
  q.rsvd_any = p->rsvd_any;
  q.loops = p->loops != 0;
  if (p->eltype != 0 && p->eltype != 1) throw catn::Error(-2, "eltype must be 0 (real) or 1 (complex)");
  if (p->precision < 0 || p->precision > 2)
    throw catn::Error(-2, "precision must be 0 (FP64), 1 (FP32, split-BF16 3-pass) or 2 (FP32, 6-pass)");
  net->cplx = p->eltype == 1;
  net->f32 = p->precision != 0;
  // ComplexF32 storage (eltype 1, precision 1 or 2): FP32 (re, im) pairs in
  // the arena, every product in FP64, so the storage's noise floor (level 2)
  // applies whichever precision level was asked for
  net->bf16_terms = (p->precision == 2 || net->cplx) ? 3 : 2;
  net->arena_cap = p->arena_elems;
  net->scratch_cap = p->scratch_elems;
}

int catn_ising_network(catn_ctx* ctx, int32_t n, int32_t nedges, const int32_t* edges,
                       const double* w, const double* h, double beta, int32_t pos1,
                       int32_t val1, int32_t pos2, int32_t val2, const catn_params* p,
                       catn_net** out) {
  CATN_TRY
  std::unique_ptr<catn_net> net(new catn_net);
  net->attach(ctx);
  apply_params(net.get(), p);
  catn::build_ising(net.get(), n, nedges, edges, w, h, beta, pos1, val1, pos2, val2);
  *out = net.release();
  return 0;
  CATN_CATCH
}

int catn_ising_network_batch(catn_ctx* ctx, int32_t ninst, const int32_t* inst_n,
                             const int32_t* inst_nedges, const int32_t* edges, const double* w,
                             const double* h, int32_t nnets, const int32_t* net_inst,
                             const double* net_beta, const int32_t* net_pins,
                             const catn_params* p, catn_net** out) {
  CATN_TRY
  if (ninst <= 0 || nnets <= 0) throw catn::Error(-2, "ising_network_batch: empty batch");
  std::vector<int64_t> eoff(ninst + 1, 0), foff(ninst + 1, 0);
  for (int q = 0; q < ninst; ++q) {
    if (inst_n[q] <= 0 || inst_nedges[q] < 0)
      throw catn::Error(-2, "ising_network_batch: instance sizes must be positive");
    eoff[q + 1] = eoff[q] + inst_nedges[q];
    foff[q + 1] = foff[q] + inst_n[q];
  }
  std::vector<std::unique_ptr<catn_net>> nets(nnets);
  for (int k = 0; k < nnets; ++k) {
    const int q = net_inst[k];
    if (q < 0 || q >= ninst) throw catn::Error(-2, "ising_network_batch: instance index out of range");
    nets[k].reset(new catn_net);
    nets[k]->attach(ctx);
    apply_params(nets[k].get(), p);
    const int32_t* pin = net_pins ? net_pins + 4 * k : nullptr;
    catn::build_ising(nets[k].get(), inst_n[q], inst_nedges[q], edges + 2 * eoff[q], w + eoff[q],
                      h + foff[q], net_beta[k], pin ? pin[0] : 0, pin ? pin[1] : 0,
                      pin ? pin[2] : 0, pin ? pin[3] : 0, true);
  }
  for (auto& n : nets) n->ensure_device();
  // site records per instance from its first network; a network whose arena
  // layout differs (it cannot for equal structure and parameters) gets its own
  catn::IsingBuild b;
  b.edges = edges; b.w = w; b.h = h;
  b.nedges = eoff[ninst]; b.nfields = foff[ninst];
  std::vector<int> rec_of(ninst, -1);   // instance -> record index
  auto add_inst = [&](const catn_net* n, int q) {
    catn::IsingInstRec ir;
    ir.site_begin = (int32_t)b.sites.size();
    for (const auto& is : n->isites)
      b.sites.push_back({n->off0[is.slot], is.kind, is.edge, is.side, is.node});
    ir.site_end = (int32_t)b.sites.size();
    ir.edge_off = eoff[q];
    ir.field_off = foff[q];
    b.max_sites = std::max(b.max_sites, ir.site_end - ir.site_begin);
    b.insts.push_back(ir);
    return (int)b.insts.size() - 1;
  };
  for (int k = 0; k < nnets; ++k) {
    const catn_net* n = nets[k].get();
    const int q = net_inst[k];
    int r = rec_of[q];
    if (r < 0) {
      r = rec_of[q] = add_inst(n, q);
    } else {
      const catn::IsingInstRec& ir = b.insts[r];
      bool same = (int)n->isites.size() == ir.site_end - ir.site_begin;
      for (size_t t = 0; same && t < n->isites.size(); ++t)
        same = n->off0[n->isites[t].slot] == b.sites[ir.site_begin + t].off;
      if (!same) r = add_inst(n, q);
    }
    const int32_t* pin = net_pins ? net_pins + 4 * k : nullptr;
    catn::IsingNetRec nr;
    nr.arena = n->hs.arena;
    nr.beta = net_beta[k];
    nr.inst = r;
    nr.f32 = n->f32 ? 1 : 0;
    nr.pos1 = pin ? pin[0] : 0; nr.val1 = pin ? pin[1] : 0;
    nr.pos2 = pin ? pin[2] : 0; nr.val2 = pin ? pin[3] : 0;
    b.nets.push_back(nr);
  }
  ctx->ex->build_ising(b);
  for (int k = 0; k < nnets; ++k) out[k] = nets[k].release();
  return 0;
  CATN_CATCH
}

int catn_tensor_network(catn_ctx* ctx, int32_t ntensors, const int32_t* ndims,
                        const int64_t* dims, const double* data, const int64_t* labels,
                        const catn_params* p, catn_net** out) {
  CATN_TRY
  std::unique_ptr<catn_net> net(new catn_net);
  net->attach(ctx);
  apply_params(net.get(), p);
  // label -> occurrences (src/tensor_network.jl:59-102)
  std::map<int64_t, std::vector<std::pair<int, int>>> occ;
  std::vector<int64_t> first_seen;
  size_t li = 0;
  for (int t = 0; t < ntensors; ++t) {
    std::set<int64_t> seen;
    for (int k = 0; k < ndims[t]; ++k) {
      const int64_t l = labels[li + k];
      if (seen.count(l))
        throw catn::Error(-2, "Self-loop detected: label " + std::to_string(l) +
                                  " appears twice in tensor " + std::to_string(t + 1) + "'s index list");
      seen.insert(l);
      if (!occ.count(l)) first_seen.push_back(l);
      occ[l].push_back({t + 1, k});
    }
    li += ndims[t];
  }
  for (auto& kv : occ)
    if (kv.second.size() > 2)
      throw catn::Error(-2, "Label " + std::to_string(kv.first) + " appears in " +
                                std::to_string(kv.second.size()) + " tensors (max 2 allowed)");
  std::vector<std::vector<int>> nbv(ntensors);
  for (int t = 0; t < ntensors; ++t) nbv[t].assign(ndims[t], 0);
  int sentinel = 0;
  for (int64_t l : first_seen) {
    auto& o = occ[l];
    if (o.size() == 2) {
      nbv[o[0].first - 1][o[0].second] = o[1].first;
      nbv[o[1].first - 1][o[1].second] = o[0].first;
    } else {
      nbv[o[0].first - 1][o[0].second] = --sentinel;
    }
  }
  size_t di = 0, dd = 0;
  int iso = 0;
  for (int t = 0; t < ntensors; ++t) {
    int64_t total = 1;
    for (int k = 0; k < ndims[t]; ++k) total *= dims[di + k];
    catn::add_tensor_node(net.get(), t + 1, ndims[t], dims + di, data + dd, nbv[t]);
    bool all_neg = true;
    for (int v : nbv[t]) all_neg = all_neg && v < 0;
    if (all_neg) ++iso;
    di += ndims[t];
    dd += (size_t)total * net->ed();
  }
  net->num_isolated = iso;
  net->n = ntensors;
  net->beta = 1.0;
  if (!net->pl.prog.empty()) {
    // a real network records its raw2mps splits now, for a later reverse sweep
    if (!net->cplx) {
      net->ensure_device();
      int nsplit = 0;
      for (const catn::Op& op : net->pl.prog) nsplit += op.code == catn::OP_RAW_SPLIT;
      net->alloc_tape(nsplit);
      net->hs.record = 1;
    }
    catn::run_one(net.get());
    net->hs.record = 0;
    net->build_tape = net->hs.tape_len;
  }
  net->tensors_taped = !net->cplx;
  *out = net.release();
  return 0;
  CATN_CATCH
}

// braket_network(tensors, ixs; ...) (src/braket.jl:369-496) for real tensors:
// one BraKetNode per tensor (braket_node, :75-185, as device split ops), the
// paired edges in the order their labels first appear.
int catn_braket_network(catn_ctx* ctx, int32_t ntensors, const int32_t* ndims,
                        const int64_t* dims, const double* data, const int64_t* labels,
                        const catn_params* p, catn_net** out) {
  CATN_TRY
  std::unique_ptr<catn_net> net(new catn_net);
  net->attach(ctx);
  apply_params(net.get(), p);
  {   // the double layer's merged-bond Dmax is opt-in (braket_dmax); `Dmax` is not used
    catn_params d;
    catn_params_default(&d);
    net->pl.p.Dmax = (p ? p : &d)->braket_dmax;
  }
  std::map<int64_t, std::vector<int>> owners;   // label -> tensor ids (1-based)
  std::vector<int64_t> first_seen;
  size_t li = 0;
  for (int t = 0; t < ntensors; ++t) {
    for (int k = 0; k < ndims[t]; ++k) {
      const int64_t l = labels[li + k];
      if (!owners.count(l)) first_seen.push_back(l);
      owners[l].push_back(t + 1);
    }
    li += ndims[t];
  }
  for (auto& kv : owners)
    if (kv.second.size() > 2)
      throw catn::Error(-2, "braket_network: label " + std::to_string(kv.first) + " appears " +
                                std::to_string(kv.second.size()) + " times (expected 1 or 2)");
  // acyclic virtual-bond graph only (:402-420), unless loops are enabled;
  // two bonds between one pair of tensors are rejected either way
  std::set<std::pair<int, int>> pairs;
  std::vector<int> parent(ntensors + 1);
  for (int t = 0; t <= ntensors; ++t) parent[t] = t;
  auto root = [&](int x) {
    while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; }
    return x;
  };
  std::vector<std::pair<int, int>> edges;
  for (int64_t l : first_seen) {
    const auto& o = owners[l];
    if (o.size() != 2) continue;
    const std::pair<int, int> pr = {std::min(o[0], o[1]), std::max(o[0], o[1])};
    if (pairs.count(pr))
      throw catn::Error(-2, "braket_network: tensors " + std::to_string(pr.first) + " and " +
                                std::to_string(pr.second) + " share more than one bond");
    pairs.insert(pr);
    const int a = root(o[0]), b = root(o[1]);
    if (a == b && !net->pl.p.loops)
      throw catn::Error(-2, "braket_network: virtual-bond graph contains a cycle - cyclic graphs "
                            "are not supported (loops=1 enables them)");
    parent[a] = b;
    edges.push_back({o[0], o[1]});
  }
  catn::Planner& pl = net->pl;
  size_t di = 0, dd = 0;
  li = 0;
  for (int t = 0; t < ntensors; ++t) {
    const int nd = ndims[t];
    int phys_pos = -1, nphys = 0;
    std::vector<int> kn, virt;
    for (int k = 0; k < nd; ++k) {
      const auto& o = owners[labels[li + k]];
      if (o.size() == 1) { phys_pos = k; ++nphys; }
      else { virt.push_back(k); kn.push_back(o[0] == t + 1 ? o[1] : o[0]); }
    }
    if (nphys != 1)
      throw catn::Error(-2, "braket_network: tensor " + std::to_string(t + 1) + " has " +
                                std::to_string(nphys) + " physical (open) labels, expected 1");
    const int deg = nd - 1;
    if (deg < 1) throw catn::Error(-2, "braket_node: degree-0 site (no virtual legs) is not supported.");
    const int64_t* tdims = dims + di;
    int64_t total = 1;
    for (int k = 0; k < nd; ++k) total *= tdims[k];
    // column-major permutation: out axes = perm, out[i_out] = in[i_in]
    // (complex tensors: interleaved pairs; conj negates the imaginary parts)
    const int ed = net->ed();
    auto permuted = [&](const std::vector<int>& perm, bool conj) {
      std::vector<double> outv((size_t)total * ed);
      std::vector<int64_t> stride(nd, 1);
      for (int k = 1; k < nd; ++k) stride[k] = stride[k - 1] * tdims[k - 1];
      std::vector<int64_t> idx(nd, 0);
      for (int64_t q = 0; q < total; ++q) {
        int64_t src = 0;
        for (int k = 0; k < nd; ++k) src += idx[k] * stride[perm[k]];
        outv[(size_t)q * ed] = data[dd + (size_t)src * ed];
        if (ed == 2) outv[(size_t)q * 2 + 1] = (conj ? -1.0 : 1.0) * data[dd + (size_t)src * 2 + 1];
        for (int k = 0; k < nd; ++k) {   // next output multi-index
          if (++idx[k] < tdims[perm[k]]) break;
          idx[k] = 0;
        }
      }
      return outv;
    };
    std::vector<int> pk = virt, pb;
    pk.push_back(phys_pos);
    pb.push_back(phys_pos);
    pb.insert(pb.end(), virt.begin(), virt.end());
    const std::vector<double> tk = permuted(pk, false);   // (D1, ..., D_deg, d)
    const std::vector<double> tb = permuted(pb, true);    // conj, (d, D1, ..., D_deg)
    const int d = (int)tdims[phys_pos];
    catn::PNode node;
    // ket chain (:116-130): splits over D1..D_{deg-1}; the remainder becomes the
    // last ket site (chi, D_deg, d)
    const int krem = net->new_slot(1, (int)total, 1, tk.data());
    for (int k = 0; k < deg - 1; ++k) {
      const int s2 = net->new_slot(0, 0, 0, nullptr);
      pl.emit(catn::OP_RAW_SPLIT, krem, s2, (int)tdims[virt[k]], pl.p.chi, 0, 0, pl.p.cutoff);
      node.slot.push_back(s2);
    }
    pl.emit(catn::OP_RESHAPE, krem, 0, 0, (int)tdims[virt[deg - 1]]);
    node.slot.push_back(krem);
    // bra chain (:155-172): the first bra site's left bond is d
    const int brem = net->new_slot(d, (int)(total / d), 1, tb.data());
    for (int k = 0; k < deg - 1; ++k) {
      const int s2 = net->new_slot(0, 0, 0, nullptr);
      pl.emit(catn::OP_RAW_SPLIT, brem, s2, (int)tdims[virt[k]], pl.p.chi, 0, 0, pl.p.cutoff);
      node.slot.push_back(s2);
    }
    node.slot.push_back(brem);
    for (int side = 0; side < 2; ++side)
      for (int k = 0; k < deg; ++k) {
        node.nb.push_back(kn[k]);
        node.phys.push_back((int)tdims[virt[k]]);
        node.layer.push_back(side == 0 ? 1 : 0);
      }
    node.cano = 1;
    pl.nodes[t + 1] = node;
    pl.left_canonical(pl.nodes[t + 1]);   // :183
    di += nd;
    dd += (size_t)total * ed;
    li += nd;
  }
  pl.braket = true;
  pl.bk_edges = edges;
  pl.maxdim_intermediate = 0;
  net->n = ntensors;
  net->beta = 1.0;
  if (!pl.prog.empty()) catn::run_one(net.get());
  *out = net.release();
  return 0;
  CATN_CATCH
}

int catn_net_create(catn_ctx* ctx, const catn_params* p, catn_net** out) {
  CATN_TRY
  std::unique_ptr<catn_net> net(new catn_net);
  net->attach(ctx);
  apply_params(net.get(), p);
  *out = net.release();
  return 0;
  CATN_CATCH
}

int catn_net_add_node(catn_net* net, int32_t id, int32_t nsites, const int32_t* dims,
                      const double* data, const int32_t* neighbor, int32_t cano) {
  CATN_TRY
  catn::PNode node;
  size_t off = 0;
  for (int k = 0; k < nsites; ++k) {
    const int dl = dims[3 * k], d = dims[3 * k + 1], dr = dims[3 * k + 2];
    node.slot.push_back(net->new_slot(dl, d, dr, data + off));
    node.phys.push_back(d);
    node.nb.push_back(neighbor[k]);
    off += (size_t)dl * d * dr * net->ed();
  }
  node.cano = cano;
  net->pl.nodes[id] = node;
  return 0;
  CATN_CATCH
}

int catn_net_add_tensor_node(catn_net* net, int32_t id, int32_t ndim, const int64_t* dims,
                             const double* data, const int32_t* neighbor) {
  CATN_TRY
  std::vector<int> nb(neighbor, neighbor + ndim);
  catn::add_tensor_node(net, id, ndim, dims, data, nb);
  if (!net->pl.prog.empty()) catn::run_one(net);
  return 0;
  CATN_CATCH
}

int catn_net_set_fault(catn_net* net, int32_t flags) {
  CATN_TRY
  if (flags & ~1) throw catn::Error(-2, "set_fault: unknown flag bits");
  if (net->cplx && flags) throw catn::Error(-2, "set_fault: the SVD fallback is built for real networks");
  net->ensure_device();
  net->hs.fault = flags;
  return 0;
  CATN_CATCH
}

int catn_net_prepare_gradient(catn_net* net, int32_t tape_cap) {
  CATN_TRY
  if (tape_cap < 0) throw catn::Error(-2, "prepare_gradient: tape_cap must be >= 0");
  if (net->cplx)
    throw catn::Error(-2, "prepare_gradient: the reverse sweep is built for real networks");
  net->alloc_record(tape_cap > 0 ? tape_cap : catn::kDefaultTapeCap);
  return 0;
  CATN_CATCH
}

int catn_net_upload(catn_net* net) {
  CATN_TRY
  net->ensure_device();
  return 0;
  CATN_CATCH
}

int catn_net_free(catn_net* net) {
  delete net;
  return 0;
}

int catn_contraction(catn_net* net, double* lnZ, double* error, double* psi) {
  CATN_TRY
  catn::contract_nets(catn::exec_of(net), {net});
  if (lnZ) *lnZ = net->hs.lnZ + net->hs.lntail;
  if (error) *error = net->hs.err;
  if (psi) {
    psi[0] = net->hs.psi;
    if (net->cplx) psi[1] = net->hs.psi_i;
  }
  return 0;
  CATN_CATCH
}

int catn_contraction_batch(catn_ctx* ctx, int32_t count, catn_net* const* nets, double* lnZ,
                           double* error, double* psi) {
  CATN_TRY
  std::vector<catn_net*> v(nets, nets + count);
  for (int k = 1; k < count; ++k)
    if (v[k]->cplx != v[0]->cplx)
      throw catn::Error(-2, "contraction_batch: real and complex networks in one batch");
  if (!catn::contract_nets_grouped(ctx->ex.get(), v)) catn::contract_nets(ctx->ex.get(), v);
  const bool cx = count > 0 && v[0]->cplx;
  for (int k = 0; k < count; ++k) {
    if (lnZ) lnZ[k] = v[k]->hs.lnZ + v[k]->hs.lntail;
    if (error) error[k] = v[k]->hs.err;
    if (psi && cx) {
      psi[2 * k] = v[k]->hs.psi;
      psi[2 * k + 1] = v[k]->hs.psi_i;
    } else if (psi) {
      psi[k] = v[k]->hs.psi;
    }
  }
  return 0;
  CATN_CATCH
}

int catn_gradient_batch(catn_ctx* ctx, int32_t count, catn_net* const* nets, int32_t tape_cap,
                        double* lnZ, double* error, double* psi) {
  CATN_TRY
  std::vector<catn_net*> v(nets, nets + count);
  catn::gradient_nets(ctx->ex.get(), v, tape_cap > 0 ? tape_cap : catn::kDefaultTapeCap);
  for (int k = 0; k < count; ++k) {
    if (lnZ) lnZ[k] = v[k]->hs.lnZ + v[k]->hs.lntail;
    if (error) error[k] = v[k]->hs.err;
    if (psi) psi[k] = v[k]->hs.psi;
  }
  return 0;
  CATN_CATCH
}

int catn_ising_gradient(catn_net* net, double* dw, double* dh, double* dbeta) {
  CATN_TRY
  if (!net->is_ising) throw catn::Error(-2, "ising gradient: network was not built by ising_network");
  if (!net->have_adjoint) throw catn::Error(-5, "ising gradient: run catn_gradient_batch first");
  int64_t lo = INT64_MAX, hi = 0;
  for (const auto& is : net->isites) {
    const int64_t o = net->off0[is.slot];
    lo = std::min(lo, o);
    hi = std::max(hi, o + 8);
  }
  hi = std::min<int64_t>(hi, net->arena_cap);
  std::vector<double> adj0((size_t)(hi - lo), 0.0);
  if (hi > lo) catn::exec_of(net)->d2h(adj0.data(), net->hs.adj + lo, sizeof(double) * (size_t)(hi - lo));
  catn::ising_chain_rule(net, adj0, lo, dw, dh, dbeta);
  return 0;
  CATN_CATCH
}

int catn_net_site_adjoint(catn_net* net, int32_t slot, double* out, int64_t cap) {
  CATN_TRY
  if (!net->have_adjoint) throw catn::Error(-5, "site adjoint: run catn_gradient_batch first");
  if (slot < 0 || slot >= (int)net->off0.size() || net->off0[slot] < 0)
    throw catn::Error(-2, "site adjoint: slot was not an initial site");
  if (cap > 0) catn::exec_of(net)->d2h(out, net->hs.adj + net->off0[slot], sizeof(double) * (size_t)cap);
  return 0;
  CATN_CATCH
}

int catn_net_tensor_adjoint(catn_net* net, int32_t k, double* out, int64_t cap) {
  CATN_TRY
  if (!net->have_adjoint) throw catn::Error(-5, "tensor adjoint: run catn_gradient_batch first");
  if (!net->tensors_taped)
    throw catn::Error(-2, "tensor adjoint: the network was not built by catn_tensor_network");
  if (k < 0 || k >= (int)net->tensor_slot.size())
    throw catn::Error(-2, "tensor adjoint: no tensor " + std::to_string(k) + " in this network");
  const auto ts = net->tensor_slot[k];
  if (cap < 0 || cap > ts.second)
    throw catn::Error(-2, "tensor adjoint: cap exceeds the tensor's " + std::to_string(ts.second) +
                              " entries");
  if (cap > 0) {
    if (ts.first < 0 || ts.first >= (int)net->off0.size() || net->off0[ts.first] < 0)
      throw catn::Error(-2, "tensor adjoint: the tensor's data slot has no initial version");
    catn::exec_of(net)->d2h(out, net->hs.adj + net->off0[ts.first], sizeof(double) * (size_t)cap);
  }
  return 0;
  CATN_CATCH
}

int catn_net_tape(catn_net* net, int32_t* rec, double* val, int32_t cap) {
  CATN_TRY
  const int n = std::min<int>(cap, net->hs.tape_len);
  if (n <= 0 || !net->hs.tape) return 0;
  std::vector<catn::TapeRec> t((size_t)n);
  catn::exec_of(net)->d2h(t.data(), net->hs.tape, sizeof(catn::TapeRec) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    int32_t* r = rec + 16 * i;
    r[0] = t[i].code; r[1] = t[i].k; r[2] = t[i].path; r[3] = t[i].flag;
    r[4] = t[i].aux0; r[5] = t[i].aux1; r[6] = t[i].nin; r[7] = t[i].nout;
    for (int q = 0; q < 2; ++q) {
      r[8 + 4 * q] = t[i].in[q].dl; r[9 + 4 * q] = t[i].in[q].d; r[10 + 4 * q] = t[i].in[q].dr;
      r[11 + 4 * q] = (int32_t)t[i].in[q].off;
    }
    if (val) val[i] = t[i].val;
  }
  return 0;
  CATN_CATCH
}

int catn_net_bwd_profile(catn_net* net, int64_t* ticks, double* replay_dev, int32_t* tape_len) {
  for (int k = 0; k < 16; ++k)
    if (ticks) ticks[k] = net->hs.bwd_ticks[k];
  if (replay_dev) *replay_dev = net->hs.replay_dev;
  if (tape_len) *tape_len = net->hs.tape_len;
  return 0;
}

int catn_net_info(catn_net* net, double* sign, int32_t* maxdim, int32_t* num_isolated, int32_t* n,
                  double* beta) {
  if (sign) {
    sign[0] = net->hs.sign;
    if (net->cplx) sign[1] = net->hs.sign_i;
  }
  if (maxdim) *maxdim = net->plan().maxdim_intermediate;
  if (num_isolated) *num_isolated = net->num_isolated;
  if (n) *n = net->n;
  if (beta) *beta = net->beta;
  return 0;
}

int catn_net_edge_log(catn_net* net, int32_t* pairs, int32_t* count) {
  const int cap = *count;
  const catn::Planner& pl = net->plan();
  const int tot = (int)pl.edge_log.size();
  for (int k = 0; k < tot && k < cap; ++k) {
    pairs[2 * k] = pl.edge_log[k].first;
    pairs[2 * k + 1] = pl.edge_log[k].second;
  }
  *count = tot;
  return 0;
}

int catn_net_plan_trace(catn_net* net, int32_t* steps, int32_t* nsteps, int32_t* dups,
                        int32_t* ndups, int32_t* cuts, int32_t* ncuts) {
  const auto give = [](const std::vector<int32_t>& v, int w, int32_t* out, int32_t* n) {
    if (!n) return;
    const int tot = (int)v.size() / w;
    if (out)
      for (int k = 0; k < tot && k < *n; ++k)
        for (int q = 0; q < w; ++q) out[w * k + q] = v[(size_t)w * k + q];
    *n = tot;
  };
  const catn::Planner& pl = net->plan();
  give(pl.trace_steps, 8, steps, nsteps);
  give(pl.trace_dups, 3, dups, ndups);
  give(pl.cut_kept, 1, cuts, ncuts);
  return 0;
}

int catn_net_profile(catn_net* net, int64_t* op_ticks, int32_t* op_count, int64_t* svd_ticks,
                     int32_t* svd_count, double* tick_hz) {
  for (int k = 0; k < 16; ++k) {
    if (op_ticks) op_ticks[k] = net->hs.op_ticks[k];
    if (op_count) op_count[k] = net->hs.op_count[k];
  }
  for (int k = 0; k < 8; ++k) {
    if (svd_ticks) svd_ticks[k] = net->hs.svd_ticks[k];
    if (svd_count) svd_count[k] = net->hs.svd_count[k];
  }
  if (tick_hz) *tick_hz = net->ctx ? net->ctx->ex->tick_hz() : 1.0;
  return 0;
}

int catn_net_svd_profile(catn_net* net, int64_t* sub_ticks) {
  for (int k = 0; k < 16; ++k) sub_ticks[k] = net->hs.sub_ticks[k];
  return 0;
}

int catn_net_work(catn_net* net, double* flops, double* bytes, int64_t* scratch_hwm) {
  if (flops) *flops = net->hs.alg_flops;
  if (bytes) *bytes = net->hs.alg_bytes;
  if (scratch_hwm) *scratch_hwm = net->hs.scratch_hwm;
  return 0;
}

int catn_kernel_stats(catn_ctx* ctx, int64_t* launches, double* seconds, int32_t reset) {
  if (launches) *launches = ctx->ex->launches;
  if (seconds) *seconds = ctx->ex->kernel_seconds;
  if (reset) {
    ctx->ex->launches = 0;
    ctx->ex->kernel_seconds = 0.0;
    ctx->ex->plan_seconds = 0.0;
    ctx->ex->run_seconds = 0.0;
  }
  return 0;
}

int catn_bwd_kernel_stats(catn_ctx* ctx, int64_t* launches, double* seconds, int32_t reset) {
  if (launches) *launches = ctx->ex->bwd_launches;
  if (seconds) *seconds = ctx->ex->bwd_seconds;
  if (reset) {
    ctx->ex->bwd_launches = 0;
    ctx->ex->bwd_seconds = 0.0;
  }
  return 0;
}

int catn_host_stats(catn_ctx* ctx, double* plan_seconds, double* run_seconds) {
  if (plan_seconds) *plan_seconds = ctx->ex->plan_seconds;
  if (run_seconds) *run_seconds = ctx->ex->run_seconds;
  return 0;
}

static bool lanes_ok(int32_t t) {
  return t == 64 || t == 128 || t == 256 || t == 512 || t == 1024;
}

int catn_set_threads(catn_ctx* ctx, int32_t threads) {
  if (!lanes_ok(threads))
    return catn::fail(-2, "catn_set_threads: threads must be 64, 128, 256, 512 or 1024");
  ctx->ex->threads = threads;
  return 0;
}

int catn_set_launch(catn_ctx* ctx, int32_t threads, int32_t lds_doubles, int32_t panel,
                    int32_t jacobi_lanes) {
  if (!lanes_ok(threads))
    return catn::fail(-2, "catn_set_launch: threads must be 64, 128, 256, 512 or 1024");
  const int lim = ctx->ex->lds_limit();
  if (lds_doubles != 0 && (lds_doubles < catn::kLdsMinDoubles || lds_doubles > lim))
    return catn::fail(-2, "catn_set_launch: lds_doubles must be 0 (automatic) or in [" +
                              std::to_string(catn::kLdsMinDoubles) + ", " + std::to_string(lim) +
                              "] on this device");
  if (panel != 0 && panel != 8 && panel != 16 && panel != 32)
    return catn::fail(-2, "catn_set_launch: panel must be 0 (default 16), 8, 16 or 32");
  if (jacobi_lanes != 0 && jacobi_lanes != 8 && jacobi_lanes != 16 && jacobi_lanes != 32 &&
      jacobi_lanes != 64)
    return catn::fail(-2, "catn_set_launch: jacobi_lanes must be 0 (default 16), 8, 16, 32 or 64");
  if (jacobi_lanes > threads)
    return catn::fail(-2, "catn_set_launch: jacobi_lanes exceeds the lanes per workgroup");
  ctx->ex->threads = threads;
  ctx->ex->lds_override = lds_doubles;
  ctx->ex->qb = panel ? panel : catn::kQBDefault;
  ctx->ex->jac_lanes = jacobi_lanes ? jacobi_lanes : catn::kJacLanesDefault;
  return 0;
}

int catn_set_batch_streams(catn_ctx* ctx, int32_t streams) {
  if (streams < 1 || streams > catn::kGroupStreams)
    return catn::fail(-2, "catn_set_batch_streams: streams must be in [1, " +
                              std::to_string(catn::kGroupStreams) + "]");
  ctx->ex->batch_streams = streams;
  return 0;
}

int catn_get_launch(catn_ctx* ctx, int32_t* threads, int32_t* lds_doubles, int32_t* lds_limit,
                    int32_t* panel, int32_t* jacobi_lanes) {
  if (threads) *threads = ctx->ex->threads;
  if (lds_doubles) *lds_doubles = ctx->ex->lds_override;
  if (lds_limit) *lds_limit = ctx->ex->lds_limit();
  if (panel) *panel = ctx->ex->qb;
  if (jacobi_lanes) *jacobi_lanes = ctx->ex->jac_lanes;
  return 0;
}

int catn_net_mem(catn_net* net, int64_t* used, int64_t* cap) {
  if (used) *used = net->arena_hwm;
  if (cap) *cap = net->arena_cap;
  return 0;
}

#define NODE_OP(body)                      \
  CATN_TRY                                 \
  net->own_plan();                         \
  catn::PNode& nd = net->pl.node(id);      \
  (void)nd;                                \
  net->ensure_device();                    \
  net->reset_acc();                        \
  net->pl.acc_err = true;                  \
  body;                                    \
  catn::run_one(net);                      \
  if (error) *error = net->hs.err;         \
  return 0;                                \
  CATN_CATCH

int catn_node_swap(catn_net* net, int32_t id, int32_t i, int32_t j, double* error) {
  NODE_OP(net->pl.swap(nd, i, j))
}
int catn_node_move(catn_net* net, int32_t id, int32_t a, int32_t b, double* error) {
  NODE_OP(net->pl.move(nd, a, b))
}
int catn_node_move2tail(catn_net* net, int32_t id, int32_t idx, double* error) {
  NODE_OP(net->pl.move2tail(nd, idx))
}
int catn_node_move2head(catn_net* net, int32_t id, int32_t idx, double* error) {
  NODE_OP(net->pl.move2head(nd, idx))
}
int catn_node_merge(catn_net* net, int32_t id, int32_t j, int32_t cross, double* error) {
  NODE_OP(net->pl.merge(nd, j, cross != 0))
}
int catn_node_compress(catn_net* net, int32_t id, int32_t opt, double* error) {
  NODE_OP(net->pl.compress_node(nd, opt != 0))
}
int catn_node_cano_to(catn_net* net, int32_t id, int32_t idx) {
  double* error = nullptr;
  NODE_OP(net->pl.cano_to(nd, idx))
}
int catn_node_left_canonical(catn_net* net, int32_t id) {
  double* error = nullptr;
  NODE_OP(net->pl.left_canonical(nd))
}
int catn_node_reverse(catn_net* net, int32_t id) {
  double* error = nullptr;
  NODE_OP(net->pl.reverse(nd))
}

int catn_node_eat(catn_net* net, int32_t id_i, int32_t id_j, int32_t idx, int32_t idxi,
                  double* lognorm, double* error, double* phase) {
  CATN_TRY
  net->own_plan();
  catn::PNode& ni = net->pl.node(id_i);
  catn::PNode& nj = net->pl.node(id_j);
  net->ensure_device();
  net->reset_acc();
  net->pl.acc_err = true;
  net->pl.eat(ni, nj, idx, idxi);
  catn::run_one(net);
  if (lognorm) *lognorm = net->hs.lnZ;
  if (error) *error = net->hs.err;
  if (phase) {
    phase[0] = net->hs.psi;
    if (net->cplx) phase[1] = net->hs.psi_i;
  }
  return 0;
  CATN_CATCH
}

int catn_cut_bondim(catn_net* net, int32_t i, int32_t idx_j_in_i, int32_t opt, double* error) {
  CATN_TRY
  net->own_plan();
  if (net->pl.p.Dmax < 0) {   // no-op in exact mode (:343, :400)
    if (error) *error = 0.0;
    return 0;
  }
  net->ensure_device();
  net->reset_acc();
  net->pl.emit_cut(i, idx_j_in_i, opt != 0);
  catn::run_one(net);
  net->pl.apply_cut(net->read_result(net->pl.pending_cut_result()));
  if (error) *error = net->hs.err;
  return 0;
  CATN_CATCH
}

int catn_node_get(catn_net* net, int32_t id, int32_t* order, int32_t* cano, int32_t* neighbor,
                  int32_t* dims, int32_t cap) {
  CATN_TRY
  net->own_plan();
  catn::PNode& nd = net->pl.node(id);
  if (order) *order = nd.order();
  if (cano) *cano = nd.cano;
  if (neighbor)
    for (int k = 0; k < (int)nd.nb.size() && k < cap; ++k) neighbor[k] = nd.nb[k];
  if (dims) {
    net->ensure_device();
    for (int k = 0; k < nd.order() && k < cap; ++k) {
      catn::SiteDesc d = net->read_desc(nd.slot[k]);
      dims[3 * k] = d.dl; dims[3 * k + 1] = d.d; dims[3 * k + 2] = d.dr;
    }
  }
  return 0;
  CATN_CATCH
}

int catn_node_site(catn_net* net, int32_t id, int32_t k, double* out, int64_t cap) {
  CATN_TRY
  net->own_plan();
  catn::PNode& nd = net->pl.node(id);
  if (k < 1 || k > nd.order()) throw catn::Error(-2, "site index out of range");
  net->ensure_device();
  catn::SiteDesc d = net->read_desc(nd.slot[k - 1]);
  const int64_t cnt = (int64_t)d.dl * d.d * d.dr;
  if (cnt * net->ed() > cap) throw catn::Error(-2, "output buffer too small");
  if (net->f32) {
    const int64_t nv = cnt * net->ed();
    std::vector<float> fb(nv);
    catn::exec_of(net)->d2h(fb.data(), (float*)net->hs.arena + d.off * net->ed(), sizeof(float) * nv);
    for (int64_t q = 0; q < nv; ++q) out[q] = fb[q];
  } else {
    catn::exec_of(net)->d2h(out, net->hs.arena + d.off * net->ed(), net->elem_bytes() * cnt);
  }
  return 0;
  CATN_CATCH
}

int catn_result_tensor(catn_net* net, double* out, int64_t cap, int32_t* rank, int32_t* dims,
                       int32_t* labels, int32_t dims_cap) {
  CATN_TRY
  net->own_plan();
  // open-leg nodes: those with a site whose physical dimension is not 1
  // (src/tensor_network.jl:518-519), in ascending id
  std::vector<int> open;
  for (auto& kv : net->pl.nodes) {
    bool all1 = true;
    for (int d : kv.second.phys)
      if (d != 1) { all1 = false; break; }
    if (!all1) open.push_back(kv.first);
  }
  if (open.empty()) {   // closed network: 0-d one() (:520)
    if (rank) *rank = 0;
    if (!out) return 0;
    if (cap < net->ed()) throw catn::Error(-2, "result_tensor: output buffer too small");
    out[0] = 1.0;
    if (net->cplx) out[1] = 0.0;
    return 0;
  }
  if (open.size() != 1)   // :521
    throw catn::Error(-2, "result_tensor: expected one surviving open-leg node, got " +
                              std::to_string(open.size()) +
                              " (disconnected/multi-open networks are unsupported)");
  catn::PNode& nd = net->pl.node(open[0]);
  const int r = nd.order();
  if (r > dims_cap) throw catn::Error(-2, "result_tensor: dims buffer too small");
  int64_t total = 1;
  for (int k = 0; k < r; ++k) {
    if (dims) dims[k] = nd.phys[k];
    if (labels) labels[k] = nd.nb[k];
    total *= nd.phys[k];
  }
  if (rank) *rank = r;
  if (!out) return 0;   // query: rank, dims and labels only
  if (total * net->ed() > cap) throw catn::Error(-2, "result_tensor: output buffer too small");
  net->ensure_device();
  if (r > 1) {   // mps2raw on the device: the sites fused into one
    const double lz = net->hs.lnZ, er = net->hs.err, ps = net->hs.psi, sg = net->hs.sign;
    const double psi_i = net->hs.psi_i, sg_i = net->hs.sign_i;
    net->pl.fuse_all(nd);
    catn::run_one(net);
    net->hs.lnZ = lz; net->hs.err = er; net->hs.psi = ps; net->hs.sign = sg;
    net->hs.psi_i = psi_i; net->hs.sign_i = sg_i;
  }
  const catn::SiteDesc d = net->read_desc(nd.slot[0]);
  const int64_t cnt = (int64_t)d.dl * d.d * d.dr;
  if (cnt != total) throw catn::Error(-3, "result_tensor: fused site has the wrong size");
  if (net->f32) {
    const int64_t nv = cnt * net->ed();
    std::vector<float> fb(nv);
    catn::exec_of(net)->d2h(fb.data(), (float*)net->hs.arena + d.off * net->ed(), sizeof(float) * nv);
    for (int64_t q = 0; q < nv; ++q) out[q] = fb[q];
  } else {
    catn::exec_of(net)->d2h(out, net->hs.arena + d.off * net->ed(), net->elem_bytes() * cnt);
  }
  return 0;
  CATN_CATCH
}

int catn_tsvd(catn_ctx* ctx, int32_t m, int32_t n, const double* A, double cutoff, int32_t maxdim,
              int32_t* k, double* U, double* S, double* V, double* discarded) {
  CATN_TRY
  catn_net net;
  net.attach(ctx);
  catn_params p;
  catn_params_default(&p);
  apply_params(&net, &p);
  const int64_t need = (int64_t)m * n;
  net.arena_cap = 4 * need + 4 * ((int64_t)m + n) * std::min(m, n) + 4096;
  net.scratch_cap = 8 * need + 8 * ((int64_t)std::max(m, n)) * std::max(m, n) + 4096;
  const int a = net.new_slot(m, 1, n, A);
  const int su = net.new_slot(0, 0, 0, nullptr);
  const int sv = net.new_slot(0, 0, 0, nullptr);
  const int ss = net.new_slot(0, 0, 0, nullptr);
  net.pl.nresults = 1;
  net.pl.emit(catn::OP_TSVD, a, su, sv, maxdim <= 0 ? (1 << 30) : maxdim, 0, ss, cutoff);
  catn::run_one(&net);
  const int kk = net.read_result(0);
  double disc = 0.0;
  ctx->ex->d2h(&disc, net.hs.dresults, sizeof(double));
  catn::SiteDesc du = net.read_desc(su), dv = net.read_desc(sv), dsd = net.read_desc(ss);
  if (U) ctx->ex->d2h(U, net.hs.arena + du.off, sizeof(double) * (size_t)m * kk);
  if (V) ctx->ex->d2h(V, net.hs.arena + dv.off, sizeof(double) * (size_t)n * kk);
  if (S) ctx->ex->d2h(S, net.hs.arena + dsd.off, sizeof(double) * (size_t)dsd.dl);
  if (k) *k = kk;
  if (discarded) *discarded = disc;
  return 0;
  CATN_CATCH
}

int catn_tsvd_fixed(catn_ctx* ctx, int32_t m, int32_t n, const double* dA, double cutoff,
                    int32_t maxdim, int32_t* k, double* dU, double* dS, double* dV,
                    double* discarded) {
  CATN_TRY
  if (m <= 0 || n <= 0) throw catn::Error(-2, "tsvd_fixed: m and n must be positive");
  catn::Executor* ex = ctx->ex.get();
  size_t cap = 0;
  char* base = (char*)ex->fixed_region(&cap);
  if (!base) throw catn::Error(-2, "tsvd_fixed: this library has no fixed device region");
  // layout of the fixed region, 256-byte aligned blocks: instance state, op,
  // op range, 4 site descriptors, result slots, arena, scratch (the same
  // arena and scratch sizes catn_tsvd allocates)
  const int64_t mn = (int64_t)m * n, r = std::min(m, n), mx = std::max(m, n);
  const int64_t acap = 4 * mn + 4 * ((int64_t)m + n) * r + 4096;
  const int64_t scap = 8 * mn + 8 * mx * mx + 4096;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t o_state = 0;
  const size_t o_op = al(o_state + sizeof(catn::InstState));
  const size_t o_rng = al(o_op + sizeof(catn::Op));
  const size_t o_sites = al(o_rng + 2 * sizeof(int32_t));
  const size_t o_res = al(o_sites + 4 * sizeof(catn::SiteDesc));
  const size_t o_dres = al(o_res + sizeof(int32_t));
  const size_t o_arena = al(o_dres + sizeof(double));
  const size_t o_scr = al(o_arena + sizeof(double) * (size_t)acap);
  const size_t need = o_scr + sizeof(double) * (size_t)scap;
  if (need > cap)
    throw catn::Error(-3, "tsvd_fixed: a " + std::to_string(m) + " x " + std::to_string(n) +
                              " matrix needs " + std::to_string(need) + " bytes of the fixed "
                              "device region, which holds " + std::to_string(cap) +
                              " (at most about 512 x 512)");
  catn::InstState hs{};
  hs.psi = 1.0;
  hs.sign = 1.0;
  hs.arena = (double*)(base + o_arena);
  hs.scratch = (double*)(base + o_scr);
  hs.sites = (catn::SiteDesc*)(base + o_sites);
  hs.results = (int32_t*)(base + o_res);
  hs.dresults = (double*)(base + o_dres);
  hs.arena_cap = acap;
  hs.scratch_cap = scap;
  hs.nsites = 4;
  hs.nresults = 1;
  hs.seed = 1;
  hs.arena_top = (mn + 31) & ~int64_t(31);   // A at offset 0
  // A into the arena, the descriptors (A, then three placeholders written by the op)
  ex->d2d(hs.arena, dA, sizeof(double) * (size_t)mn);
  catn::SiteDesc sd[4] = {};
  sd[0].off = 0; sd[0].dl = m; sd[0].d = 1; sd[0].dr = n; sd[0].live = 1;
  ex->h2d(hs.sites, sd, sizeof(sd));
  catn::Op op{};
  op.code = catn::OP_TSVD;
  op.a = 0; op.b = 1; op.c = 2; op.f = 3; op.e = 0;
  op.d = maxdim <= 0 ? (1 << 30) : maxdim;
  op.x = cutoff;
  ex->run_fixed(&hs, base + o_state, op, base + o_op, base + o_rng);
  if (hs.status != catn::STATUS_OK)
    throw catn::Error(-10 - hs.status, "tsvd_fixed: device op failed with status " +
                                           std::to_string(hs.status));
  int32_t kk = 0;
  double disc = 0.0;
  ex->d2h(&kk, hs.results, sizeof(int32_t));
  ex->d2h(&disc, hs.dresults, sizeof(double));
  ex->d2h(sd, hs.sites, sizeof(sd));
  if (dU) ex->d2d(dU, hs.arena + sd[1].off, sizeof(double) * (size_t)m * kk);
  if (dV) ex->d2d(dV, hs.arena + sd[2].off, sizeof(double) * (size_t)n * kk);
  if (dS) ex->d2d(dS, hs.arena + sd[3].off, sizeof(double) * (size_t)sd[3].dl);
  if (k) *k = kk;
  if (discarded) *discarded = disc;
  return 0;
  CATN_CATCH
}

int catn_rsvd(catn_ctx* ctx, int32_t m, int32_t n, const double* A, int32_t k, int32_t oversample,
              int32_t power, int32_t seed, int32_t* kk, double* U, double* S, double* V) {
  CATN_TRY
  if (m <= 0 || n <= 0 || k <= 0 || oversample <= 0 || power < 0)
    throw catn::Error(-2, "rsvd: dimensions, k and oversample must be positive, power >= 0");
  const int64_t p = std::min<int64_t>(n, (int64_t)oversample * k);
  if (p > m) throw catn::Error(-2, "rsvd: the sketch has more columns than A has rows");
  catn_net net;
  net.attach(ctx);
  catn_params prm;
  catn_params_default(&prm);
  prm.seed = seed;
  apply_params(&net, &prm);
  const int64_t need = (int64_t)m * n;
  net.arena_cap = 2 * need + 4 * ((int64_t)m + n) * p + 4096;
  net.scratch_cap = 4 * need + 8 * ((int64_t)m + n) * p + 8 * p * p + 4096;
  const int a = net.new_slot(m, 1, n, A);
  const int su = net.new_slot(0, 0, 0, nullptr);
  const int sv = net.new_slot(0, 0, 0, nullptr);
  const int ss = net.new_slot(0, 0, 0, nullptr);
  net.pl.nresults = 1;
  net.pl.emit(catn::OP_TSVD, a, su, sv, k, 0, ss, 0.0, 1, oversample, power);
  catn::run_one(&net);
  const int kq = net.read_result(0);
  catn::SiteDesc du = net.read_desc(su), dv = net.read_desc(sv), dsd = net.read_desc(ss);
  if (U) ctx->ex->d2h(U, net.hs.arena + du.off, sizeof(double) * (size_t)m * kq);
  if (V) ctx->ex->d2h(V, net.hs.arena + dv.off, sizeof(double) * (size_t)n * kq);
  if (S) ctx->ex->d2h(S, net.hs.arena + dsd.off, sizeof(double) * (size_t)kq);
  if (kk) *kk = kq;
  return 0;
  CATN_CATCH
}

}  // extern "C"
