// Multi-threaded CPU emulator of libcatn_hip — TEST BUILD ONLY.
//
// Runs every workgroup of a launch with one host thread per lane (256 by
// default, the product's workgroup size), so that the lane-parallel code
// paths of engine.h / adjoint.h (lane groups, strided loops, cross-lane
// reductions, LDS budgets) execute on the CPU the way they do on the device.
// Built with AddressSanitizer it turns an out-of-bounds access of any lane
// into a report with a stack trace instead of a GPU memory fault:
//   * device buffers are heap blocks of exactly the requested size;
//   * the dynamic LDS region is a heap block of exactly the launch's budget;
//   * every barrier checks that all lanes arrive from the same call site
//     (a divergent barrier is reported with both sites), and a barrier or a
//     lane-group reduction that some lanes never reach is reported after a
//     time limit instead of hanging.
// Cross-lane reductions give every lane the same value in a fixed order, as
// the device's butterfly reductions do.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#define CATN_EMU_MT 1
#include "../../catn.jl_amd/csrc/engine.h"
#include "../../catn.jl_amd/csrc/capi_impl.h"

namespace catn {

static double emu_timeout_s() {
  const char* e = getenv("CATN_EMU_TIMEOUT");
  return e ? atof(e) : 120.0;
}

struct EmuShared {
  int nthr = 256;
  int lds_cap = 0;
  double* lds = nullptr;
  std::vector<double> xb;          // per-lane exchange slots of the reductions
  // checked barrier
  std::mutex m;
  std::condition_variable cv;
  int count = 0;
  long gen = 0;
  std::vector<const void*> sites;
  long divergent = 0;
  // lane-group reductions: one slot per possible group leader
  struct alignas(64) GSlot {
    std::atomic<int> cnt{0};
    std::atomic<long> gen{0};
  };
  std::vector<GSlot> gs;
  std::vector<double> gx;
  // wavefront primitives (CATN_EMU_DEV): per-lane shuffle slots, one counter
  // per offset 1..32, and matrix-core operand slots
  struct alignas(64) ShSlot {
    std::atomic<long> cnt[6] = {};
    double val[6][2] = {};
  };
  struct alignas(64) MfSlot {
    std::atomic<long> cnt{0};
    double a[2] = {}, b[2] = {};
    short ha[2][4] = {}, hb[2][4] = {};
  };
  std::vector<ShSlot> sx;
  std::vector<MfSlot> mf;
  explicit EmuShared(int n, int lds_doubles)
      : nthr(n), lds_cap(lds_doubles), xb(n), sites(n), gs(n), gx(n), sx(n), mf(n) {
    lds = new double[lds_doubles > 0 ? lds_doubles : 1];
    const uint64_t sentinel = 0x7ff4deadbeef0001ull;   // a NaN no kernel writes
    for (int i = 0; i < lds_doubles; ++i) memcpy(lds + i, &sentinel, 8);
  }
  // one past the highest LDS double the launch wrote
  int lds_high_water() const {
    const uint64_t sentinel = 0x7ff4deadbeef0001ull;
    for (int i = lds_cap - 1; i >= 0; --i)
      if (memcmp(lds + i, &sentinel, 8) != 0) return i + 1;
    return 0;
  }
  ~EmuShared() { delete[] lds; }
};

struct EmuLane {
  EmuShared* sh = nullptr;
  int tid = 0;
  const void* hist[32] = {};   // the lane's last barrier sites (ring)
  long nb = 0;
};
static thread_local EmuLane t_lane;

// A barrier site is the (file, line) of the sync() call, interned to a
// stable pointer so that sites compare by address.
struct Site {
  const char* file;
  int line;
};
static std::mutex g_site_m;
static const Site* intern_site(const char* file, int line) {
  static std::vector<Site*> all;
  std::lock_guard<std::mutex> lk(g_site_m);
  for (Site* s : all)
    if (s->line == line && strcmp(s->file, file) == 0) return s;
  Site* s = new Site{file, line};
  all.push_back(s);
  return s;
}
static const char* site_str(const void* p, char* buf, size_t n) {
  const Site* s = (const Site*)p;
  if (!s) return "(none)";
  const char* f = strrchr(s->file, '/');
  snprintf(buf, n, "%s:%d", f ? f + 1 : s->file, s->line);
  return buf;
}

static void emu_die(const char* what) {
  fprintf(stderr, "[catn-emu-mt] %s (lane %d)\n", what, t_lane.tid);
  fflush(stderr);
  abort();
}

static std::vector<EmuLane*>* g_lanes = nullptr;   // lanes of the running launch
extern std::atomic<long> g_divergent;
extern std::atomic<long> g_lds_hwm;
static void note_divergent() { g_divergent.fetch_add(1); }

static void dump_history(int lane) {
  if (!g_lanes || lane >= (int)g_lanes->size()) return;
  EmuLane* L = (*g_lanes)[lane];
  fprintf(stderr, "[catn-emu-mt] lane %d: last barrier sites (oldest first):", lane);
  char b[256];
  for (long i = L->nb > 32 ? L->nb - 32 : 0; i < L->nb; ++i)
    fprintf(stderr, " %s", site_str(L->hist[i & 31], b, sizeof b));
  fprintf(stderr, "\n");
}

static void barrier_at(const void* site) {
  EmuShared* sh = t_lane.sh;
  t_lane.hist[t_lane.nb++ & 31] = site;
  std::unique_lock<std::mutex> lk(sh->m);
  sh->sites[t_lane.tid] = site;
  const long g = sh->gen;
  if (++sh->count == sh->nthr) {
    for (int i = 1; i < sh->nthr; ++i) {
      if (sh->sites[i] != sh->sites[0]) {
        note_divergent();
        if (sh->divergent++ < 8)
        {
          char b0[256], b1[256];
          fprintf(stderr, "[catn-emu-mt] divergent barrier: lane 0 at %s, lane %d at %s\n",
                  site_str(sh->sites[0], b0, sizeof b0), i, site_str(sh->sites[i], b1, sizeof b1));
        }
        if (getenv("CATN_EMU_STRICT")) {
          dump_history(0);
          dump_history(i);
          emu_die("divergent barrier");
        }
        break;
      }
    }
    sh->count = 0;
    sh->gen++;
    sh->cv.notify_all();
    return;
  }
  // plain wait (intercepted by ThreadSanitizer); the launch's watchdog
  // reports a barrier that some lanes never reach
  sh->cv.wait(lk, [&] { return sh->gen != g; });
}

// Watchdog of one launch: a barrier generation that does not advance within
// the time limit while lanes wait in it is reported with the missing lanes'
// last barrier sites.
static void watchdog(EmuShared* sh, std::atomic<bool>* done) {
  long last = -1;
  auto since = std::chrono::steady_clock::now();
  while (!done->load()) {
    std::this_thread::sleep_for(std::chrono::milliseconds(200));
    long g;
    int cnt;
    {
      std::lock_guard<std::mutex> lk(sh->m);
      g = sh->gen;
      cnt = sh->count;
    }
    if (g != last || cnt == 0) {
      last = g;
      since = std::chrono::steady_clock::now();
      continue;
    }
    if (std::chrono::duration<double>(std::chrono::steady_clock::now() - since).count() >
        emu_timeout_s()) {
      fprintf(stderr, "[catn-emu-mt] barrier not reached by all lanes: %d of %d arrived\n", cnt,
              sh->nthr);
      dump_history(0);
      for (int i = 1; i < sh->nthr; ++i)
        if (g_lanes && (*g_lanes)[i] && (*g_lanes)[0] &&
            (*g_lanes)[i]->nb != (*g_lanes)[0]->nb) {
          dump_history(i);
          break;
        }
      fflush(stderr);
      abort();
    }
  }
}

void emu_sync(const char* file, int line) { barrier_at(intern_site(file, line)); }

static double wg_reduce(double v, bool is_max) {
  EmuShared* sh = t_lane.sh;
  sh->xb[t_lane.tid] = v;
  static const Site* s1 = intern_site("emu_wg_reduce", 1);
  static const Site* s2 = intern_site("emu_wg_reduce", 2);
  barrier_at(s1);
  double s = sh->xb[0];
  for (int i = 1; i < sh->nthr; ++i) s = is_max ? fmax(s, sh->xb[i]) : s + sh->xb[i];
  barrier_at(s2);
  return s;
}

double emu_wg_sum(double v) { return wg_reduce(v, false); }
double emu_wg_max(double v) { return wg_reduce(v, true); }

static void group_wait(EmuShared::GSlot& s, int L) {
  const long g0 = s.gen.load(std::memory_order_acquire);
  if (s.cnt.fetch_add(1, std::memory_order_acq_rel) == L - 1) {
    s.cnt.store(0, std::memory_order_relaxed);
    s.gen.fetch_add(1, std::memory_order_acq_rel);
    return;
  }
  const auto t0 = std::chrono::steady_clock::now();
  int spins = 0;
  while (s.gen.load(std::memory_order_acquire) == g0) {
    if (++spins > 64) {
      std::this_thread::yield();
      if ((spins & 4095) == 0 &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >
              emu_timeout_s()) {
        fprintf(stderr, "[catn-emu-mt] lane group of %d not complete\n", L);
        emu_die("group reduction hang");
      }
    }
  }
}

double emu_grp_sum(double v, int L) {
  if (L <= 1) return v;
  EmuShared* sh = t_lane.sh;
  const int tid = t_lane.tid;
  if (L > 64 || (L & (L - 1))) emu_die("group size is not a power of two <= 64");
  const int base = tid & ~(L - 1);
  if (base + L > sh->nthr) emu_die("lane group extends past the workgroup");
  sh->gx[tid] = v;
  group_wait(sh->gs[base], L);
  double s = 0.0;
  for (int i = 0; i < L; ++i) s += sh->gx[base + i];
  group_wait(sh->gs[base], L);
  return s;
}

double* emu_lds_base() { return t_lane.sh->lds; }

// CATN_EMU_REPOISON=1: at the start of every op and reverse step, lane 0
// refills the instance's scratch up to its high-water mark and the launch's
// LDS with a NaN pattern (all lanes wait before and after)
static bool g_repoison = getenv("CATN_EMU_REPOISON") != nullptr;
void emu_op_begin(const WG& g, InstState* st) {
  if (!g_repoison) return;
  g.sync();
  if (g.tid == 0) {
    const int64_t n = std::min<int64_t>(st->scratch_cap, std::max<int64_t>(st->scratch_hwm, 1 << 18));
    memset(st->scratch, 0xFF, sizeof(double) * (size_t)n);
    EmuShared* sh = t_lane.sh;
    memset(sh->lds, 0xFF, sizeof(double) * (size_t)sh->lds_cap);
  }
  g.sync();
}

#if defined(CATN_EMU_DEV)
// Wavefront primitives of the device code paths (CATN_EMU_DEV build). Each
// call is a rendezvous of the lanes the hardware instruction reads:
//   * shfl_xor(v, o): this lane and lane ^ o, matched per offset by a call
//     counter, so both lanes must execute the same number of shuffles at
//     that offset (on the device an inactive partner lane returns a stale
//     register);
//   * the matrix-core products: all 64 lanes of the wavefront (the device
//     loops around them have wavefront-uniform bounds).
// Values are exchanged through per-lane slots alternating between two
// parities: a lane overwrites a parity only after its partners have passed
// the following call, which they reach only after reading it. A lane that
// waits longer than the time limit (a partner never arrives) is reported.
std::atomic<long> g_mfma_f64{0}, g_mfma_bf16{0}, g_shfl{0};   // emulated instructions

static void spin_until(const std::atomic<long>& c, long target, const char* what) {
  const auto t0 = std::chrono::steady_clock::now();
  for (int spins = 0; c.load(std::memory_order_acquire) < target;) {
    if (++spins > 64) {
      std::this_thread::yield();
      if ((spins & 4095) == 0 &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >
              emu_timeout_s())
        emu_die(what);
    }
  }
}

double emu_shfl_xor(double v, int o) {
  EmuShared* sh = t_lane.sh;
  if (t_lane.tid == 0) g_shfl.fetch_add(1, std::memory_order_relaxed);
  const int t = t_lane.tid, p = t ^ o;
  if (o <= 0 || o >= 64 || (o & (o - 1))) emu_die("shuffle offset is not a power of two < 64");
  if (p >= sh->nthr) emu_die("shuffle partner outside the workgroup");
  int q = 0;
  while ((1 << q) != o) ++q;
  EmuShared::ShSlot& me = sh->sx[t];
  const long n = me.cnt[q].load(std::memory_order_relaxed);
  me.val[q][n & 1] = v;
  me.cnt[q].store(n + 1, std::memory_order_release);
  EmuShared::ShSlot& other = sh->sx[p];
  spin_until(other.cnt[q], n + 1, "shuffle partner never arrived");
  return other.val[q][n & 1];
}

// the call number of this lane's next matrix-core product
static long wave_deposit(EmuShared* sh, int t) {
  return sh->mf[t].cnt.load(std::memory_order_relaxed);
}
// wait until every lane of this lane's wavefront deposited call n
static void wave_wait(EmuShared* sh, int t, long n) {
  const int w0 = t & ~63;
  if (w0 + 64 > sh->nthr) emu_die("matrix-core product in a partial wavefront");
  for (int j = w0; j < w0 + 64; ++j)
    spin_until(sh->mf[j].cnt, n + 1, "matrix-core product: a lane of the wavefront never arrived");
}

catn_d4 emu_mfma_f64(double a, double b, catn_d4 c) {
  EmuShared* sh = t_lane.sh;
  if ((t_lane.tid & 63) == 0) g_mfma_f64.fetch_add(1, std::memory_order_relaxed);
  const int t = t_lane.tid, l = t & 63, w0 = t & ~63;
  const long n = wave_deposit(sh, t);
  EmuShared::MfSlot& me = sh->mf[t];
  me.a[n & 1] = a;
  me.b[n & 1] = b;
  me.cnt.store(n + 1, std::memory_order_release);
  wave_wait(sh, t, n);
  catn_d4 d;
  for (int v = 0; v < 4; ++v) {
    const int r = l / 16 + 4 * v, col = l % 16;
    double s = c[v];
    for (int k = 0; k < 4; ++k) s += sh->mf[w0 + r + 16 * k].a[n & 1] * sh->mf[w0 + col + 16 * k].b[n & 1];
    d[v] = s;
  }
  return d;
}

static float bf16_to_f32(short h) {
  const uint32_t u = (uint32_t)(uint16_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

catn_f32x4 emu_mfma_bf16(catn_bf16x4 a, catn_bf16x4 b, catn_f32x4 c) {
  EmuShared* sh = t_lane.sh;
  if ((t_lane.tid & 63) == 0) g_mfma_bf16.fetch_add(1, std::memory_order_relaxed);
  const int t = t_lane.tid, l = t & 63, w0 = t & ~63;
  const long n = wave_deposit(sh, t);
  EmuShared::MfSlot& me = sh->mf[t];
  for (int v = 0; v < 4; ++v) {
    me.ha[n & 1][v] = a[v];
    me.hb[n & 1][v] = b[v];
  }
  me.cnt.store(n + 1, std::memory_order_release);
  wave_wait(sh, t, n);
  catn_f32x4 d;
  for (int v = 0; v < 4; ++v) {
    const int r = 4 * (l / 16) + v, col = l % 16;
    float s = c[v];
    for (int k = 0; k < 16; ++k) {
      const short x = sh->mf[w0 + r + 16 * (k / 4)].ha[n & 1][k % 4];
      const short y = sh->mf[w0 + col + 16 * (k / 4)].hb[n & 1][k % 4];
      s += bf16_to_f32(x) * bf16_to_f32(y);
    }
    d[v] = s;
  }
  return d;
}
#endif

class MtExecutor : public Executor {
 public:
  MtExecutor() { threads = 256; }
#if defined(CATN_EMU_DEV)
  const char* name() const override { return "cpu-emulator-dev"; }
#else
  const char* name() const override { return "cpu-emulator-mt"; }
#endif
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

  template <class F>
  void launch(int lds_cap, InstState* st, F body) {
    EmuShared sh(threads, lds_cap);
    double red[32];
    int flag = 0;
    std::vector<std::thread> th;
    std::vector<EmuLane*> lanes(threads, nullptr);
    g_lanes = &lanes;
    std::atomic<int> ready{0};
    th.reserve(threads);
    for (int t = 0; t < threads; ++t) {
      th.emplace_back([&, t] {
        t_lane = EmuLane();
        t_lane.sh = &sh;
        t_lane.tid = t;
        lanes[t] = &t_lane;
        ready.fetch_add(1);
        while (ready.load() < threads) std::this_thread::yield();
        WG g;
        g.tid = t;
        g.nthr = threads;
        g.red = red;
        g.flag = &flag;
        g.lds_cap = lds_cap;
        g.qb = qb;
        g.jac_lanes = jac_lanes;
        g.sub = st->sub_ticks;
        body(g);
      });
    }
    std::atomic<bool> done{false};
    std::thread wd(watchdog, &sh, &done);
    for (auto& x : th) x.join();
    done.store(true);
    wd.join();
    g_lanes = nullptr;
    const int hw = sh.lds_high_water();
    long cur = g_lds_hwm.load();
    while (hw > cur && !g_lds_hwm.compare_exchange_weak(cur, hw)) {
    }
  }

  void run(const std::vector<InstState*>& states,
           const std::vector<const std::vector<Op>*>& programs) override {
    int lds = lds_for((int)states.size());
    if (getenv("CATN_EMU_LDS")) lds = atoi(getenv("CATN_EMU_LDS"));   // a device batch's budget
    for (size_t b = 0; b < states.size(); ++b) {
      const std::vector<Op>& p = *programs[b];
      InstState* st = states[b];
      launch(lds, st, [&](const WG& g) { run_program(g, st, p.data(), 0, (int)p.size()); });
    }
  }
  void run_backward(const std::vector<InstState*>& states, int lds_cap) override {
    if (lds_cap <= 0 || lds_cap > kLdsMaxDoubles) lds_cap = kLdsMaxDoubles;
    for (InstState* st : states)
      launch(lds_cap, st, [&](const WG& g) { catn::run_backward(g, st); });
  }
};

Executor* make_executor(int) { return new MtExecutor(); }

std::atomic<long> g_divergent{0};
std::atomic<long> g_lds_hwm{0};

}  // namespace catn

// barriers reached from different call sites since the library was loaded
extern "C" long catn_emu_divergent_barriers(void) { return catn::g_divergent.load(); }
#if defined(CATN_EMU_DEV)
// wavefront instructions emulated since the last call (FP64 matrix-core,
// BF16 matrix-core, shuffles of lane 0); resets
extern "C" void catn_emu_wave_ops(long* out) {
  out[0] = catn::g_mfma_f64.exchange(0);
  out[1] = catn::g_mfma_bf16.exchange(0);
  out[2] = catn::g_shfl.exchange(0);
}
#endif
// highest LDS extent (doubles) any launch wrote since the last call; resets
extern "C" long catn_emu_lds_high_water(void) { return catn::g_lds_hwm.exchange(0); }
