// Workgroup-level numerics of the MPS edge step and the op interpreter.
//
// Single source: compiled by hipcc for gfx950 (one workgroup executes one
// instance's program; lanes cooperate through LDS and wavefront shuffles) and
// by g++ for the CPU emulator used in tests (one "workgroup" of one thread,
// barriers become no-ops). Every routine is written in terms of
// (tid, nthr, barrier) so both builds run the same algorithm.
//
// Numerics (all FP64 here):
//   * theta of swap!            ein"ijk,kab->iajb" without a materialised permute
//   * truncated SVD (tsvd)       Householder QR for tall/wide operands, then a
//                               one-sided (Hestenes) Jacobi SVD with round-robin
//                               pair ordering; singular values carry absolute
//                               error O(eps*||A||) like LAPACK gesdd, so the
//                               reference's absolute cutoff keeps its meaning
//   * folds, fuses, permutes, norms (streaming, coalesced over the leading index)
#pragma once
#include <math.h>
#include <stdint.h>
#include "ops.h"

#if defined(__HIPCC__)
#define CATN_D __device__ __forceinline__
#define CATN_DN __device__ __noinline__
#define CATN_HD __host__ __device__ inline
#else
#define CATN_D inline
#define CATN_DN
#define CATN_HD inline
#endif

// CATN_DEV selects the device code paths: FP64/BF16 matrix-core products,
// wavefront shuffles and the one-barrier reductions. They are compiled for
// gfx950 and, with CATN_EMU_DEV, by the test-only wavefront emulator
// (tests/emul/catn_emul_mt.cpp), which supplies the three wavefront
// primitives below on host threads so that the device's own operand and
// accumulator index arithmetic runs under the host sanitizers.
#if defined(__HIP_DEVICE_COMPILE__) || defined(CATN_EMU_DEV)
#define CATN_DEV 1
#else
#define CATN_DEV 0
#endif

namespace catn {

#if defined(CATN_EMU_MT)
// Test-only multi-threaded CPU emulation (tests/emul/catn_emul_mt.cpp): every
// lane of a workgroup is a host thread; these hooks stand in for the
// workgroup barrier, the cross-lane reductions and the LDS base.
void emu_sync(const char* file, int line);
double emu_wg_sum(double v);
double emu_wg_max(double v);
double emu_grp_sum(double v, int L);
double* emu_lds_base();
// start of every op and reverse step: with CATN_EMU_REPOISON set, the
// scratch used so far and the LDS are refilled with a NaN pattern, so a read
// of either before this op wrote it shows up in the results
struct InstState;
struct WG;
void emu_op_begin(const WG& g, InstState* st);
#define CATN_EMU_OP_BEGIN(g, st) emu_op_begin(g, st)
#else
#define CATN_EMU_OP_BEGIN(g, st) ((void)0)
#endif

// wavefront primitives of the device paths:
//   shfl_xor(v, o)            value of v in lane (this lane ^ o), o < 64;
//   mfma_f64_16x16x4(a, b, c) v_mfma_f64_16x16x4f64: lane l holds A(l % 16,
//                             l / 16), B(l / 16, l % 16) and D = A B + C at
//                             (l / 16 + 4 v, l % 16) in element v;
//   mfma_bf16_16x16x16(a, b, c) v_mfma_f32_16x16x16bf16_1k: lane l holds
//                             A(l % 16, 4 (l / 16) + v), B(4 (l / 16) + v,
//                             l % 16) and D at (4 (l / 16) + v, l % 16).
#if defined(__HIP_DEVICE_COMPILE__)
typedef double catn_d4 __attribute__((ext_vector_type(4)));
typedef short catn_bf16x4 __attribute__((ext_vector_type(4)));
typedef float catn_f32x4 __attribute__((ext_vector_type(4)));
CATN_D double shfl_xor(double v, int o) { return __shfl_xor(v, o, 64); }
CATN_D catn_d4 mfma_f64_16x16x4(double a, double b, catn_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
CATN_D catn_f32x4 mfma_bf16_16x16x16(catn_bf16x4 a, catn_bf16x4 b, catn_f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
#elif defined(CATN_EMU_DEV)
typedef double catn_d4 __attribute__((vector_size(32)));
typedef short catn_bf16x4 __attribute__((vector_size(8)));
typedef float catn_f32x4 __attribute__((vector_size(16)));
double emu_shfl_xor(double v, int o);
catn_d4 emu_mfma_f64(double a, double b, catn_d4 c);
catn_f32x4 emu_mfma_bf16(catn_bf16x4 a, catn_bf16x4 b, catn_f32x4 c);
inline double shfl_xor(double v, int o) { return emu_shfl_xor(v, o); }
inline catn_d4 mfma_f64_16x16x4(double a, double b, catn_d4 c) { return emu_mfma_f64(a, b, c); }
inline catn_f32x4 mfma_bf16_16x16x16(catn_bf16x4 a, catn_bf16x4 b, catn_f32x4 c) {
  return emu_mfma_bf16(a, b, c);
}
#endif

// ---------------------------------------------------------------------------
// workgroup context
// ---------------------------------------------------------------------------
// tile parameters of a launch (set with the launch shape, picked by the
// autotuner): Householder panel width and lanes per column pair of the
// global-memory Jacobi
constexpr int kQBMax = 32;            // widest panel (T factors are stored kQBMax x n)
constexpr int kQBDefault = 16;
constexpr int kJacLanesDefault = 16;

struct WG {
  int tid, nthr;
  double* red;     // LDS, 32 doubles (two halves of one per wavefront, <= 16 wavefronts)
  int* flag;       // LDS int
  int lds_cap;     // doubles available in the dynamic LDS region
  int64_t* sub;    // sub-step tick accumulators (instance state), or null
  int qb = kQBDefault;               // Householder panel width (8, 16 or 32)
  int jac_lanes = kJacLanesDefault;  // global Jacobi lanes per pair (8..64)
  int svd_fault = 0;                 // 1: every SVD's primary path counts as not converged
  // Jacobi pair test floor (cosine): 0 for FP64 storage (4 eps sqrt(m) only);
  // FP32 storage stops rotating at 2^-26, below the rounding of its outputs
  double jac_tol = 0.0;
#if defined(CATN_EMU_MT)
  // the emulator checks that every lane arrives from the same call site
  void sync(const char* file = __builtin_FILE(), int line = __builtin_LINE()) const {
    emu_sync(file, line);
  }
#else
  CATN_D void sync() const {
#if defined(__HIP_DEVICE_COMPILE__)
    __syncthreads();
#endif
  }
#endif
};

CATN_D double wg_sum(const WG& g, double v) {
#if CATN_DEV
  for (int o = 32; o > 0; o >>= 1) v += shfl_xor(v, o);
  const int w = g.tid >> 6, l = g.tid & 63;
  g.sync();
  if (l == 0) g.red[w] = v;
  g.sync();
  double s = 0.0;
  const int nw = (g.nthr + 63) >> 6;
  for (int i = 0; i < nw; ++i) s += g.red[i];
  return s;
#elif defined(CATN_EMU_MT)
  (void)g;
  return emu_wg_sum(v);
#else
  (void)g;
  return v;
#endif
}

// wg_sum with one barrier: the wavefront sums go to half `parity` of the
// reduction slots (16 per half, nthr <= 1024). The caller alternates the
// parity and keeps a workgroup barrier between the reads of one half and its
// next write.
CATN_D double wg_sum_db(const WG& g, double v, int parity) {
#if CATN_DEV
  for (int o = 32; o > 0; o >>= 1) v += shfl_xor(v, o);
  const int w = g.tid >> 6, l = g.tid & 63;
  double* r = g.red + 16 * (parity & 1);
  if (l == 0) r[w] = v;
  g.sync();
  double s = 0.0;
  const int nw = (g.nthr + 63) >> 6;
  for (int i = 0; i < nw; ++i) s += r[i];
  return s;
#else
  (void)parity;
  return wg_sum(g, v);
#endif
}

CATN_D double wg_max(const WG& g, double v) {
#if CATN_DEV
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, shfl_xor(v, o));
  const int w = g.tid >> 6, l = g.tid & 63;
  g.sync();
  if (l == 0) g.red[w] = v;
  g.sync();
  double s = g.red[0];
  const int nw = (g.nthr + 63) >> 6;
  for (int i = 1; i < nw; ++i) s = fmax(s, g.red[i]);
  return s;
#elif defined(CATN_EMU_MT)
  (void)g;
  return emu_wg_max(v);
#else
  (void)g;
  return v;
#endif
}

// Sum over an aligned group of L lanes (L a power of two <= 64). The xor
// butterfly gives every lane the bitwise-identical sum (a+b == b+a), so
// decisions taken on it are uniform across the group.
CATN_D double grp_sum(double v, int L) {
#if CATN_DEV
  for (int o = L >> 1; o > 0; o >>= 1) v += shfl_xor(v, o);
#elif defined(CATN_EMU_MT)
  v = emu_grp_sum(v, L);
#else
  (void)L;
#endif
  return v;
}

CATN_D int64_t now_ticks() {
#if defined(__HIP_DEVICE_COMPILE__)
  return (int64_t)wall_clock64();
#else
  return 0;
#endif
}

// add to a profiling counter without waiting for it: on the device a global
// atomic whose result is unused (no read-modify-write latency on the lane)
CATN_D void count_add(int64_t* p, int64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicAdd((unsigned long long*)p, (unsigned long long)v);
#else
  *p += v;
#endif
}
CATN_D void count_add(int32_t* p, int32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicAdd((int*)p, (int)v);
#else
  *p += v;
#endif
}

CATN_D void sub_tick(const WG& g, int k, int64_t t0) {
  if (g.tid == 0 && g.sub) count_add(g.sub + k, now_ticks() - t0);
}

CATN_D int size_bucket(int r) {
  int b = 0, lim = 4;
  while (b < 7 && r > lim) { ++b; lim *= 2; }
  return b;
}

CATN_D int pow2floor(int x) {
  int p = 1;
  while (p * 2 <= x) p *= 2;
  return p;
}

// ---------------------------------------------------------------------------
// scratch bump allocator (per op)
// ---------------------------------------------------------------------------
struct Scratch {
  double* base;
  int64_t cap, top;
  bool fail;
  CATN_D double* alloc(int64_t n) {
    int64_t a = (top + 31) & ~int64_t(31);
    if (a + n > cap) { fail = true; return base; }
    top = a + n;
    return base + a;
  }
};

// ---------------------------------------------------------------------------
// small dense kernels
// ---------------------------------------------------------------------------
CATN_D void wg_copy(const WG& g, double* dst, const double* src, int64_t n) {
  for (int64_t i = g.tid; i < n; i += g.nthr) dst[i] = src[i];
}

CATN_D void wg_fill(const WG& g, double* dst, double v, int64_t n) {
  for (int64_t i = g.tid; i < n; i += g.nthr) dst[i] = v;
}

// C (m x n, ldc) = A (m x k, lda) * B (k x n, ldb); column-major.
CATN_D void wg_gemm_nn(const WG& g, int m, int n, int k, const double* A, int64_t lda,
                       const double* B, int64_t ldb, double* C, int64_t ldc) {
  const int64_t tot = (int64_t)m * n;
  for (int64_t idx = g.tid; idx < tot; idx += g.nthr) {
    const int i = (int)(idx % m), j = (int)(idx / m);
    double s = 0.0;
    const double* a = A + i;
    const double* b = B + (int64_t)j * ldb;
    for (int p = 0; p < k; ++p) s += a[(int64_t)p * lda] * b[p];
    C[i + (int64_t)j * ldc] = s;
  }
}

// transpose: dst (n x m) = src (m x n)^T
CATN_D void wg_transpose(const WG& g, double* dst, const double* src, int m, int n) {
  const int64_t tot = (int64_t)m * n;
  for (int64_t idx = g.tid; idx < tot; idx += g.nthr) {
    const int j = (int)(idx % n), i = (int)(idx / n);  // dst[j + n*i] = src[i + m*j]
    dst[j + (int64_t)n * i] = src[i + (int64_t)m * j];
  }
}

// ---------------------------------------------------------------------------
// Householder QR (LAPACK dgeqr2 reflectors: H = I - tau v v^T, v(0) = 1)
// A (m x n, m >= n) overwritten by R (upper) and v (below the diagonal).
// ---------------------------------------------------------------------------
CATN_DN void wg_geqr2(const WG& g, double* A, int m, int n, int64_t lda, double* tau) {
  // Three barriers per column: the reflector v = [1; scal * x(1:)] is applied
  // with the scaling folded into the update, and the stored sub-diagonal of
  // column j is scaled during the norm pass of column j+1 (disjoint memory).
  // The diagonal entry beta of column j is stored in the same deferred pass:
  // every lane reads alpha = A(j, j) after the norm barrier of column j, and
  // an earlier store could overtake a lagging wavefront's read.
  int pend_col = -1;
  double pend_scal = 1.0, pend_beta = 0.0;
  for (int j = 0; j < n; ++j) {
    double* x = A + j + (int64_t)j * lda;
    double part = 0.0;
    for (int i = 1 + g.tid; i < m - j; i += g.nthr) part += x[i] * x[i];
    if (pend_col >= 0) {
      double* xp = A + pend_col + (int64_t)pend_col * lda;
      for (int i = 1 + g.tid; i < m - pend_col; i += g.nthr) xp[i] *= pend_scal;
      if (g.tid == 0) xp[0] = pend_beta;
    }
    // one barrier: the end-of-column barrier separates this column's reads of
    // the reduction slots from the next writes; column 0 writes the half that
    // no other reduction uses, so a caller's last wg_sum needs no barrier after
    const double s = wg_sum_db(g, part, (j + 1) & 1);
    const double alpha = x[0];
    double t = 0.0, beta = alpha, scal = 0.0;
    if (s > 0.0) {
      const double nrm = sqrt(alpha * alpha + s);
      beta = alpha >= 0.0 ? -nrm : nrm;
      t = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
      const int ncol = n - j - 1;
      if (ncol > 0) {
        int L = pow2floor(g.nthr / ncol < 1 ? 1 : g.nthr / ncol);
        if (L > 64) L = 64;
        const int ng = g.nthr / L, grp = g.tid / L, lane = g.tid % L;
        for (int c = grp; c < ncol; c += ng) {
          double* y = A + j + (int64_t)(j + 1 + c) * lda;
          const double y0 = y[0];   // read by the whole group before lane 0 updates it
          double w = 0.0;
          for (int i = 1 + lane; i < m - j; i += L) w += x[i] * y[i];
          w = grp_sum(w, L);
          w = t * (y0 + scal * w);
          if (lane == 0) y[0] = y0 - w;
          const double ws = w * scal;
          for (int i = 1 + lane; i < m - j; i += L) y[i] -= ws * x[i];
        }
      }
    }
    if (g.tid == 0) tau[j] = t;
    g.sync();
    pend_col = s > 0.0 ? j : -1;
    pend_scal = scal;
    pend_beta = beta;
  }
  if (pend_col >= 0) {
    double* xp = A + pend_col + (int64_t)pend_col * lda;
    for (int i = 1 + g.tid; i < m - pend_col; i += g.nthr) xp[i] *= pend_scal;
    if (g.tid == 0) xp[0] = pend_beta;
  }
  g.sync();
}

// U (m x k) := Q * U where Q = H_0 ... H_{n-1} from wg_geqr2 (stored in A);
// U is given as its top n x k block followed by zeros (rows n..m-1 set here).
// identity_in: U is [I; 0] with k = n (forming Q): H_j leaves the unit
// columns c < j alone (their nonzero is above row j), so only the columns
// c >= j are updated (LAPACK dorg2r).
CATN_DN void wg_apply_q(const WG& g, const double* A, int m, int n, int64_t lda,
                        const double* tau, double* U, int k, int64_t ldu,
                        bool identity_in = false) {
  for (int64_t idx = g.tid; idx < (int64_t)(m - n) * k; idx += g.nthr) {
    const int i = n + (int)(idx % (m - n)), c = (int)(idx / (m - n));
    U[i + (int64_t)c * ldu] = 0.0;
  }
  g.sync();
  int L = pow2floor(g.nthr / k < 1 ? 1 : g.nthr / k);
  if (L > 64) L = 64;
  const int ng = g.nthr / L, grp = g.tid / L, lane = g.tid % L;
  for (int j = n - 1; j >= 0; --j) {
    const double t = tau[j];
    if (t != 0.0) {
      const double* v = A + j + (int64_t)j * lda;   // v[0] implicit 1
      for (int c = (identity_in ? j : 0) + grp; c < k; c += ng) {
        double* y = U + j + (int64_t)c * ldu;
        double w = 0.0;
        for (int i = lane; i < m - j; i += L) w += (i == 0 ? 1.0 : v[i]) * y[i];
        w = grp_sum(w, L) * t;
        for (int i = lane; i < m - j; i += L) y[i] -= w * (i == 0 ? 1.0 : v[i]);
      }
    }
    g.sync();
  }
}

// ---------------------------------------------------------------------------
// dynamic LDS region of the interpreter workgroup. On the device the base is
// the kernel's dynamic shared array (so pointers derived from it compile to
// ds_* instructions after inlining); the CPU emulator uses a static buffer.
// ---------------------------------------------------------------------------
constexpr int kLdsMaxDoubles = 20480;   // 160 KB: the most a launch may choose
constexpr int kLdsMinDoubles = 2048;    // the least a launch shape may set

// Default LDS budget per workgroup (doubles). MI355X has 160 KB of LDS per CU:
// up to 512 workgroups get ~80 KB each (2 per CU on 256 CUs); larger batches
// get ~40 KB so that 4 workgroups co-reside per CU. The static part (reduction
// slots, flags) is kept out of the budget. A launch shape set on the context
// (catn_set_launch, picked by the autotuner) overrides it.
inline int lds_doubles_for_batch(int nb) { return nb <= 512 ? 10176 : 5056; }

#if defined(__HIP_DEVICE_COMPILE__)
CATN_D double* lds_base() {
  extern __shared__ double catn_dyn_lds[];
  return catn_dyn_lds;
}
#elif defined(__HIPCC__)
CATN_D double* lds_base() { return nullptr; }   // host pass of the HIP build: unused
#elif defined(CATN_EMU_MT)
inline double* lds_base() { return emu_lds_base(); }
#else
inline double* lds_base() {
  static thread_local double buf[kLdsMaxDoubles];
  return buf;
}
#endif

// ---------------------------------------------------------------------------
// LDS-tiled FP64 GEMM for one workgroup:
//   C[out(r, c)] = alpha * sum_k opA(r, k) * opB(k, c) + beta * C[out(r, c)]
// opA = A (M x K, lda) or A^T (A stored K x M); opB = B (K x N) or B^T.
// 64 x 64 output tiles, K chunks of 16 staged in LDS (16 KB at the top of the
// dynamic LDS region), 4 x 4 register micro-tile per lane (256 "virtual
// lanes" per tile; with fewer real lanes each one loops over several).
// ---------------------------------------------------------------------------
constexpr int kTM = 64, kTN = 64, kTK = 16;

struct IdxPlain {
  int64_t ldc;
  CATN_D int64_t operator()(int r, int c) const { return r + ldc * c; }
};

// Tile shape TM x TN, K chunk TK, micro-tile (TM/16) x (TN/16) per lane over a
// 16 x 16 grid of 256 virtual lanes. Square 64x64x16 for general products;
// 16x128x32 for the tall-skinny V^T C of the Householder updates (4x fewer
// barriers for long K).
template <int TM, int TN, int TK, class Out>
CATN_D void wg_gemm_tile(const WG& g, int M, int N, int K, double alpha, const double* A,
                         int64_t lda, bool tA, const double* B, int64_t ldb, bool tB, double beta,
                         double* C, Out out) {
  constexpr int MU = TM / 16, MV = TN / 16;
  double* As = lds_base() + (g.lds_cap - TK * (TM + TN));   // [TK][TM]
  double* Bs = As + TK * TM;                                 // [TK][TN]
  for (int r0 = 0; r0 < M; r0 += TM) {
    for (int c0 = 0; c0 < N; c0 += TN) {
      for (int vbase = 0; vbase < 256; vbase += g.nthr) {
        const int vt = vbase + g.tid;
        const int tx = vt & 15, ty = vt >> 4;
        double acc[MU][MV];
        for (int u = 0; u < MU; ++u)
          for (int v = 0; v < MV; ++v) acc[u][v] = 0.0;
        for (int k0 = 0; k0 < K; k0 += TK) {
          g.sync();
          // tile loads walk the operand's contiguous (leading) index across
          // consecutive lanes so that global reads coalesce
          for (int e = g.tid; e < TK * TM; e += g.nthr) {
            int kk, rr;
            if (tA) { kk = e % TK; rr = e / TK; }   // A[k + lda r]: k contiguous
            else { kk = e / TM; rr = e % TM; }      // A[r + lda k]: r contiguous
            const int r = r0 + rr, k = k0 + kk;
            double va = 0.0;
            if (r < M && k < K) va = tA ? A[k + lda * (int64_t)r] : A[r + lda * (int64_t)k];
            As[kk * TM + rr] = va;
          }
          for (int e = g.tid; e < TK * TN; e += g.nthr) {
            int kk, cc;
            if (tB) { kk = e / TN; cc = e % TN; }   // B[c + ldb k]: c contiguous
            else { kk = e % TK; cc = e / TK; }      // B[k + ldb c]: k contiguous
            const int c = c0 + cc, k = k0 + kk;
            double vb = 0.0;
            if (c < N && k < K) vb = tB ? B[c + ldb * (int64_t)k] : B[k + ldb * (int64_t)c];
            Bs[kk * TN + cc] = vb;
          }
          g.sync();
          if (vt < 256) {
            const int kmax = (K - k0) < TK ? (K - k0) : TK;
            for (int kk = 0; kk < kmax; ++kk) {
              double a[MU], b[MV];
              for (int u = 0; u < MU; ++u) a[u] = As[kk * TM + ty + 16 * u];
              for (int v = 0; v < MV; ++v) b[v] = Bs[kk * TN + tx + 16 * v];
              for (int u = 0; u < MU; ++u)
                for (int v = 0; v < MV; ++v) acc[u][v] += a[u] * b[v];
            }
          }
        }
        if (vt < 256) {
          for (int u = 0; u < MU; ++u) {
            const int r = r0 + ty + 16 * u;
            if (r >= M) continue;
            for (int v = 0; v < MV; ++v) {
              const int c = c0 + tx + 16 * v;
              if (c >= N) continue;
              const int64_t o = out(r, c);
              C[o] = beta == 0.0 ? alpha * acc[u][v] : alpha * acc[u][v] + beta * C[o];
            }
          }
        }
      }
    }
  }
  g.sync();
}

#if CATN_DEV
// The same product on the FP64 matrix cores (v_mfma_f64_16x16x4f64; lane l
// holds A(l % 16, l / 16), B(l / 16, l % 16) and D(l / 16 + 4 v, l % 16) in
// accumulator element v, checked on gfx950): one wavefront per 16 x 16 output
// tile, K in steps of 4 with operands loaded straight from memory, so there is
// no LDS staging and no workgroup barrier inside. Loop bounds are
// wave-uniform; out-of-range operand entries are loaded as 0.
template <class Out>
CATN_D void wg_gemm_mfma(const WG& g, int M, int N, int K, double alpha, const double* A,
                         int64_t lda, bool tA, const double* B, int64_t ldb, bool tB, double beta,
                         double* C, Out out) {
  const int nw = g.nthr >> 6, wv = g.tid >> 6, l = g.tid & 63;
  const int li = l & 15, lk = l >> 4;
  const int rt = (M + 15) >> 4, ct = (N + 15) >> 4;
  for (int t = wv; t < rt * ct; t += nw) {
    const int r0 = (t % rt) * 16, c0 = (t / rt) * 16;
    const int ra = r0 + li, cb = c0 + li;
    const bool rok = ra < M, cok = cb < N;
    const double* ap = tA ? A + lda * (int64_t)(rok ? ra : 0) : A + (rok ? ra : 0);
    const double* bp = tB ? B + (cok ? cb : 0) : B + ldb * (int64_t)(cok ? cb : 0);
    const int64_t sa = tA ? 1 : lda, sb = tB ? ldb : 1;   // strides along k
    catn_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k0 = 0; k0 < K; k0 += 4) {
      const int k = k0 + lk;
      const double a = (rok && k < K) ? ap[sa * k] : 0.0;
      const double b = (cok && k < K) ? bp[sb * k] : 0.0;
      acc = mfma_f64_16x16x4(a, b, acc);
    }
    for (int v = 0; v < 4; ++v) {
      const int r = r0 + lk + 4 * v;
      if (r < M && cok) {
        const int64_t o = out(r, cb);
        C[o] = beta == 0.0 ? alpha * acc[v] : alpha * acc[v] + beta * C[o];
      }
    }
  }
  g.sync();
}
#endif

// general product: FP64 matrix cores on the device, the LDS-tiled FMA GEMM on
// the host
template <class Out>
CATN_D void wg_gemm_t(const WG& g, int M, int N, int K, double alpha, const double* A, int64_t lda,
                      bool tA, const double* B, int64_t ldb, bool tB, double beta, double* C,
                      Out out) {
#if CATN_DEV
  wg_gemm_mfma(g, M, N, K, alpha, A, lda, tA, B, ldb, tB, beta, C, out);
#else
  wg_gemm_tile<kTM, kTN, kTK>(g, M, N, K, alpha, A, lda, tA, B, ldb, tB, beta, C, out);
#endif
}

CATN_DN void wg_gemm(const WG& g, int M, int N, int K, double alpha, const double* A, int64_t lda,
                     bool tA, const double* B, int64_t ldb, bool tB, double beta, double* C,
                     int64_t ldc) {
  IdxPlain o;
  o.ldc = ldc;
  wg_gemm_t(g, M, N, K, alpha, A, lda, tA, B, ldb, tB, beta, C, o);
}

// C = A * B (no transposes) through the tiled GEMM when the product is large
// enough to amortise its barriers, else the direct per-element loop.
CATN_D void wg_mm(const WG& g, int M, int N, int K, const double* A, int64_t lda,
                  const double* B, int64_t ldb, double* C, int64_t ldc) {
  if ((int64_t)M * N >= 2048 && K >= 8 && g.lds_cap >= 2 * kTK * kTM + 64)
    wg_gemm(g, M, N, K, 1.0, A, lda, false, B, ldb, false, 0.0, C, ldc);
  else
    wg_gemm_nn(g, M, N, K, A, lda, B, ldb, C, ldc);
}

// C = alpha * op(A) op(B) + beta * C; tiled GEMM for products large enough to
// amortise its barriers, else a direct loop (every lane owns whole outputs).
CATN_DN void wg_gemm_any(const WG& g, int M, int N, int K, double alpha, const double* A,
                         int64_t lda, bool tA, const double* B, int64_t ldb, bool tB, double beta,
                         double* C, int64_t ldc) {
  if (M <= 0 || N <= 0) return;
  if ((int64_t)M * N >= 2048 && K >= 8 && g.lds_cap >= 2 * kTK * kTM + 64) {
    wg_gemm(g, M, N, K, alpha, A, lda, tA, B, ldb, tB, beta, C, ldc);
    return;
  }
  const int64_t tot = (int64_t)M * N;
  for (int64_t idx = g.tid; idx < tot; idx += g.nthr) {
    const int i = (int)(idx % M), j = (int)(idx / M);
    double s = 0.0;
    for (int p = 0; p < K; ++p) {
      const double a = tA ? A[p + lda * (int64_t)i] : A[i + lda * (int64_t)p];
      const double b = tB ? B[j + ldb * (int64_t)p] : B[p + ldb * (int64_t)j];
      s += a * b;
    }
    double* c = C + i + ldc * (int64_t)j;
    *c = beta == 0.0 ? alpha * s : alpha * s + beta * *c;
  }
  g.sync();
}

// ---------------------------------------------------------------------------
// Split-BF16 GEMM for FP32 operands: C[out(r, c)] = sum_k A(r, k) B(k, c) with
// A (M x K, lda) and B (K x N, ldb) in FP32, products on the BF16 matrix
// cores, accumulated in FP32, widened to FP64. Each operand is split into
// BF16 terms by successive round-to-nearest residuals:
//   NT = 2: x = hi + lo (to 2^-17 relative); hi*hi + hi*lo + lo*hi, 3 passes;
//   NT = 3: x = hi + mid + lo (to FP32 precision); the six products whose
//           order is at most 2^-16 relative, 6 passes.
// One wavefront per 16 x 16 output tile, v_mfma_f32_16x16x16_bf16 per
// 16 x 16 x 16 block and pass.
// ---------------------------------------------------------------------------
CATN_D uint32_t f32_bits(float x) {
  union { float f; uint32_t u; } v;
  v.f = x;
  return v.u;
}
CATN_D float f32_from_bits(uint32_t u) {
  union { float f; uint32_t u; } v;
  v.u = u;
  return v.f;
}
// round to nearest even BF16 (the operands are finite)
CATN_D uint16_t bf16_rne(float x) {
  const uint32_t u = f32_bits(x);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
CATN_D float bf16_val(uint16_t h) { return f32_from_bits((uint32_t)h << 16); }
template <int NT>
CATN_D void bf16_split(float x, uint16_t* t) {
  float r = x;
  for (int q = 0; q < NT; ++q) {
    t[q] = bf16_rne(r);
    r -= bf16_val(t[q]);
  }
}
// the (a-term, b-term) pairs of each pass
template <int NT>
struct Bf16Passes;
template <>
struct Bf16Passes<2> {
  static constexpr int n = 3;
  CATN_D static int a(int p) { return p == 2 ? 1 : 0; }
  CATN_D static int b(int p) { return p == 1 ? 1 : 0; }
};
template <>
struct Bf16Passes<3> {
  static constexpr int n = 6;   // (0,0) (0,1) (1,0) (0,2) (1,1) (2,0)
  CATN_D static int a(int p) { return (0x210100 >> (4 * p)) & 15; }
  CATN_D static int b(int p) { return (0x012010 >> (4 * p)) & 15; }
};

template <int NT, class Out>
CATN_DN void wg_gemm_bf16(const WG& g, int M, int N, int K, const float* A, int64_t lda,
                          const float* B, int64_t ldb, double* C, Out out) {
  typedef Bf16Passes<NT> P;
#if CATN_DEV
  // 16 x 16 x 16 MFMA operand layout: lane l holds A(l % 16, 4 (l / 16) + v),
  // B(4 (l / 16) + v, l % 16) and result C(4 (l / 16) + v, l % 16), v = 0..3
  const int nw = g.nthr >> 6, w = g.tid >> 6, l = g.tid & 63;
  const int tm = (M + 15) >> 4, tn = (N + 15) >> 4;
  const int lr = l & 15, lk = (l >> 4) << 2;
  for (int t = w; t < tm * tn; t += nw) {
    const int i0 = (t % tm) << 4, j0 = (t / tm) << 4;
    catn_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 16) {
      catn_bf16x4 at[NT], bt[NT];
      for (int v = 0; v < 4; ++v) {
        const int k = k0 + lk + v;
        const int i = i0 + lr, j = j0 + lr;
        const float a = (i < M && k < K) ? A[i + lda * (int64_t)k] : 0.0f;
        const float b = (k < K && j < N) ? B[k + ldb * (int64_t)j] : 0.0f;
        uint16_t ta[NT], tb[NT];
        bf16_split<NT>(a, ta);
        bf16_split<NT>(b, tb);
        for (int q = 0; q < NT; ++q) {
          at[q][v] = (short)ta[q];
          bt[q][v] = (short)tb[q];
        }
      }
      for (int p = 0; p < P::n; ++p)
        acc = mfma_bf16_16x16x16(at[P::a(p)], bt[P::b(p)], acc);
    }
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + lk + v, j = j0 + lr;
      if (i < M && j < N) C[out(i, j)] = (double)acc[v];
    }
  }
#else
  // the same products in FP32, k in order (the emulator's rendering)
  for (int64_t idx = g.tid; idx < (int64_t)M * N; idx += g.nthr) {
    const int i = (int)(idx % M), j = (int)(idx / M);
    float acc = 0.0f;
    for (int k = 0; k < K; ++k) {
      uint16_t ta[NT], tb[NT];
      bf16_split<NT>(A[i + lda * (int64_t)k], ta);
      bf16_split<NT>(B[k + ldb * (int64_t)j], tb);
      for (int p = 0; p < P::n; ++p) acc += bf16_val(ta[P::a(p)]) * bf16_val(tb[P::b(p)]);
    }
    C[out(i, j)] = (double)acc;
  }
#endif
  g.sync();
}

// output index of the swap theta: X[(i + d0 j), (a + d1 b)] -> M[(i + d0 a), (j + d2 b)]
struct IdxTheta {
  int d0, d1, d2;
  int64_t m;
  CATN_D int64_t operator()(int r, int c) const {
    const int i = r % d0, j = r / d0, a = c % d1, b = c / d1;
    return (int64_t)(i + d0 * a) + m * (int64_t)(j + d2 * b);
  }
};

// ---------------------------------------------------------------------------
// Blocked Householder QR (LAPACK dgeqrf structure): panels of g.qb columns are
// factorised by wg_geqr2; the trailing matrix is updated with the compact WY
// form Q_p^T C = C - V (T^T (V^T C)) through two skinny products, so the trailing
// matrix is streamed twice per panel instead of twice per column. The panel
// T factors are kept (T_all, kQBMax x n) for wg_apply_q_blocked.
// ---------------------------------------------------------------------------

#if CATN_DEV
// FP64 matrix-core forms of the two products of a compact-WY update, for
// v_mfma_f64_16x16x4f64: a wavefront computes a 16 x 16 tile D += A B with
// A 16 x 4 and B 4 x 16; lane l holds A(l % 16, l / 16), B(l / 16, l % 16) and
// D(l / 16 + 4 v, l % 16) in accumulator element v (checked on gfx950). Every loop below has
// wave-uniform bounds (the matrix instruction runs for the whole wavefront);
// out-of-range operand entries are loaded as 0.

// entry (k, q) of a reflector panel: the stored value, or, for a unit-lower
// panel read in place from the factored matrix, 1 on the diagonal and 0 above
CATN_D double refl_at(const double* V, int64_t ldv, bool unit, int k, int q) {
  if (unit && k <= q) return k == q ? 1.0 : 0.0;
  return V[k + ldv * (int64_t)q];
}

// W (jb x N, ldw) = V^T C, V rows x jb (ldv), C rows x N (ldc), either of them
// a unit-lower reflector panel read in place (vunit / cunit): one wavefront
// per 16 x 16 tile of W walking the rows 4 at a time, so no workgroup barrier
// is needed inside. When the tiles are fewer than the wavefronts the rows are
// split among them and the partial tiles are summed through LDS (one per
// wavefront up to 4 partial tiles: the top 256 min(wavefronts, 4) doubles of
// the free LDS region, which the callers leave free).
// With T given (jb <= 16, upper triangular, ld kQBMax) and the rows not split,
// W = op(T) V^T C is formed in registers (op(T) = T^T when tT): the
// accumulator layout of a tile (lane l, element v: row l/16 + 4v, column l%16)
// is the B-operand layout of K-step v of the product with op(T). The return
// value says whether T was applied.
CATN_D bool wg_vtc_mfma(const WG& g, int jb, int N, int rows, const double* V, int64_t ldv,
                        bool vunit, const double* C, int64_t ldc, bool cunit, double* W,
                        int64_t ldw, const double* T = nullptr, bool tT = false) {
  const int nw = g.nthr >> 6, wv = g.tid >> 6, l = g.tid & 63;
  const int li = l & 15, lk = l >> 4;
  const int qt = (jb + 15) >> 4, ct = (N + 15) >> 4, tasks = qt * ct;
  constexpr int kPartials = 4;
  int ks = tasks >= nw ? 1 : nw / tasks;
  if (ks > 1 && tasks * ks > kPartials) ks = kPartials / tasks > 1 ? kPartials / tasks : 1;
  const bool fuse = T != nullptr && ks == 1 && jb <= 16;
  const int chunk = ((rows + ks - 1) / ks + 3) & ~3;
  double* P = lds_base() + (g.lds_cap - 256 * (nw < kPartials ? nw : kPartials));
  for (int t = wv; t < tasks * ks; t += nw) {
    const int tile = t % tasks, part = t / tasks;
    const int q0 = (tile % qt) * 16, c0 = (tile / qt) * 16;
    const int r0 = part * chunk;
    const int r1 = (r0 + chunk) < rows ? (r0 + chunk) : rows;
    const int q = q0 + li, c = c0 + li;
    const bool qok = q < jb, cok = c < N;
    const int qs = qok ? q : 0, cs = cok ? c : 0;
    catn_d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = r0; k0 < r1; k0 += 4) {
      const int k = k0 + lk;
      const double a = (qok && k < r1) ? refl_at(V, ldv, vunit, k, qs) : 0.0;
      const double b = (cok && k < r1) ? refl_at(C, ldc, cunit, k, cs) : 0.0;
      acc = mfma_f64_16x16x4(a, b, acc);
    }
    if (fuse) {
      catn_d4 o = {0.0, 0.0, 0.0, 0.0};
      for (int st = 0; st < 4; ++st) {
        const int kk = 4 * st + lk;
        double a = 0.0;
        if (li < jb && kk < jb) {
          if (tT) a = kk <= li ? T[kk + kQBMax * li] : 0.0;   // T^T(li, kk) = T(kk, li)
          else a = li <= kk ? T[li + kQBMax * kk] : 0.0;
        }
        o = mfma_f64_16x16x4(a, acc[st], o);
      }
      acc = o;
    }
    if (ks == 1) {
      for (int v = 0; v < 4; ++v) {
        const int qq = q0 + lk + 4 * v;
        if (qq < jb && cok) W[qq + ldw * (int64_t)c] = acc[v];
      }
    } else {
      for (int v = 0; v < 4; ++v) P[256 * t + 16 * (lk + 4 * v) + li] = acc[v];
    }
  }
  if (ks > 1) {
    g.sync();
    for (int e = g.tid; e < tasks * 256; e += g.nthr) {
      const int tile = e >> 8, rr = (e >> 4) & 15, cc = e & 15;
      const int qq = (tile % qt) * 16 + rr, c = (tile / qt) * 16 + cc;
      double s = 0.0;
      for (int part = 0; part < ks; ++part) s += P[256 * (part * tasks + tile) + 16 * rr + cc];
      if (qq < jb && c < N) W[qq + ldw * (int64_t)c] = s;
    }
  }
  g.sync();
  return fuse;
}

// C (rows x N, ldc) -= V W2, V rows x jb (ldv; unit-lower read in place when
// vunit), W2 jb x N (ldw), jb <= kQBMax: one wavefront per 16 x 16 tile of C,
// jb / 4 matrix instructions per tile accumulating onto the loaded tile with
// -V as the A operand.
CATN_D void wg_c_minus_vw_mfma(const WG& g, int rows, int N, int jb, const double* V,
                               int64_t ldv, bool vunit, const double* W2, int64_t ldw,
                               double* C, int64_t ldc) {
  constexpr int kS = kQBMax / 4;
  const int nw = g.nthr >> 6, wv = g.tid >> 6, l = g.tid & 63;
  const int li = l & 15, lk = l >> 4;
  const int rt = (rows + 15) >> 4, ct = (N + 15) >> 4;
  const int steps = (jb + 3) >> 2;
  for (int t = wv; t < rt * ct; t += nw) {
    const int r0 = (t % rt) * 16, c0 = (t / rt) * 16;
    const int ra = r0 + li, c = c0 + li;
    catn_d4 acc;
    for (int v = 0; v < 4; ++v) {
      const int r = r0 + lk + 4 * v;
      acc[v] = (r < rows && c < N) ? C[r + ldc * (int64_t)c] : 0.0;
    }
    for (int st = 0; st < kS; ++st) {
      if (st >= steps) break;
      const int k = 4 * st + lk;
      const double a = (ra < rows && k < jb) ? -refl_at(V, ldv, vunit, ra, k) : 0.0;
      const double b = (k < jb && c < N) ? W2[k + ldw * (int64_t)c] : 0.0;
      acc = mfma_f64_16x16x4(a, b, acc);
    }
    for (int v = 0; v < 4; ++v) {
      const int r = r0 + lk + 4 * v;
      if (r < rows && c < N) C[r + ldc * (int64_t)c] = acc[v];
    }
  }
  g.sync();
}
#endif

// C (jb x N) = V^T B: FP64 matrix cores on the device; on the host the skinny
// LDS tile (budget permitting)
CATN_DN void wg_gemm_skinny_tn(const WG& g, int jb, int N, int rows, const double* V, int64_t ldv,
                               const double* B, int64_t ldb, double* C, int64_t ldc) {
#if CATN_DEV
  wg_vtc_mfma(g, jb, N, rows, V, ldv, false, B, ldb, false, C, ldc);
#else
  IdxPlain o;
  o.ldc = ldc;
  if (g.lds_cap >= 32 * (16 + 128) + 64)
    wg_gemm_tile<16, 128, 32>(g, jb, N, rows, 1.0, V, ldv, true, B, ldb, false, 0.0, C, o);
  else
    wg_gemm_tile<kTM, kTN, kTK>(g, jb, N, rows, 1.0, V, ldv, true, B, ldb, false, 0.0, C, o);
#endif
}

// C (rows x N) -= V W2 with V rows x jb (jb <= kQBMax): FP64 matrix cores on
// the device, the tiled GEMM on the host
CATN_DN void wg_update_cvw(const WG& g, int rows, int N, int jb, const double* V, int64_t ldv,
                           const double* W2, int64_t ldw, double* C, int64_t ldc) {
#if CATN_DEV
  wg_c_minus_vw_mfma(g, rows, N, jb, V, ldv, false, W2, ldw, C, ldc);
#else
  wg_gemm(g, rows, N, jb, -1.0, V, ldv, false, W2, ldw, false, 1.0, C, ldc);
#endif
}

// T (jb x jb, upper) of H_0..H_{jb-1} = I - V T V^T from the unit-lower V
// (rows x jb, ld ldv; read in place from the factored panel when vunit) and
// tau (LAPACK dlarft, forward columnwise).
CATN_D void wg_larft(const WG& g, const double* V, int64_t ldv, bool vunit, int rows, int jb,
                     const double* tau, double* T, double* Y) {
  // Y = V^T V (jb x jb)
#if CATN_DEV
  wg_vtc_mfma(g, jb, jb, rows, V, ldv, vunit, V, ldv, vunit, Y, jb);
#else
  (void)vunit;
  wg_gemm_skinny_tn(g, jb, jb, rows, V, ldv, V, ldv, Y, jb);
#endif
  // T(j, i) = -tau_i sum_{l=j}^{i-1} T(j, l) Y(l, i): row j of T needs only its
  // own earlier entries, so lane j builds row j in registers.
  for (int j = g.tid; j < jb; j += g.nthr) {
    double row[kQBMax];
    for (int i = 0; i < kQBMax; ++i) row[i] = 0.0;
    row[j] = tau[j];
    for (int i = j + 1; i < jb; ++i) {
      double acc = 0.0;
      for (int l = j; l < i; ++l) acc += row[l] * Y[l + jb * i];
      row[i] = -tau[i] * acc;
    }
    for (int i = 0; i < jb; ++i) T[j + kQBMax * i] = row[i];
  }
  g.sync();
}

// W2 (jb x N, ld jb) = op(T) W (jb x N, ld jb); T upper triangular, ld kQBMax;
// op(T) = T^T when trans.
CATN_D void wg_gemm_nn_t(const WG& g, int jb, int N, const double* T, const double* W, double* W2,
                         bool trans) {
  for (int64_t idx = g.tid; idx < (int64_t)jb * N; idx += g.nthr) {
    const int r = (int)(idx % jb), c = (int)(idx / jb);
    double s = 0.0;
    if (trans) {
      for (int q = 0; q <= r; ++q) s += T[q + kQBMax * r] * W[q + (int64_t)jb * c];
    } else {
      for (int q = r; q < jb; ++q) s += T[r + kQBMax * q] * W[q + (int64_t)jb * c];
    }
    W2[idx] = s;
  }
  g.sync();
}

// C (rows x N, ldc) := (I - V op(T) V^T) C with op(T) = T^T when tT (the
// trailing update of wg_geqrf, Q_p^T C) or T (wg_apply_q_blocked, Q_p C).
// W, W2: jb x N scratch. On the device V is the factored panel read in place
// (vunit) and op(T) is folded into the V^T C tiles when they are not split;
// on the host V is the materialised unit-lower copy.
CATN_DN void wg_wy_apply(const WG& g, int rows, int N, int jb, const double* V, int64_t ldv,
                         bool vunit, const double* T, bool tT, double* C, int64_t ldc, double* W,
                         double* W2) {
#if CATN_DEV
  const bool fuse = jb <= 16 && ((N + 15) >> 4) >= (g.nthr >> 6);
  if (fuse) {
    wg_vtc_mfma(g, jb, N, rows, V, ldv, vunit, C, ldc, false, W2, jb, T, tT);
  } else {
    wg_vtc_mfma(g, jb, N, rows, V, ldv, vunit, C, ldc, false, W, jb);
    wg_gemm_nn_t(g, jb, N, T, W, W2, tT);
  }
  wg_c_minus_vw_mfma(g, rows, N, jb, V, ldv, vunit, W2, jb, C, ldc);
#else
  (void)vunit;
  wg_gemm_skinny_tn(g, jb, N, rows, V, ldv, C, ldc, W, jb);
  wg_gemm_nn_t(g, jb, N, T, W, W2, tT);
  wg_update_cvw(g, rows, N, jb, V, ldv, W2, jb, C, ldc);
#endif
}

// unit-lower copy of the panel's Householder vectors
CATN_D void wg_panel_v(const WG& g, const double* A, int rows, int jb, int64_t lda, double* V) {
  for (int64_t idx = g.tid; idx < (int64_t)rows * jb; idx += g.nthr) {
    const int i = (int)(idx % rows), j = (int)(idx / rows);
    V[idx] = i > j ? A[i + lda * j] : (i == j ? 1.0 : 0.0);
  }
  g.sync();
}

// wg_geqr2 of an operand in global memory. When it fits the LDS budget it is
// factorised in LDS (the 3 barriers per column then wait on LDS, not on
// global memory) and copied back. The LDS must be free: every caller of
// wg_geqrf keeps its operands in scratch.
CATN_DN void wg_geqr2_staged(const WG& g, double* A, int m, int n, int64_t lda, double* tau) {
  if ((int64_t)m * n > g.lds_cap) {
    wg_geqr2(g, A, m, n, lda, tau);
    return;
  }
  double* P = lds_base();
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr)
    P[idx] = A[(idx % m) + lda * (idx / m)];
  g.sync();
  wg_geqr2(g, P, m, n, m, tau);
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr)
    A[(idx % m) + lda * (idx / m)] = P[idx];
  g.sync();
}

CATN_DN bool wg_geqrf(const WG& g, Scratch& sc, double* A, int m, int n, int64_t lda, double* tau,
                      double* Tall) {
  const int qb = g.qb;
  if (n <= qb) {
    wg_geqr2_staged(g, A, m, n, lda, tau);
    return false;   // unblocked: no T factors
  }
#if !CATN_DEV
  double* V = sc.alloc((int64_t)m * qb);   // unit-lower panel copy (device: read in place)
#endif
  double* W = sc.alloc((int64_t)qb * n);
  double* W2 = sc.alloc((int64_t)qb * n);
  double* Y = sc.alloc(qb * qb);
  if (sc.fail) return false;
  for (int p = 0; p < n; p += qb) {
    const int jb = (n - p) < qb ? (n - p) : qb;
    const int rows = m - p;
    double* Ap = A + p + lda * (int64_t)p;
    int64_t t0 = now_ticks();
    wg_geqr2_staged(g, Ap, rows, jb, lda, tau + p);
    sub_tick(g, 8, t0);
    t0 = now_ticks();
#if CATN_DEV
    const double* Vp = Ap;   // reflectors read in place (unit lower)
    const int64_t ldv = lda;
    const bool vunit = true;
#else
    wg_panel_v(g, Ap, rows, jb, lda, V);
    const double* Vp = V;
    const int64_t ldv = rows;
    const bool vunit = false;
#endif
    double* T = Tall + (int64_t)kQBMax * p;   // jb x jb block, ld kQBMax
    wg_larft(g, Vp, ldv, vunit, rows, jb, tau + p, T, Y);
    const int nc = n - p - jb;
    // C := Q_p^T C = C - V T^T (V^T C)
    if (nc > 0) wg_wy_apply(g, rows, nc, jb, Vp, ldv, vunit, T, true, Ap + lda * (int64_t)jb, lda, W, W2);
    sub_tick(g, 9, t0);
  }
  return true;
}

// X (m x k, ld ldx) := Q X with Q from wg_geqrf (blocked: Tall given) or
// wg_geqr2 (Tall null); X's rows n..m-1 must be set by the caller.
CATN_DN void wg_apply_q_blocked(const WG& g, Scratch& sc, const double* A, int m, int n,
                                int64_t lda, const double* tau, const double* Tall, double* X,
                                int k, int64_t ldx, bool identity_in = false) {
  if (!Tall) {
    wg_apply_q(g, A, m, n, lda, tau, X, k, ldx, identity_in);
    return;
  }
  const int64_t save = sc.top;
  const int qb = g.qb;   // the panel width wg_geqrf used (same launch)
#if !CATN_DEV
  double* V = sc.alloc((int64_t)m * qb);   // unit-lower panel copy (device: read in place)
#endif
  double* W = sc.alloc((int64_t)qb * k);
  double* W2 = sc.alloc((int64_t)qb * k);
  if (sc.fail) return;
  const int64_t t0 = now_ticks();
  const int last = ((n - 1) / qb) * qb;
  for (int p = last; p >= 0; p -= qb) {
    const int jb = (n - p) < qb ? (n - p) : qb;
    const int rows = m - p;
#if CATN_DEV
    const double* Vp = A + p + lda * (int64_t)p;   // reflectors read in place (unit lower)
    const int64_t ldv = lda;
    const bool vunit = true;
#else
    wg_panel_v(g, A + p + lda * (int64_t)p, rows, jb, lda, V);
    const double* Vp = V;
    const int64_t ldv = rows;
    const bool vunit = false;
#endif
    const double* T = Tall + (int64_t)kQBMax * p;
    // Q_p X = X - V (T (V^T X)); from [I; 0] the columns left of the panel
    // are still unit vectors above its rows, so only X(p:, p:) changes
    // (LAPACK dorgqr)
    if (identity_in)
      wg_wy_apply(g, rows, k - p, jb, Vp, ldv, vunit, T, false, X + p + ldx * (int64_t)p, ldx, W, W2);
    else
      wg_wy_apply(g, rows, k, jb, Vp, ldv, vunit, T, false, X + p, ldx, W, W2);
  }
  sub_tick(g, 10, t0);
  sc.top = save;
}

// ---------------------------------------------------------------------------
// one-sided Jacobi SVD: A (m x n, m >= n) -> A V with mutually orthogonal
// columns; V (n x n) accumulates the rotations. Returns sweeps used, or -1
// when not converged within maxsweeps. Round-robin ordering: n/2 disjoint
// column pairs per step, one aligned group of L lanes per pair.
// ---------------------------------------------------------------------------
// lanes_per_pair = 0 spreads all lanes over the n/2 pairs of a round (short
// LDS-resident columns); a wavefront-sized value gives each pair one
// wavefront that walks its rows contiguously, which global-memory operands
// need: coalesced loads and full-width reductions, each wavefront taking its
// share of the round's pairs in turn.
// rows per lane held in registers by the Jacobi pair step (m, n <= kJacReg * L)
constexpr int kJacReg = 8;

CATN_D int jacobi_core(const WG& g, double* A, int m, int n, int64_t lda, double* V,
                       int64_t ldv, int maxsweeps, double thr_in = -1.0, int lanes_per_pair = 0) {
  for (int64_t idx = g.tid; idx < (int64_t)n * n; idx += g.nthr) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    V[i + (int64_t)j * ldv] = (i == j) ? 1.0 : 0.0;
  }
  g.sync();
  if (n < 2) return 0;
  const int N = n + (n & 1);
  const int npairs = N / 2;
  int L = g.nthr / npairs;
  L = L < 1 ? 1 : pow2floor(L);
  if (lanes_per_pair > L) L = pow2floor(lanes_per_pair < g.nthr ? lanes_per_pair : g.nthr);
  if (L > 64) L = 64;
  const int ng = g.nthr / L, grp = g.tid / L, lane = g.tid % L;
  // pair convergence test on the cosine |a_p.a_q| / (|a_p||a_q|); a rotation
  // below that level is below working precision and can repeat forever
  const double tol = fmax(4.0 * 2.220446049250313e-16 * sqrt((double)(m > 1 ? m : 1)), g.jac_tol);
  // columns below 1e-18 ||A||_F are rounding noise of a rank-deficient input:
  // they cannot be made orthogonal (the column space is full) and only shrink
  // towards denormals, so they are treated as zero columns
  double thr = thr_in;
  if (thr < 0.0) {
    double fpart = 0.0;
    for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
      const double v = A[(idx % m) + (idx / m) * lda];
      fpart += v * v;
    }
    thr = 1e-36 * wg_sum(g, fpart);
  }
  for (int sweep = 1; sweep <= maxsweeps; ++sweep) {
    int rotated = 0;
    for (int r = 0; r < N - 1; ++r) {
      for (int k = grp; k < npairs; k += ng) {
        int p, q;
        if (k == 0) { p = r; q = N - 1; }
        else { p = (r + k) % (N - 1); q = (r - k + N - 1) % (N - 1); }
        if (p > q) { int t = p; p = q; q = t; }
        if (q >= n) continue;
        double* ap = A + (int64_t)p * lda;
        double* aq = A + (int64_t)q * lda;
        double* vp = V + (int64_t)p * ldv;
        double* vq = V + (int64_t)q * ldv;
        double al = 0.0, be = 0.0, ga = 0.0;
        if (m <= kJacReg * L && n <= kJacReg * L) {
          // the pair's columns (and V's) stay in registers: one round of
          // loads in flight, rotation from registers, one round of stores
          // (named scalars: indexed arrays were placed on the stack)
#define CATN_JR8(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)
#define CATN_JLOAD(u)                                  \
  const int i##u = lane + u * L;                       \
  const double xa##u = i##u < m ? ap[i##u] : 0.0;      \
  const double ya##u = i##u < m ? aq[i##u] : 0.0;      \
  const double xv##u = i##u < n ? vp[i##u] : 0.0;      \
  const double yv##u = i##u < n ? vq[i##u] : 0.0;
#define CATN_JDOT(u) al += xa##u * xa##u; be += ya##u * ya##u; ga += xa##u * ya##u;
#define CATN_JROT(u)                                   \
  if (i##u < m) {                                      \
    ap[i##u] = c * xa##u - s * ya##u;                  \
    aq[i##u] = s * xa##u + c * ya##u;                  \
  }                                                    \
  if (i##u < n) {                                      \
    vp[i##u] = c * xv##u - s * yv##u;                  \
    vq[i##u] = s * xv##u + c * yv##u;                  \
  }
          CATN_JR8(CATN_JLOAD)
          CATN_JR8(CATN_JDOT)
          al = grp_sum(al, L); be = grp_sum(be, L); ga = grp_sum(ga, L);
          if (al <= thr || be <= thr) continue;
          if (fabs(ga) <= tol * sqrt(al) * sqrt(be)) continue;
          rotated = 1;
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
          CATN_JR8(CATN_JROT)
#undef CATN_JROT
#undef CATN_JDOT
#undef CATN_JLOAD
#undef CATN_JR8
          continue;
        }
        // streamed pair (rows beyond the registers): unrolled by 4 so that
        // four rows' loads are in flight per lane (the sums keep their order)
#pragma unroll 4
        for (int i = lane; i < m; i += L) {
          const double x = ap[i], y = aq[i];
          al += x * x; be += y * y; ga += x * y;
        }
        al = grp_sum(al, L); be = grp_sum(be, L); ga = grp_sum(ga, L);
        if (al <= thr || be <= thr) continue;
        if (fabs(ga) <= tol * sqrt(al) * sqrt(be)) continue;
        rotated = 1;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
#pragma unroll 4
        for (int i = lane; i < m; i += L) {
          const double x = ap[i], y = aq[i];
          ap[i] = c * x - s * y;
          aq[i] = s * x + c * y;
        }
#pragma unroll 4
        for (int i = lane; i < n; i += L) {
          const double x = vp[i], y = vq[i];
          vp[i] = c * x - s * y;
          vq[i] = s * x + c * y;
        }
      }
      g.sync();
    }
    if (!(wg_max(g, rotated ? 1.0 : 0.0) > 0.0)) return sweep;
  }
  return -1;
}

// Jacobi on an operand in global memory: a group of g.jac_lanes lanes per
// column pair (default 16: a wavefront rotates four pairs at once, each lane
// holding up to kJacReg rows of both columns in registers for n <= 128);
// consecutive lanes of a group read consecutive doubles.

CATN_DN int wg_jacobi(const WG& g, double* A, int m, int n, int64_t lda, double* V,
                      int64_t ldv, int maxsweeps) {
  const int sw = jacobi_core(g, A, m, n, lda, V, ldv, maxsweeps, -1.0, g.jac_lanes);
  if (g.tid == 0 && g.sub && sw > 0) count_add(g.sub + 11, sw);
  return sw;
}

// Block width of the LDS-staged Jacobi below for an m x n operand: the largest
// b with a 2b-column slice (max(m, n) rows) plus its 2b x 2b rotation in the
// LDS budget, or 0 when the blocking would not pay (blocks under 8 columns,
// or more than 12 blocks: the barrier rounds grow with the block count).
CATN_D int jacobi_block_width(const WG& g, int m, int n) {
  const int64_t rows = m > n ? m : n;
  int b = 0;
  while (2 * (int64_t)(b + 1) * rows + 4 * (int64_t)(b + 1) * (b + 1) <= g.lds_cap) ++b;
  if (b < 8 || b >= n) return 0;
  return (n + b - 1) / b <= 12 ? b : 0;
}

// One-sided Jacobi of an operand in global memory by column blocks: the
// columns are cut into nb blocks of b; the outer sweep walks the round-robin
// schedule of block pairs, and each block pair (I, J) is loaded into LDS, given
// one round-robin sweep over its |I| + |J| columns there (jacobi_core, which
// accumulates the pair's rotations into a small matrix Q), and written back,
// V's columns as V[:, I u J] Q. Every column pair is tested at least once per
// outer sweep with the same threshold and cosine test as the plain sweep, so
// an outer sweep with no rotation is the plain convergence test. A and V
// cross HBM once per outer round (nb - 1 per sweep, nb - 1 + (nb & 1) with a
// padding block) instead of once per column round (n - 1 per sweep).
CATN_DN int wg_jacobi_blocked(const WG& g, double* A, int m, int n, int64_t lda, double* V,
                              int64_t ldv, int maxsweeps, int b) {
  for (int64_t idx = g.tid; idx < (int64_t)n * n; idx += g.nthr) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    V[i + (int64_t)j * ldv] = (i == j) ? 1.0 : 0.0;
  }
  double fpart = 0.0;
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const double v = A[(idx % m) + (idx / m) * lda];
    fpart += v * v;
  }
  const double thr = 1e-36 * wg_sum(g, fpart);   // the plain sweep's noise threshold
  const int rows = m > n ? m : n;
  const int nb = (n + b - 1) / b, NB = nb + (nb & 1);
  double* Xs = lds_base();                     // |I|+|J| columns of A, then of V (ld m / n)
  double* Qs = Xs + (int64_t)rows * 2 * b;     // (|I|+|J|)^2 rotation, ld |I|+|J|
  for (int sweep = 1; sweep <= maxsweeps; ++sweep) {
    bool rotated = false;
    for (int r = 0; r < NB - 1; ++r) {
      for (int k = 0; k < NB / 2; ++k) {
        int P, Q;
        if (k == 0) { P = r; Q = NB - 1; }
        else { P = (r + k) % (NB - 1); Q = (r - k + NB - 1) % (NB - 1); }
        if (P > Q) { const int t = P; P = Q; Q = t; }
        if (Q >= nb) continue;   // the padding block
        const int i0 = P * b, wi = (n - i0) < b ? (n - i0) : b;
        const int j0 = Q * b, wj = (n - j0) < b ? (n - j0) : b;
        const int w = wi + wj;
        for (int64_t idx = g.tid; idx < (int64_t)m * w; idx += g.nthr) {
          const int i = (int)(idx % m), c = (int)(idx / m);
          const int col = c < wi ? i0 + c : j0 + c - wi;
          Xs[idx] = A[i + lda * (int64_t)col];
        }
        g.sync();
        if (jacobi_core(g, Xs, m, w, m, Qs, w, 1, thr, 0) < 0) rotated = true;
        for (int64_t idx = g.tid; idx < (int64_t)m * w; idx += g.nthr) {
          const int i = (int)(idx % m), c = (int)(idx / m);
          const int col = c < wi ? i0 + c : j0 + c - wi;
          A[i + lda * (int64_t)col] = Xs[idx];
        }
        g.sync();
        for (int64_t idx = g.tid; idx < (int64_t)n * w; idx += g.nthr) {
          const int i = (int)(idx % n), c = (int)(idx / n);
          const int col = c < wi ? i0 + c : j0 + c - wi;
          Xs[idx] = V[i + ldv * (int64_t)col];
        }
        g.sync();
        for (int64_t idx = g.tid; idx < (int64_t)n * w; idx += g.nthr) {
          const int i = (int)(idx % n), c = (int)(idx / n);
          const int col = c < wi ? i0 + c : j0 + c - wi;
          double s = 0.0;
          for (int q = 0; q < w; ++q) s += Xs[i + (int64_t)n * q] * Qs[q + (int64_t)w * c];
          V[i + ldv * (int64_t)col] = s;
        }
        g.sync();
      }
    }
    if (!rotated) {
      if (g.tid == 0 && g.sub) count_add(g.sub + 11, sweep);
      return sweep;
    }
  }
  return -1;
}

// ---------------------------------------------------------------------------
// full thin SVD: A (m x n, lda) is NOT modified. Outputs, sorted descending:
// S[r], U (m x r, ld m), V (n x r, ld n), r = min(m, n). Returns r, or -1 on
// scratch exhaustion / non-convergence.
// ---------------------------------------------------------------------------
struct SvdResult {
  int r;
  double* U;  // m x r
  double* S;  // r
  double* V;  // n x r
};

// Sorted outputs from the converged Jacobi state: W (rows x n, ld ldw) holds
// A V (orthogonal columns), V (n x n, ld ldv). Columns go to global U (top
// `rows` rows of an ldu-leading matrix), Vs, Ss in descending order of the
// column norms (stable in the column index). `key` is an n-double LDS/global
// scratch, `rk` an n-int scratch.
CATN_D void jacobi_finish(const WG& g, const double* W, int rows, int n, int64_t ldw,
                          const double* V, int64_t ldv, double* key, int* rk, double* U,
                          int64_t ldu, double* Ss, double* Vs) {
  for (int j = g.tid; j < n; j += g.nthr) {
    const double* c = W + (int64_t)j * ldw;
    double s = 0.0;
    for (int i = 0; i < rows; ++i) s += c[i] * c[i];
    key[j] = sqrt(s);
  }
  g.sync();
  for (int j = g.tid; j < n; j += g.nthr) {
    const double sj = key[j];
    int r = 0;
    for (int i = 0; i < n; ++i) {
      const double si = key[i];
      r += (si > sj) || (si == sj && i < j);
    }
    rk[j] = r;
    Ss[r] = sj;
  }
  g.sync();
  for (int64_t idx = g.tid; idx < (int64_t)rows * n; idx += g.nthr) {
    const int i = (int)(idx % rows), j = (int)(idx / rows);
    const double sj = key[j];
    const int r = rk[j];
    // a zero column keeps the unit vector e_r (LAPACK returns some
    // orthonormal completion; only the all-zero matrix keeps such a column)
    U[i + (int64_t)r * ldu] = sj > 0.0 ? W[i + (int64_t)j * ldw] / sj : (i == r ? 1.0 : 0.0);
  }
  for (int64_t idx = g.tid; idx < (int64_t)n * n; idx += g.nthr) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    Vs[i + (int64_t)rk[j] * n] = V[i + (int64_t)j * ldv];
  }
  g.sync();
}

// Tiny operands (m * n <= kTinySvd): lane 0 computes the whole thin SVD in
// LDS while the other lanes wait at one barrier instead of the dozen
// workgroup barriers of the parallel paths. Measured on the n=100 spin glass:
// a small gain at 16 elements; at 64 the serial LDS work was slower than the
// barriers. Same round-robin rotations, per-pair test, noise threshold, sweep
// limit and output conventions as jacobi_core + jacobi_finish; wide operands
// are handled through their transpose as in wg_svd.
constexpr int kTinySvd = 16;

CATN_DN int wg_svd_tiny(const WG& g, Scratch& sc, const double* A, int m, int n, int64_t lda,
                        SvdResult& out, int* sweeps_used) {
  const bool tr = m < n;
  const int tm = tr ? n : m, tn = tr ? m : n;
  double* U = sc.alloc((int64_t)tm * tn);
  double* Ss = sc.alloc(tn);
  double* Vs = sc.alloc((int64_t)tn * tn);
  if (sc.fail) return -1;
  double* W = lds_base();
  double* V = W + tm * tn;
  double* key = V + tn * tn;
  if (g.tid == 0) {
    double fro2 = 0.0;
    for (int j = 0; j < tn; ++j)
      for (int i = 0; i < tm; ++i) {
        const double v = tr ? A[j + (int64_t)i * lda] : A[i + (int64_t)j * lda];
        W[i + tm * j] = v;
        fro2 += v * v;
      }
    for (int j = 0; j < tn; ++j)
      for (int i = 0; i < tn; ++i) V[i + tn * j] = i == j ? 1.0 : 0.0;
    const double thr = 1e-36 * fro2;
    const double tol = fmax(4.0 * 2.220446049250313e-16 * sqrt((double)(tm > 1 ? tm : 1)),
                            g.jac_tol);
    const int N = tn + (tn & 1), npairs = N / 2;
    int sw = tn < 2 ? 0 : -1;
    for (int sweep = 1; sweep <= 60 && sw < 0; ++sweep) {
      int rotated = 0;
      for (int r = 0; r < N - 1; ++r)
        for (int k = 0; k < npairs; ++k) {
          int p, q;
          if (k == 0) { p = r; q = N - 1; }
          else { p = (r + k) % (N - 1); q = (r - k + N - 1) % (N - 1); }
          if (p > q) { const int t = p; p = q; q = t; }
          if (q >= tn) continue;
          double* ap = W + tm * p;
          double* aq = W + tm * q;
          double al = 0.0, be = 0.0, ga = 0.0;
          for (int i = 0; i < tm; ++i) {
            al += ap[i] * ap[i]; be += aq[i] * aq[i]; ga += ap[i] * aq[i];
          }
          if (al <= thr || be <= thr) continue;
          if (fabs(ga) <= tol * sqrt(al) * sqrt(be)) continue;
          rotated = 1;
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
          for (int i = 0; i < tm; ++i) {
            const double x = ap[i], y = aq[i];
            ap[i] = c * x - sn * y;
            aq[i] = sn * x + c * y;
          }
          double* vp = V + tn * p;
          double* vq = V + tn * q;
          for (int i = 0; i < tn; ++i) {
            const double x = vp[i], y = vq[i];
            vp[i] = c * x - sn * y;
            vq[i] = sn * x + c * y;
          }
        }
      if (!rotated) sw = sweep;
    }
    // norms, descending order stable in the column index, U = W / s (a zero
    // column keeps the unit vector e_r), V's columns in the same order
    for (int j = 0; j < tn; ++j) {
      double ss = 0.0;
      for (int i = 0; i < tm; ++i) ss += W[i + tm * j] * W[i + tm * j];
      key[j] = sqrt(ss);
    }
    for (int j = 0; j < tn; ++j) {
      const double sj = key[j];
      int rr = 0;
      for (int i = 0; i < tn; ++i) rr += (key[i] > sj) || (key[i] == sj && i < j);
      Ss[rr] = sj;
      for (int i = 0; i < tm; ++i)
        U[i + (int64_t)tm * rr] = sj > 0.0 ? W[i + tm * j] / sj : (i == rr ? 1.0 : 0.0);
      for (int i = 0; i < tn; ++i) Vs[i + (int64_t)tn * rr] = V[i + tn * j];
    }
    *g.flag = sw;
  }
  g.sync();
  const int sw = *g.flag;
  g.sync();   // every lane has read the flag before it is reused
  if (sweeps_used) *sweeps_used = sw;
  out.r = tn;
  out.S = Ss;
  out.U = tr ? Vs : U;   // m x r (ld m)
  out.V = tr ? U : Vs;   // n x r (ld n)
  return sw < 0 ? -1 : tn;
}

// LDS-resident path: the whole m x n operand, V and the norms live in LDS.
CATN_DN int wg_svd_lds(const WG& g, Scratch& sc, const double* A, int m, int n, int64_t lda,
                       bool transpose_in, SvdResult& out, int* sweeps_used) {
  double* L = lds_base();
  double* W = L;                              // m x n
  double* V = W + (int64_t)m * n;             // n x n
  double* key = V + (int64_t)n * n;           // n
  int* rk = (int*)(key + n);                  // n ints
  double* U = sc.alloc((int64_t)m * n);
  double* Ss = sc.alloc(n);
  double* Vs = sc.alloc((int64_t)n * n);
  if (sc.fail) return -1;
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int i = (int)(idx % m), j = (int)(idx / m);
    W[idx] = transpose_in ? A[j + (int64_t)i * lda] : A[i + (int64_t)j * lda];
  }
  g.sync();
  const int sw = jacobi_core(g, W, m, n, m, V, n, 60);
  if (sweeps_used) *sweeps_used = sw;
  jacobi_finish(g, W, m, n, m, V, n, key, rk, U, m, Ss, Vs);
  out.r = n;
  out.U = U;
  out.S = Ss;
  out.V = Vs;
  return sw < 0 ? -1 : n;
}

// LDS-resident Jacobi on the R factor of a global Householder QR.
CATN_DN int wg_jacobi_r_lds(const WG& g, const double* R, int n, double* U, int64_t ldu,
                            double* Ss, double* Vs, int* sweeps_used) {
  double* L = lds_base();
  double* W = L;
  double* V = W + (int64_t)n * n;
  double* key = V + (int64_t)n * n;
  int* rk = (int*)(key + n);
  for (int64_t idx = g.tid; idx < (int64_t)n * n; idx += g.nthr) W[idx] = R[idx];
  g.sync();
  const int sw = jacobi_core(g, W, n, n, n, V, n, 60);
  if (sweeps_used) *sweeps_used = sw;
  jacobi_finish(g, W, n, n, n, V, n, key, rk, U, ldu, Ss, Vs);
  return sw;
}

// Householder QR with column pivoting (LAPACK dgeqp2 structure) of A (m x n,
// m >= n, lda): A P = Q R. R overwrites the upper triangle, the reflectors
// (v(0) = 1 implicit) the part below it, tau[j] their factors; perm[j] is the
// original index of column j. The pivot is the column of largest remaining
// norm (first on ties); the trailing norms are recomputed exactly in each
// update pass, so no norm downdating is involved. cn: n doubles of scratch.
CATN_DN void wg_geqp2(const WG& g, double* A, int m, int n, int64_t lda, double* tau, int* perm,
                      double* cn) {
  {
    int L = pow2floor(g.nthr / n < 1 ? 1 : g.nthr / n);
    if (L > 64) L = 64;
    const int ng = g.nthr / L, grp = g.tid / L, lane = g.tid % L;
    for (int c = grp; c < n; c += ng) {
      const double* y = A + (int64_t)c * lda;
      double w = 0.0;
      for (int i = lane; i < m; i += L) w += y[i] * y[i];
      w = grp_sum(w, L);
      if (lane == 0) { cn[c] = w; perm[c] = c; }
    }
  }
  g.sync();
  for (int j = 0; j < n; ++j) {
    // pivot: every lane scans the same norms, so the choice is uniform
    int p = j;
    double best = cn[j];
    for (int c = j + 1; c < n; ++c)
      if (cn[c] > best) { best = cn[c]; p = c; }
    g.sync();   // all lanes have read cn before lane 0 swaps it
    if (p != j) {
      double* a = A + (int64_t)j * lda;
      double* b = A + (int64_t)p * lda;
      for (int i = g.tid; i < m; i += g.nthr) {
        const double t = a[i];
        a[i] = b[i];
        b[i] = t;
      }
      if (g.tid == 0) {
        const double t = cn[j]; cn[j] = cn[p]; cn[p] = t;
        const int q = perm[j]; perm[j] = perm[p]; perm[p] = q;
      }
    }
    g.sync();
    double* x = A + j + (int64_t)j * lda;
    double part = 0.0;
    for (int i = 1 + g.tid; i < m - j; i += g.nthr) part += x[i] * x[i];
    const double s = wg_sum(g, part);
    const double alpha = x[0];
    double t = 0.0, beta = alpha, scal = 0.0;
    const int ncol = n - j - 1;
    if (s > 0.0) {
      const double nrm = sqrt(alpha * alpha + s);
      beta = alpha >= 0.0 ? -nrm : nrm;
      t = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (ncol > 0) {
      int L = pow2floor(g.nthr / ncol < 1 ? 1 : g.nthr / ncol);
      if (L > 64) L = 64;
      const int ng = g.nthr / L, grp = g.tid / L, lane = g.tid % L;
      for (int c = grp; c < ncol; c += ng) {
        double* y = A + j + (int64_t)(j + 1 + c) * lda;
        const double y0 = y[0];
        double w = 0.0;
        if (s > 0.0) {
          for (int i = 1 + lane; i < m - j; i += L) w += x[i] * y[i];
          w = grp_sum(w, L);
          w = t * (y0 + scal * w);
        }
        const double ws = w * scal;
        double nq = 0.0;   // squared norm of the updated column below row j
        for (int i = 1 + lane; i < m - j; i += L) {
          const double v = y[i] - ws * x[i];
          if (s > 0.0) y[i] = v;
          nq += v * v;
        }
        nq = grp_sum(nq, L);
        if (lane == 0) {
          y[0] = y0 - w;
          cn[j + 1 + c] = nq;
        }
      }
    }
    g.sync();   // every group has read x (unscaled) before it is scaled
    if (s > 0.0)
      for (int i = 1 + g.tid; i < m - j; i += g.nthr) x[i] *= scal;
    if (g.tid == 0) {
      tau[j] = t;
      x[0] = beta;
    }
    g.sync();
  }
}

CATN_DN int wg_svd_tall(const WG& g, Scratch& sc, const double* A, int m, int n, int64_t lda,
                        bool transpose_in, SvdResult& out, int* sweeps_used) {
  // SVD of B = A (m x n) or B = A^T when transpose_in (then m, n are B's).
  // Requires m >= n. Produces U (m x n), S, V (n x n) of B, sorted.
  // Operands that fit run entirely in LDS. Otherwise (the QR-preconditioned
  // one-sided Jacobi of Drmac and Veselic): blocked Householder QR reduces B
  // to R0 (n x n; B itself when square), a column-pivoted QR gives
  // R0 P = Q2 R, and the Jacobi sweeps run on R^T, whose graded columns
  // converge in a few sweeps: R^T Vj = U' S, so B = (Q1 Q2 Vj) S (P U')^T.
  // The Jacobi runs in LDS when R and its rotations fit, else by column blocks
  // staged in LDS (wg_jacobi_blocked), or column pair by column pair in global
  // memory when blocking would not pay.
  const int64_t lds_direct = (int64_t)m * n + (int64_t)n * n + 2 * (int64_t)n;
  if (lds_direct <= g.lds_cap) return wg_svd_lds(g, sc, A, m, n, lda, transpose_in, out, sweeps_used);
  double* W = sc.alloc((int64_t)m * n);
  double* U = sc.alloc((int64_t)m * n);
  double* Ss = sc.alloc(n);
  double* Vs = sc.alloc((int64_t)n * n);
  if (sc.fail) return -1;
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int i = (int)(idx % m), j = (int)(idx / m);
    W[idx] = transpose_in ? A[j + (int64_t)i * lda] : A[i + (int64_t)j * lda];
  }
  g.sync();
  const bool use_qr = m > n;
  int sw = 0;
  double* tau = nullptr;
  double* R = W;   // n x n, ld n (B itself when square)
  int64_t t0 = now_ticks();
  double* Tall = nullptr;
  bool blocked = false;
  if (use_qr) {
    tau = sc.alloc(n);
    R = sc.alloc((int64_t)n * n);
    Tall = sc.alloc((int64_t)kQBMax * n);
    if (sc.fail) return -1;
    blocked = wg_geqrf(g, sc, W, m, n, m, tau, Tall);
    if (sc.fail) return -1;
    for (int64_t idx = g.tid; idx < (int64_t)n * n; idx += g.nthr) {
      const int i = (int)(idx % n), j = (int)(idx / n);
      R[idx] = (i <= j) ? W[i + (int64_t)j * m] : 0.0;
    }
    g.sync();
  }
  // R0 P = Q2 R (in place), then X = R^T
  double* tau2 = sc.alloc(n);
  double* cn = sc.alloc(n);
  int* perm = (int*)sc.alloc(n);
  double* X = sc.alloc((int64_t)n * n);
  double* Y = sc.alloc((int64_t)n * n);
  double* Vj = sc.alloc((int64_t)n * n);
  if (sc.fail) return -1;
  wg_geqp2(g, R, n, n, n, tau2, perm, cn);
  for (int64_t idx = g.tid; idx < (int64_t)n * n; idx += g.nthr) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    X[idx] = (j <= i) ? R[j + (int64_t)i * n] : 0.0;
  }
  g.sync();
  sub_tick(g, 0, t0);
  t0 = now_ticks();
  if (2 * (int64_t)n * n + 2 * (int64_t)n <= g.lds_cap) {
    sw = wg_jacobi_r_lds(g, X, n, Y, n, Ss, Vj, sweeps_used);
    sub_tick(g, 1, t0);
    t0 = now_ticks();
  } else {
    double* V = sc.alloc((int64_t)n * n);
    double* key = sc.alloc(n);
    int* rk = (int*)sc.alloc(n);
    if (sc.fail) return -1;
    const int jb = jacobi_block_width(g, n, n);
    sw = jb ? wg_jacobi_blocked(g, X, n, n, n, V, n, 60, jb) : wg_jacobi(g, X, n, n, n, V, n, 60);
    if (sweeps_used) *sweeps_used = sw;
    sub_tick(g, 1, t0);
    t0 = now_ticks();
    jacobi_finish(g, X, n, n, n, V, n, key, rk, Y, n, Ss, Vj);
  }
  // V of B: rows permuted back; U of B: Q1 Q2 Vj
  for (int64_t idx = g.tid; idx < (int64_t)n * n; idx += g.nthr) {
    const int i = (int)(idx % n), c = (int)(idx / n);
    Vs[perm[i] + (int64_t)c * n] = Y[idx];
    U[i + (int64_t)c * m] = Vj[idx];
  }
  g.sync();
  wg_apply_q(g, R, n, n, n, tau2, U, n, m);
  if (use_qr) {
    for (int64_t idx = g.tid; idx < (int64_t)(m - n) * n; idx += g.nthr) {
      const int i = n + (int)(idx % (m - n)), c = (int)(idx / (m - n));
      U[i + (int64_t)c * m] = 0.0;
    }
    g.sync();
    wg_apply_q_blocked(g, sc, W, m, n, m, tau, blocked ? Tall : nullptr, U, n, m);
  }
  sub_tick(g, 2, t0);
  if (g.tid == 0 && g.sub && sw > 0) count_add(g.sub + 3, sw);
  out.r = n;
  out.U = U;
  out.S = Ss;
  out.V = Vs;
  return sw < 0 ? -1 : n;
}

// Fallback SVD (tsvd's retry with LAPACK's QR-iteration driver when the
// divide-and-conquer one fails, src/linalg_utils.jl:16-21): one-sided Jacobi
// straight on B = A, or A^T when m < n, held in scratch, with no QR
// preconditioner and up to 200 sweeps (the primary paths stop at 60). Same
// output conventions as wg_svd. Returns min(m, n), or -1.
CATN_DN int wg_svd_fallback(const WG& g, Scratch& sc, const double* A, int m, int n, int64_t lda,
                            SvdResult& out, int* sweeps_used) {
  const bool tr = m < n;
  const int tm = tr ? n : m, tn = tr ? m : n;
  double* W = sc.alloc((int64_t)tm * tn);
  double* V = sc.alloc((int64_t)tn * tn);
  double* U = sc.alloc((int64_t)tm * tn);
  double* Ss = sc.alloc(tn);
  double* Vs = sc.alloc((int64_t)tn * tn);
  double* key = sc.alloc(tn);
  int* rk = (int*)sc.alloc(tn);
  if (sc.fail) return -1;
  for (int64_t idx = g.tid; idx < (int64_t)tm * tn; idx += g.nthr) {
    const int i = (int)(idx % tm), j = (int)(idx / tm);
    W[idx] = tr ? A[j + (int64_t)i * lda] : A[i + (int64_t)j * lda];
  }
  g.sync();
  const int sw = wg_jacobi(g, W, tm, tn, tm, V, tn, 200);
  if (sweeps_used) *sweeps_used = sw;
  jacobi_finish(g, W, tm, tn, tm, V, tn, key, rk, U, tm, Ss, Vs);
  if (g.tid == 0 && g.sub) count_add(g.sub + 12, 1);
  out.r = tn;
  out.S = Ss;
  out.U = tr ? Vs : U;   // m x r (ld m)
  out.V = tr ? U : Vs;   // n x r (ld n)
  return sw < 0 ? -1 : tn;
}

// A (m x n, lda) -> SVD of A with U (m x r), V (n x r), r = min(m,n). A
// primary path that does not converge (or, with g.svd_fault, any primary
// path) hands over to wg_svd_fallback.
CATN_DN int wg_svd(const WG& g, Scratch& sc, const double* A, int m, int n, int64_t lda,
                   SvdResult& out, int* sweeps_used) {
  int r;
  if ((int64_t)m * n <= kTinySvd) {
    r = wg_svd_tiny(g, sc, A, m, n, lda, out, sweeps_used);
  } else if (m >= n) {
    r = wg_svd_tall(g, sc, A, m, n, lda, false, out, sweeps_used);
  } else {
    // A^T = U' S V'^T  ->  A = V' S U'^T
    SvdResult t;
    r = wg_svd_tall(g, sc, A, n, m, lda, true, t, sweeps_used);
    out.r = t.r;
    out.U = t.V;   // m x r (ld m)
    out.S = t.S;
    out.V = t.U;   // n x r (ld n)
  }
  if (sc.fail || (r >= 0 && !g.svd_fault)) return r;
  return wg_svd_fallback(g, sc, A, m, n, lda, out, sweeps_used);
}


// ---------------------------------------------------------------------------
// Randomized SVD (src/linalg_utils.jl:44-58).
//
// The sketch is standard normal, drawn by a counter-based generator: entry i
// of stream `stream` is a SplitMix64 hash of (seed, stream, i) turned into two
// uniforms and a Box-Muller normal, so every lane draws its entries
// independently and a replay draws the same sketch. (The reference draws
// from Julia's global RNG; its stream is not reproducible elsewhere.)
// ---------------------------------------------------------------------------
CATN_D uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

CATN_D double gauss_at(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t h1 = splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) ^ (2 * i));
  const uint64_t h2 = splitmix64(h1 ^ (2 * i + 1));
  const double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);   // (0, 1)
  const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);           // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

CATN_DN bool wg_thin_qr(const WG& g, Scratch& sc, double* Wk, int m, int n, double** Qo, double** Ro);

// Y (m x p, ld m) := thin Q of Y, through scratch that is released again
CATN_DN bool wg_orthonormalise(const WG& g, Scratch& sc, double* Y, int m, int p) {
  const int64_t save = sc.top;
  double *Q = nullptr, *R = nullptr;
  if (!wg_thin_qr(g, sc, Y, m, p, &Q, &R)) return false;
  wg_copy(g, Y, Q, (int64_t)m * p);
  g.sync();
  sc.top = save;
  return true;
}

// rsvd of A (m x n, lda): U (m x kk, ld m), S (kk), V (n x kk, ld n) in
// scratch, kk = min(k, p) with p = min(n, oversample k). Requires p <= m.
// Returns kk, or -1 (scratch exhausted or the small SVD did not converge).
CATN_DN int wg_rsvd(const WG& g, Scratch& sc, const double* A, int m, int n, int64_t lda, int k,
                    int oversample, int power, uint64_t seed, uint64_t stream, double** Uo,
                    double** So, double** Vo) {
  const int p = n < oversample * k ? n : oversample * k;
  if (p > m || p <= 0) return -1;
  double* G = sc.alloc((int64_t)n * p);
  double* Y = sc.alloc((int64_t)m * p);
  double* Z = sc.alloc((int64_t)n * p);
  if (sc.fail) return -1;
  for (int64_t idx = g.tid; idx < (int64_t)n * p; idx += g.nthr) G[idx] = gauss_at(seed, stream, idx);
  g.sync();
  wg_gemm_any(g, m, p, n, 1.0, A, lda, false, G, n, false, 0.0, Y, m);          // Y = A G
  for (int it = 0; it < power; ++it) {
    wg_gemm_any(g, n, p, m, 1.0, A, lda, true, Y, m, false, 0.0, Z, n);         // Z = A' Y
    wg_gemm_any(g, m, p, n, 1.0, A, lda, false, Z, n, false, 0.0, Y, m);        // Y = A Z
    if (!wg_orthonormalise(g, sc, Y, m, p)) return -1;
  }
  if (!wg_orthonormalise(g, sc, Y, m, p)) return -1;                            // Q = thinQ(Y)
  double* B = Z;                                                                  // p x n (reuses Z)
  wg_gemm_any(g, p, n, m, 1.0, Y, m, true, A, lda, false, 0.0, B, p);           // B = Q' A
  SvdResult F;
  if (wg_svd(g, sc, B, p, n, p, F, nullptr) < 0) return -1;
  const int kk = k < F.r ? k : F.r;
  double* U = sc.alloc((int64_t)m * kk);
  if (sc.fail) return -1;
  wg_gemm_any(g, m, kk, p, 1.0, Y, m, false, F.U, p, false, 0.0, U, m);         // U = Q U_B
  *Uo = U;
  *So = F.S;
  *Vo = F.V;   // n x r (ld n); the first kk columns are the result
  return kk;
}

// ---------------------------------------------------------------------------
// Certified full-rank factorisation (no truncation possible).
//
// When every singular value of A (m x n) provably exceeds the truncation
// threshold and min(m, n) <= maxdim, the reference's truncated SVD keeps all
// min(m, n) singular triplets and discards nothing; its factors U and S V'
// then differ from any other factorisation A = L * Rt with the same
// isometric side only by a unitary gauge on the new bond, which every later
// contraction and SVD is invariant to (same bond dimension, same canonical
// form, same represented tensor). Such steps are factorised by Householder QR
// instead of the Jacobi SVD. The certificate is a lower bound on the smallest
// singular value from the QR's triangular factor,
//     sigma_min(A) = sigma_min(R) >= 1 / ||R^{-1}||_F,
// accepted only above max(100 * cutoff, 1e-12 * ||A||_F), where the inverse
// is computed to better than 1e-3 relative accuracy; any step that is not
// certified takes the SVD path.
//
// iso_left: L (m x r) has orthonormal columns (U side), Rt is the centre.
// otherwise Rt (r x n) has orthonormal rows (V' side), L is the centre.
// Outputs live in scratch: L (ld m), Rt (ld r).
// ---------------------------------------------------------------------------
struct Factor {
  int r;
  double* L;
  double* Rt;
};

// 1 / ||R^{-1}||_F of an upper-triangular n x n R (ld ldr); 0 if singular.
// X = R^{-1} (n x n, ld n) is built row by row from the bottom: row i needs
// only rows > i, so each step is one parallel pass over the columns j > i.
CATN_DN double wg_inv_fro_bound_blocked(const WG& g, const double* R, int n, int64_t ldr,
                                        double* X);

CATN_DN double wg_inv_fro_bound(const WG& g, const double* R, int n, int64_t ldr, double* X) {
  if (n > 32) return wg_inv_fro_bound_blocked(g, R, n, ldr, X);
  double part = 0.0;
  for (int i = n - 1; i >= 0; --i) {
    const double rii = R[i + (int64_t)i * ldr];
    if (rii == 0.0) return 0.0;      // uniform: every lane reads the same value
    const double inv = 1.0 / rii;
    for (int j = i + g.tid; j < n; j += g.nthr) {
      double x;
      if (j == i) {
        x = inv;
      } else {
        double sacc = 0.0;
        for (int k = i + 1; k <= j; ++k) sacc += R[i + (int64_t)k * ldr] * X[k + (int64_t)j * n];
        x = -sacc * inv;
      }
      X[i + (int64_t)j * n] = x;
      part += x * x;
    }
    g.sync();
  }
  const double tot = wg_sum(g, part);
  if (!(tot < INFINITY) || tot <= 0.0) return 0.0;
  return 1.0 / sqrt(tot);
}

// The same bound for larger R, by blocks of kInvB rows from the bottom:
// X(I, I) = R(I, I)^{-1} by back substitution, one lane per column of the
// block; then X(I, c) = -X(I, I) (R(I, i1:c) X(i1:c, c)) for the columns c
// right of the block, one lane per column. Two barriers per block instead of
// one per row.
constexpr int kInvB = 16;

CATN_DN double wg_inv_fro_bound_blocked(const WG& g, const double* R, int n, int64_t ldr,
                                        double* X) {
  for (int i = 0; i < n; ++i)          // uniform: every lane reads the same diagonal
    if (R[i + (int64_t)i * ldr] == 0.0) return 0.0;
  double part = 0.0;
  const int nb = (n + kInvB - 1) / kInvB;
  for (int ib = nb - 1; ib >= 0; --ib) {
    const int i0 = ib * kInvB;
    const int i1 = (i0 + kInvB) < n ? (i0 + kInvB) : n;
    const int bs = i1 - i0;
    // diagonal block: column j of R(I,I)^{-1}, rows i0 + j down to i0
    for (int j = g.tid; j < bs; j += g.nthr) {
      const int c = i0 + j;
      for (int r = i0 + bs - 1; r > c; --r) X[r + (int64_t)c * n] = 0.0;
      for (int r = c; r >= i0; --r) {
        double acc = r == c ? 1.0 : 0.0;
        for (int q = r + 1; q <= c; ++q) acc -= R[r + (int64_t)q * ldr] * X[q + (int64_t)c * n];
        const double x = acc / R[r + (int64_t)r * ldr];
        X[r + (int64_t)c * n] = x;
        part += x * x;
      }
    }
    g.sync();
    // columns right of the block
#if CATN_DEV
    // on the FP64 matrix cores, one wavefront per 16-column tile: the tile
    // T = R(I, i1:c) X(i1:c, tile) accumulates over K = i1 .. tile end (the
    // diagonal stage of each block zeroed X below its diagonal, so the
    // triangular K range is exact), then X(I, tile) = -X(I, I) T. The
    // accumulator layout of T (lane l, element v: row l/16 + 4v, column l%16)
    // is the B-operand layout of K-step v of the second product.
    {
      const int nw = g.nthr >> 6, wv = g.tid >> 6, l = g.tid & 63;
      const int li = l & 15, lk = l >> 4;
      for (int c0 = i1 + 16 * wv; c0 < n; c0 += 16 * nw) {
        const int c = c0 + li;
        const int kend = (c0 + 16) < n ? (c0 + 16) : n;
        catn_d4 tt = {0.0, 0.0, 0.0, 0.0};
        for (int q0 = i1; q0 < kend; q0 += 4) {
          const int q = q0 + lk;
          const double a = (li < bs && q < kend) ? R[(i0 + li) + (int64_t)q * ldr] : 0.0;
          const double b = (q < kend && c < n) ? X[q + (int64_t)c * n] : 0.0;
          tt = mfma_f64_16x16x4(a, b, tt);
        }
        catn_d4 xo = {0.0, 0.0, 0.0, 0.0};
        for (int st = 0; st < 4; ++st) {
          const int k = 4 * st + lk;
          const double a = (li < bs && k < bs) ? -X[(i0 + li) + (int64_t)(i0 + k) * n] : 0.0;
          xo = mfma_f64_16x16x4(a, tt[st], xo);
        }
        for (int v = 0; v < 4; ++v) {
          const int r = lk + 4 * v;
          if (r < bs && c < n) {
            X[(i0 + r) + (int64_t)c * n] = xo[v];
            part += xo[v] * xo[v];
          }
        }
      }
    }
#else
    for (int c = i1 + g.tid; c < n; c += g.nthr) {
      double t[kInvB];
      for (int r = 0; r < bs; ++r) t[r] = 0.0;
      for (int q = i1; q <= c; ++q) {
        const double xq = X[q + (int64_t)c * n];
        for (int r = 0; r < bs; ++r) t[r] += R[(i0 + r) + (int64_t)q * ldr] * xq;
      }
      for (int r = 0; r < bs; ++r) {
        double acc = 0.0;
        for (int q = r; q < bs; ++q) acc -= X[(i0 + r) + (int64_t)(i0 + q) * n] * t[q];
        X[(i0 + r) + (int64_t)c * n] = acc;
        part += acc * acc;
      }
    }
#endif
    g.sync();
  }
  const double tot = wg_sum(g, part);
  if (!(tot < INFINITY) || tot <= 0.0) return 0.0;
  return 1.0 / sqrt(tot);
}

// wg_factor_fullrank of a tiny operand (tm * r <= kTinySvd): lane 0 runs the
// same steps serially in LDS (Householder QR of the tall orientation, the
// inverse bound, the certificate, Q formation and the four output cases);
// the other lanes wait at one barrier. The scratch outputs are allocated by
// every lane, as the bump allocator is held per lane.
CATN_DN bool factor_fullrank_tiny(const WG& g, Scratch& sc, const double* A, int m, int n,
                                  int64_t lda, double cutoff, bool iso_left, Factor& F,
                                  double noise) {
  const int r = m < n ? m : n;
  const bool trans = m < n;
  const int tm = trans ? n : m;
  const bool need_q = iso_left != trans;
  const int64_t save = sc.top;
  double* L = sc.alloc((int64_t)m * r);
  double* Rt = sc.alloc((int64_t)r * n);
  double* Qs = need_q && !(iso_left && !trans) ? sc.alloc((int64_t)tm * r) : nullptr;
  if (sc.fail) return false;
  double* W = lds_base();          // tm x r
  double* R = W + tm * r;          // r x r
  double* X = R + r * r;           // r x r
  double* tau = X + r * r;         // r
  if (g.tid == 0) {
    double fro2 = 0.0;
    for (int j = 0; j < r; ++j)
      for (int i = 0; i < tm; ++i) {
        const double v = trans ? A[j + (int64_t)i * lda] : A[i + (int64_t)j * lda];
        W[i + tm * j] = v;
        fro2 += v * v;
      }
    for (int j = 0; j < r; ++j) {   // dgeqr2: H = I - tau v v^T, v(0) = 1
      double* x = W + j + tm * j;
      double ss = 0.0;
      for (int i = 1; i < tm - j; ++i) ss += x[i] * x[i];
      double t = 0.0;
      if (ss > 0.0) {
        const double alpha = x[0];
        const double nrm = sqrt(alpha * alpha + ss);
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        t = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        for (int i = 1; i < tm - j; ++i) x[i] *= scal;
        x[0] = beta;
        for (int c = j + 1; c < r; ++c) {
          double* y = W + j + tm * c;
          double w = y[0];
          for (int i = 1; i < tm - j; ++i) w += x[i] * y[i];
          w *= t;
          y[0] -= w;
          for (int i = 1; i < tm - j; ++i) y[i] -= w * x[i];
        }
      }
      tau[j] = t;
    }
    for (int j = 0; j < r; ++j)
      for (int i = 0; i < r; ++i) R[i + r * j] = i <= j ? W[i + tm * j] : 0.0;
    // 1 / ||R^{-1}||_F by back substitution from the bottom row
    double tot = 0.0;
    bool sing = false;
    for (int i = r - 1; i >= 0 && !sing; --i) {
      const double rii = R[i + r * i];
      if (rii == 0.0) { sing = true; break; }
      const double inv = 1.0 / rii;
      for (int j = i; j < r; ++j) {
        double xv;
        if (j == i) {
          xv = inv;
        } else {
          double acc = 0.0;
          for (int k = i + 1; k <= j; ++k) acc += R[i + r * k] * X[k + r * j];
          xv = -acc * inv;
        }
        X[i + r * j] = xv;
        tot += xv * xv;
      }
    }
    const double lb = (sing || !(tot < INFINITY) || tot <= 0.0) ? 0.0 : 1.0 / sqrt(tot);
    const double fro = sqrt(fro2);
    const double thr = fmax(fmax(100.0 * cutoff, 1e-12 * fro), noise * fro);
    const bool ok = lb > thr;
    if (ok) {
      double* Q = need_q ? ((iso_left && !trans) ? L : Qs) : nullptr;
      if (need_q) {   // Q = H_0 ... H_{r-1} [I_r; 0]
        for (int c = 0; c < r; ++c)
          for (int i = 0; i < tm; ++i) Q[i + (int64_t)tm * c] = i == c ? 1.0 : 0.0;
        for (int j = r - 1; j >= 0; --j) {
          const double t = tau[j];
          if (t == 0.0) continue;
          const double* v = W + j + tm * j;
          for (int c = 0; c < r; ++c) {
            double* y = Q + j + (int64_t)tm * c;
            double w = y[0];
            for (int i = 1; i < tm - j; ++i) w += v[i] * y[i];
            w *= t;
            y[0] -= w;
            for (int i = 1; i < tm - j; ++i) y[i] -= w * v[i];
          }
        }
      }
      if (iso_left && !trans) {          // L = Q (formed in place), Rt = R
        for (int q = 0; q < r * r; ++q) Rt[q] = R[q];
      } else if (iso_left && trans) {    // L = I_m, Rt = A
        for (int q = 0; q < m * m; ++q) L[q] = (q % m == q / m) ? 1.0 : 0.0;
        for (int q = 0; q < m * n; ++q) Rt[q] = A[(q % m) + (int64_t)(q / m) * lda];
      } else if (!iso_left && !trans) {  // L = A, Rt = I_n
        for (int q = 0; q < m * n; ++q) L[q] = A[(q % m) + (int64_t)(q / m) * lda];
        for (int q = 0; q < n * n; ++q) Rt[q] = (q % n == q / n) ? 1.0 : 0.0;
      } else {                           // L = R'^T (m x m), Rt = Q'^T (m x n)
        for (int q = 0; q < m * m; ++q) L[q] = R[(q / m) + r * (q % m)];
        for (int q = 0; q < m * n; ++q) Rt[q] = Q[(q / m) + (int64_t)tm * (q % m)];
      }
    }
    *g.flag = ok ? 1 : 0;
  }
  g.sync();
  const bool ok = *g.flag != 0;
  g.sync();   // every lane has read the flag before it is reused
  if (!ok) {
    sc.top = save;
    return false;
  }
  F.r = r;
  F.L = L;
  F.Rt = Rt;
  return true;
}

// assume_certified: the caller knows the step was certified (the reverse
// sweep's replay of a recorded certified step on the recorded input, which
// reproduces the same R), so the inverse bound is not formed again.
CATN_DN bool wg_factor_fullrank(const WG& g, Scratch& sc, const double* A, int m, int n,
                                int64_t lda, double cutoff, int maxdim, bool iso_left,
                                Factor& F, double noise = 0.0, bool assume_certified = false) {
  const int r = m < n ? m : n;
  if (r <= 0 || r > maxdim) return false;
  if ((int64_t)(m > n ? m : n) * r <= kTinySvd)
    return factor_fullrank_tiny(g, sc, A, m, n, lda, cutoff, iso_left, F, noise);
  const int64_t save = sc.top;
  // tall orientation T (tm x r): A itself when m >= n, else A^T
  const bool trans = m < n;
  const int tm = trans ? n : m;
  // the isometric factor is Q of the tall orientation when that side is the
  // one QR'd (iso_left & !trans, or !iso_left & trans); otherwise identity
  const bool need_q = iso_left != trans;
  const int64_t tmr = (int64_t)tm * r, rr = (int64_t)r * r;
  // small operands: W, R, R^{-1} in LDS, unblocked Householder; Q is formed
  // straight into its scratch destination either way
  const bool in_lds = tmr + 2 * rr + r + 64 <= g.lds_cap;
  double *W, *R, *X, *tau, *Tall = nullptr;
  if (in_lds) {
    W = lds_base();
    R = W + tmr;
    X = R + rr;
    tau = X + rr;
  } else {
    W = sc.alloc(tmr);
    tau = sc.alloc(r);
    R = sc.alloc(rr);
    X = sc.alloc(rr);
    Tall = sc.alloc((int64_t)kQBMax * r);
    if (sc.fail) return false;
  }
  double fpart = 0.0;
  for (int64_t idx = g.tid; idx < tmr; idx += g.nthr) {
    const int i = (int)(idx % tm), j = (int)(idx / tm);
    const double v = trans ? A[j + (int64_t)i * lda] : A[i + (int64_t)j * lda];
    W[idx] = v;
    fpart += v * v;
  }
  const double fro = sqrt(wg_sum(g, fpart));
  g.sync();
  int64_t t0 = now_ticks();
  bool blocked = false;
  if (in_lds) wg_geqr2(g, W, tm, r, tm, tau);
  else blocked = wg_geqrf(g, sc, W, tm, r, tm, tau, Tall);
  if (sc.fail) return false;
  sub_tick(g, 4, t0);
  t0 = now_ticks();
  for (int64_t idx = g.tid; idx < rr; idx += g.nthr) {
    const int i = (int)(idx % r), j = (int)(idx / r);
    R[idx] = (i <= j) ? W[i + (int64_t)j * tm] : 0.0;
  }
  g.sync();
  const double lb = assume_certified ? INFINITY : wg_inv_fro_bound(g, R, r, r, X);
  sub_tick(g, 5, t0);
  t0 = now_ticks();
  const double thr = fmax(fmax(100.0 * cutoff, 1e-12 * fro), noise * fro);
  if (!(lb > thr)) {
    g.sync();
    sc.top = save;
    return false;
  }
  F.r = r;
  // outputs in scratch (LDS is reused by the caller's GEMMs)
  double* L = sc.alloc((int64_t)m * r);
  double* Rt = sc.alloc((int64_t)r * n);
  double* Q = nullptr;
  if (need_q) Q = (iso_left && !trans) ? L : sc.alloc(tmr);   // Q of the tall orientation
  if (sc.fail) return false;
  if (need_q) {
    for (int64_t idx = g.tid; idx < tmr; idx += g.nthr) {
      const int i = (int)(idx % tm), j = (int)(idx / tm);
      Q[idx] = (i == j) ? 1.0 : 0.0;
    }
    g.sync();
    if (in_lds) wg_apply_q(g, W, tm, r, tm, tau, Q, r, tm, true);
    else wg_apply_q_blocked(g, sc, W, tm, r, tm, tau, blocked ? Tall : nullptr, Q, r, tm, true);
    sub_tick(g, 6, t0);
  }
  if (iso_left && !trans) {          // A = Q R: L = Q (m x n, formed in place), Rt = R (n x n)
    for (int64_t idx = g.tid; idx < rr; idx += g.nthr) Rt[idx] = R[idx];
  } else if (iso_left && trans) {    // rank m: L = I_m, Rt = A (m x n)
    for (int64_t idx = g.tid; idx < (int64_t)m * m; idx += g.nthr)
      L[idx] = (idx % m == idx / m) ? 1.0 : 0.0;
    for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr)
      Rt[idx] = A[(idx % m) + (idx / m) * lda];
  } else if (!iso_left && !trans) {  // rank n: L = A (m x n), Rt = I_n
    for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr)
      L[idx] = A[(idx % m) + (idx / m) * lda];
    for (int64_t idx = g.tid; idx < (int64_t)n * n; idx += g.nthr)
      Rt[idx] = (idx % n == idx / n) ? 1.0 : 0.0;
  } else {                           // A^T = Q' R': L = R'^T (m x m), Rt = Q'^T (m x n)
    for (int64_t idx = g.tid; idx < (int64_t)m * m; idx += g.nthr) {
      const int i = (int)(idx % m), j = (int)(idx / m);
      L[idx] = R[j + (int64_t)i * r];
    }
    for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
      const int i = (int)(idx % m), j = (int)(idx / m);
      Rt[idx] = Q[j + (int64_t)i * tm];
    }
  }
  F.L = L;
  F.Rt = Rt;
  g.sync();
  return true;
}

// ---------------------------------------------------------------------------
// ising_network site values (src/ising.jl:68-195), shared by the host builder
// and the device builder kernel. One (edge, side) site of a node: side 0 (the
// edge's first spin i) carries B = exp(beta w s_i s_j'), side 1 the identity;
// the field vector f = exp(beta h s) multiplies the leaf and first sites.
// s_node / s_other: the two spins' values (1, -1, a pinned value 0).
// kind: 0 leaf (1,2,1), 1 first (1,2,2), 2 last (2,2,1), 3 interior (2,2,2),
// 4 isolated spin (1,1,1) = 2 cosh(beta h). Returns the entry count.
// ---------------------------------------------------------------------------
CATN_HD int ising_site_entries(int kind, int side, const double* s_node, const double* s_other,
                              double w, double h, double beta, double* site) {
  if (kind == 4) {
    site[0] = exp(beta * h) + exp(-beta * h);
    return 1;
  }
  double mat[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      mat[a][b] = side == 0 ? exp(beta * w * (s_node[a] * s_other[b])) : (a == b ? 1.0 : 0.0);
  const double f[2] = {exp(beta * h * s_node[0]), exp(beta * h * s_node[1])};
  if (kind == 0) {                  // column sums of field * Q
    site[0] = f[0] * mat[0][0] + f[1] * mat[1][0];
    site[1] = f[0] * mat[0][1] + f[1] * mat[1][1];
    return 2;
  }
  if (kind == 1) {                  // (field * Q)' as (1,2,2): site[a + 2b] = (fQ)[b, a]
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) site[a + 2 * b] = f[b] * mat[b][a];
    return 4;
  }
  if (kind == 2) {                  // Q as (2,2,1)
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) site[a + 2 * b] = mat[a][b];
    return 4;
  }
  for (int q = 0; q < 8; ++q) site[q] = 0.0;   // interior: t3[a, s, b] = delta_ab Q[a, s]
  for (int a = 0; a < 2; ++a)
    for (int sv = 0; sv < 2; ++sv) site[a + 2 * sv + 4 * a] = mat[a][sv];
  return 8;
}

// spins of node `id` under pinning (src/ising.jl:273: the pinned value is a
// ghost 0, not an exclusion)
CATN_HD void ising_spins(int id, int pos1, int val1, int pos2, int val2, double* s) {
  s[0] = 1.0;
  s[1] = -1.0;
  if (id == pos1) s[val1 - 1] = 0.0;
  else if (id == pos2) s[val2 - 1] = 0.0;
}

// ---------------------------------------------------------------------------
// interpreter
// ---------------------------------------------------------------------------
struct Exec {
  WG g;
  InstState* st;
  Scratch sc;
  int64_t top;     // arena bump pointer (uniform register copy)
  int status;
  int bucket;      // size bucket of the op's main SVD (-1: none)
  double flops, bytes;   // algorithmic work of this launch segment
  CATN_D void work(double f, double elems) { flops += f; bytes += 8.0 * elems; }
  // tape record of the current op: lane 0 writes it straight into the
  // instance's tape when the instance records (rp is null elsewhere)
  TapeRec* rp;
  int tape_len;
  CATN_D void rec_begin(int code) {
    rp = nullptr;
    if (!st->record) return;
    if (tape_len >= st->tape_cap) { status = STATUS_TAPE_FULL; return; }
    if (g.tid != 0) return;
    rp = st->tape + tape_len;
    rp->code = code; rp->nin = 0; rp->nout = 0; rp->k = 0; rp->path = 0; rp->flag = 0;
    rp->aux0 = 0; rp->aux1 = g.lds_cap; rp->val = 0.0; rp->val2 = 0.0;
  }
  CATN_D void rec_in(const SiteDesc& d) {
    if (!rp) return;
    TapeSite& t = rp->in[rp->nin++];
    t.off = d.off; t.dl = d.dl; t.d = d.d; t.dr = d.dr;
  }
  CATN_D void rec_out(int64_t off, int dl, int d, int dr) {
    if (!rp) return;
    TapeSite& t = rp->out[rp->nout++];
    t.off = off; t.dl = dl; t.d = d; t.dr = dr;
  }
  CATN_D void rec_set(int k, int path, int flag, int aux0, double val) {
    if (!rp) return;
    rp->k = k; rp->path = path; rp->flag = flag; rp->aux0 = aux0; rp->val = val;
  }
  CATN_D void rec_commit() {
    if (st->record && status == STATUS_OK) ++tape_len;
  }

  // Operand descriptors of the current op, read by every lane before any
  // lane can rewrite them: lane 0 publishes the op's output descriptors at
  // the end of the op, and an op without an internal barrier would otherwise
  // let a lagging wavefront read a descriptor its first wavefront already
  // replaced.
  int op_slot[3];
  SiteDesc op_desc[3];
  CATN_D SiteDesc load_site(int s) const {
    volatile SiteDesc* p = st->sites + s;
    SiteDesc d;
    d.off = p->off; d.dl = p->dl; d.d = p->d; d.dr = p->dr; d.live = p->live;
    return d;
  }
  CATN_D void snapshot(const Op& op) {
    const int sl[3] = {op.a, op.b, op.c};
    for (int q = 0; q < 3; ++q) {
      op_slot[q] = (sl[q] >= 0 && sl[q] < st->nsites) ? sl[q] : -1;
      if (op_slot[q] >= 0) op_desc[q] = load_site(op_slot[q]);
    }
    g.sync();
  }
  CATN_D SiteDesc site(int s) const {
    for (int q = 0; q < 3; ++q)
      if (op_slot[q] == s) return op_desc[q];
    SiteDesc d;
    d.off = 0; d.dl = d.d = d.dr = 0; d.live = 0;
    return d;
  }

  // FP32 storage: an op reads FP64 copies of its input tensors (made before
  // the op runs) and writes its outputs into FP64 staging buffers, which are
  // rounded into the FP32 arena after the op. Both live in the op's scratch.
  // With FP64 storage the arena is read and written in place.
  static constexpr int kStage = 6;
  int nconv, nstage;
  int64_t conv_off[kStage], stage_off[kStage], stage_n[kStage];
  double* conv_buf[kStage];
  double* stage_buf[kStage];
  CATN_D float* arena_f() const { return (float*)st->arena; }
  // Singular values below noise_floor() * ||A||_F count as zero when the
  // tensors are stored in FP32: the storage rounds every element (relative
  // 2^-24), so much smaller singular values are rounding noise. Keeping them
  // would grow bond dimensions in exact runs and, through the 1/sigma terms
  // of the SVD adjoint, turn the FP32 orthogonality error of U into O(1)
  // gradient errors. 6-pass products (2): 2^-22, 4x the storage resolution. A
  // floor of 2^-20 dropped real spectrum of frustrated lattices: on +-J
  // lattices of width 8..22 at beta in [0.2, 1.2] it cost up to 3.6e-6 in
  // lnZ/N against FP64 (1.5e-6 on a 10 x 10 lattice at beta 1.14), 2^-22
  // keeps those within 1e-7. 3-pass products (1): 2^-18; a floor at their own
  // product error (2^-16 and above) drops real spectrum: 2^-13 cost 2e-4 in
  // lnZ/N on a 14 x 14 lattice at beta 1.1, 2^-18 keeps it within 3.4e-6 on
  // 8..16-wide lattices. With FP64 storage the reference's cutoff alone
  // decides.
  CATN_D double noise_floor() const {
    return st->f32 == 2 ? 0x1p-22 : st->f32 == 1 ? 0x1p-18 : 0.0;
  }
  // values per element: 2 (re, im) for complex networks
  CATN_D int ew() const { return st->cplx ? 2 : 1; }
  // FP64 copy of the tensor version at arena offset `off` (n elements) into
  // scratch; every lane calls it. False when the scratch is exhausted.
  CATN_D bool convert_in(int64_t off, int64_t n) {
    for (int q = 0; q < nconv; ++q)
      if (conv_off[q] == off) return true;
    if (nconv >= kStage) { status = STATUS_UNSUPPORTED; return false; }
    n *= ew();
    double* buf = sc.alloc(n > 0 ? n : 1);
    if (sc.fail) { status = STATUS_SCRATCH_FULL; return false; }
    const float* src = arena_f() + off * ew();
    for (int64_t i = g.tid; i < n; i += g.nthr) buf[i] = (double)src[i];
    conv_off[nconv] = off;
    conv_buf[nconv] = buf;
    ++nconv;
    return true;
  }
  CATN_D const double* ptr(const SiteDesc& d) const { return view(d.off); }
  CATN_D const double* view(int64_t off) const {
    if (!st->f32) return st->arena + off;
    for (int q = 0; q < nconv; ++q)
      if (conv_off[q] == off) return conv_buf[q];
    return conv_buf[0];   // not reached: every op input is converted before the op
  }
  CATN_D int64_t alloc(int64_t n) {
    int64_t a = (top + 31) & ~int64_t(31);
    if (a + n > st->arena_cap) { status = STATUS_ARENA_FULL; return 0; }
    top = a + n;
    if (st->f32) {
      if (nstage >= kStage) { status = STATUS_UNSUPPORTED; return 0; }
      double* buf = sc.alloc(n > 0 ? n * ew() : 1);
      if (sc.fail) { status = STATUS_SCRATCH_FULL; return 0; }
      stage_off[nstage] = a;
      stage_n[nstage] = n;
      stage_buf[nstage] = buf;
      ++nstage;
    }
    return a;
  }
  // where an op writes the output allocated at `off`
  CATN_D double* out(int64_t off) const {
    if (!st->f32) return st->arena + off;
    for (int q = 0; q < nstage; ++q)
      if (stage_off[q] == off) return stage_buf[q];
    return stage_buf[0];   // not reached: outputs come from alloc()
  }
  // round the staged outputs into the FP32 arena (every lane; after the op)
  CATN_D void flush_out() {
    if (!st->f32) return;
    float* dst = arena_f();
    const int w = ew();
    for (int q = 0; q < nstage; ++q)
      for (int64_t i = g.tid; i < stage_n[q] * w; i += g.nthr)
        dst[stage_off[q] * w + i] = (float)stage_buf[q][i];
  }
  CATN_D void set_site(int s, int64_t off, int dl, int d, int dr) {
    if (g.tid == 0) {
      volatile SiteDesc* p = st->sites + s;
      p->off = off; p->dl = dl; p->d = d; p->dr = dr; p->live = 1;
    }
  }
  CATN_D void acc(volatile double* dst, double v) {
    if (g.tid == 0) *dst = *dst + v;
  }
};

// count kept singular values: tsvd rule (src/linalg_utils.jl:24-25)
// textbook operation counts (Golub & Van Loan): thin SVD (R-SVD) of an
// m x n matrix, Householder QR of a tall m x n matrix
CATN_D double svd_flops(double m, double n) {
  const double M = m > n ? m : n, N = m > n ? n : m;
  return 6.0 * M * N * N + 20.0 * N * N * N;
}
CATN_D double qr_flops(double m, double n) { return 2.0 * m * n * n - 2.0 / 3.0 * n * n * n; }

CATN_D int keep_tsvd(const double* S, int r, double cutoff, int maxdim) {
  int c = 0;
  for (int i = 0; i < r; ++i) c += (S[i] > cutoff);
  if (c == 0) return 1;
  if (c > maxdim) c = maxdim;
  if (c > r) c = r;
  return c;
}

// cano_to! rule: tsvd(cutoff=0) then keep S > node.cutoff, all if none
// (src/mps_node.jl:79-82)
CATN_D int keep_cano(const double* S, int r, double cutoff) {
  int c0 = 0;
  for (int i = 0; i < r; ++i) c0 += (S[i] > 0.0);
  if (c0 == 0) c0 = 1;
  int c = 0;
  for (int i = 0; i < c0; ++i) c += (S[i] > cutoff);
  return c == 0 ? c0 : c;
}

CATN_D double fro_of(const double* S, int r) {
  double s = 0.0;
  for (int i = 0; i < r; ++i) s += S[i] * S[i];
  return sqrt(s);
}

CATN_D double tail_sum(const double* S, int from, int r) {
  double s = 0.0;
  for (int i = from; i < r; ++i) s += S[i];
  return s;
}

// Split A (m x n) into A ~= L * R with the reference's truncation rule:
// rule 0 = tsvd(cutoff, maxdim) (src/linalg_utils.jl:24-25), rule 1 = the
// cano_to! rule (src/mps_node.jl:79-82). iso_left: L = U (isometry), R = S V';
// else L = U S, R = V' (isometry rows). Certified full-rank steps are
// factorised by QR (wg_factor_fullrank); the others by the Jacobi SVD.
// L is m x k (ld m), R is k x n (ld k), both in scratch.
struct Split {
  int k;
  double* L;
  double* R;
  double disc;
  int qr;   // 1 when the certified QR factorisation was used
};

CATN_DN bool wg_split(Exec& E, const double* A, int m, int n, int64_t lda, bool iso_left, int rule,
                      double cutoff, int maxdim, Split& sp) {
  const WG& g = E.g;
  Factor F;
  const int maxd = rule == 1 ? (1 << 30) : maxdim;
  const int r0 = m < n ? m : n;
  E.bucket = size_bucket(r0);
  const double noise = E.noise_floor();
  if (wg_factor_fullrank(g, E.sc, A, m, n, lda, cutoff, maxd, iso_left, F, noise)) {
    sp.k = F.r; sp.L = F.L; sp.R = F.Rt; sp.disc = 0.0; sp.qr = 1;
    const double tm = m > n ? m : n;
    E.work(qr_flops(tm, r0) + r0 * (double)r0 * r0 / 3.0 + (iso_left != (m < n) ? qr_flops(tm, r0) : 0.0),
           0.0);
    return true;
  }
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return false; }
  SvdResult R;
  int sw = 0;
  if (wg_svd(g, E.sc, A, m, n, lda, R, &sw) < 0) {
    E.status = E.sc.fail ? STATUS_SCRATCH_FULL : STATUS_NOCONV;
    return false;
  }
  E.work(svd_flops(m, n), 0.0);
  if (noise > 0.0) cutoff = fmax(cutoff, noise * fro_of(R.S, R.r));
  const int k = rule == 1 ? keep_cano(R.S, R.r, cutoff) : keep_tsvd(R.S, R.r, cutoff, maxdim);
  double* L = E.sc.alloc((int64_t)m * k);
  double* Rm = E.sc.alloc((int64_t)k * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return false; }
  for (int64_t idx = g.tid; idx < (int64_t)m * k; idx += g.nthr) {
    const int p = (int)(idx / m);
    L[idx] = iso_left ? R.U[idx] : R.U[idx] * R.S[p];
  }
  for (int64_t idx = g.tid; idx < (int64_t)k * n; idx += g.nthr) {
    const int p = (int)(idx % k);
    const int64_t c = idx / k;
    const double v = R.V[c + (int64_t)n * p];
    Rm[idx] = iso_left ? R.S[p] * v : v;
  }
  g.sync();
  sp.k = k; sp.L = L; sp.R = Rm; sp.disc = tail_sum(R.S, k, R.r); sp.qr = 0;
  return true;
}

// swap!'s randomized-SVD branch: rsvd(M, chi, 10, 10) (src/mps_node.jl:155)
// with the reference's hand-over of the factors: going right U | S V',
// going left U S | V'. No cutoff truncation, no discarded weight (the
// reference returns 0.0 for this branch); path 2 in the tape.
CATN_DN bool swap_rsvd(Exec& E, const double* M, int m, int n, bool iso_left, int chi, Split& sp) {
  const WG& g = E.g;
  double *U, *S, *V;
  const int kk = wg_rsvd(g, E.sc, M, m, n, m, chi, 10, 10, (uint64_t)E.st->seed,
                         (uint64_t)E.top, &U, &S, &V);
  if (kk < 0) {
    E.status = E.sc.fail ? STATUS_SCRATCH_FULL : (m < (n < 10 * chi ? n : 10 * chi)
                                                       ? STATUS_UNSUPPORTED : STATUS_NOCONV);
    return false;
  }
  double* L = E.sc.alloc((int64_t)m * kk);
  double* R = E.sc.alloc((int64_t)kk * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return false; }
  for (int64_t idx = g.tid; idx < (int64_t)m * kk; idx += g.nthr) {
    const int p = (int)(idx / m);
    L[idx] = iso_left ? U[idx] : U[idx] * S[p];
  }
  for (int64_t idx = g.tid; idx < (int64_t)kk * n; idx += g.nthr) {
    const int p = (int)(idx % kk);
    const int64_t c = idx / kk;
    const double v = V[c + (int64_t)n * p];
    R[idx] = iso_left ? S[p] * v : v;
  }
  g.sync();
  const double p = n < 10.0 * chi ? n : 10.0 * chi;
  E.work(2.0 * m * n * p * (1 + 2 * 10) + 2.0 * (10 + 1) * qr_flops(m, p) + 2.0 * p * m * n +
             svd_flops(p, n) + 2.0 * m * p * kk, 0.0);
  sp.k = kk; sp.L = L; sp.R = R; sp.disc = 0.0; sp.qr = 2;
  return true;
}

CATN_DN void op_cano_r(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a), b = E.site(op.b);
  const int m = a.dl * a.d, n = a.dr;
  if (b.dl != n) { E.status = STATUS_SHAPE; return; }
  Split sp;
  if (!wg_split(E, E.ptr(a), m, n, m, true, 1, op.x, 0, sp)) return;
  const int k = sp.k;
  const int64_t nb = (int64_t)b.d * b.dr;
  const int64_t offa = E.alloc((int64_t)m * k);
  const int64_t offb = E.alloc((int64_t)k * nb);
  if (E.status) return;
  wg_copy(g, E.out(offa), sp.L, (int64_t)m * k);
  // new b = R * b   (k x n) * (n x nb)
  wg_mm(g, k, (int)nb, n, sp.R, k, E.ptr(b), n, E.out(offb), k);
  E.work(2.0 * k * n * (double)nb, (double)m * n + (double)n * nb + (double)m * k + (double)k * nb);
  E.rec_begin(OP_CANO_R); E.rec_in(a); E.rec_in(b);
  E.rec_out(offa, a.dl, a.d, k); E.rec_out(offb, k, b.d, b.dr);
  E.rec_set(k, sp.qr, 0, 0, op.x);
  E.rec_commit();
  E.set_site(op.a, offa, a.dl, a.d, k);
  E.set_site(op.b, offb, k, b.d, b.dr);
  g.sync();
}

CATN_DN void op_cano_l(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a), b = E.site(op.b);   // b = site i-1
  const int dl = a.dl, d = a.d, dr = a.dr;
  if (b.dr != dl) { E.status = STATUS_SHAPE; return; }
  const int m = d * dr, n = dl;
  double* Mt = E.sc.alloc((int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const double* ap = E.ptr(a);
  // Mt[(y + d*z), x] = a[x, y, z]   (permutedims(a, (2,3,1)) reshaped)
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int64_t row = idx % m;
    const int x = (int)(idx / m);
    const int y = (int)(row % d), z = (int)(row / d);
    Mt[idx] = ap[x + (int64_t)dl * (y + (int64_t)d * z)];
  }
  g.sync();
  Split sp;
  if (!wg_split(E, Mt, m, n, m, true, 1, op.x, 0, sp)) return;
  const int k = sp.k;
  const int64_t rb = (int64_t)b.dl * b.d;
  const int64_t offa = E.alloc((int64_t)k * m);
  const int64_t offb = E.alloc(rb * k);
  if (E.status) return;
  double* na = E.out(offa);
  double* nbp = E.out(offb);
  // new a[p, y, z] = L[y + d*z, p]
  for (int64_t idx = g.tid; idx < (int64_t)k * m; idx += g.nthr) {
    const int p = (int)(idx % k);
    const int64_t row = idx / k;
    na[idx] = sp.L[row + (int64_t)m * p];
  }
  // new b[u, p] = sum_c b[u, c] R[p, c]: (rb x n) (n x k)^T, summed over c in
  // order (tiled when large enough)
  wg_gemm_any(g, (int)rb, k, n, 1.0, E.ptr(b), rb, false, sp.R, k, true, 0.0, nbp, rb);
  E.work(2.0 * (double)rb * n * k, (double)m * n + (double)rb * n + (double)k * m + (double)rb * k);
  E.rec_begin(OP_CANO_L); E.rec_in(a); E.rec_in(b);
  E.rec_out(offa, k, d, dr); E.rec_out(offb, b.dl, b.d, k);
  E.rec_set(k, sp.qr, 0, 0, op.x);
  E.rec_commit();
  E.set_site(op.a, offa, k, d, dr);
  E.set_site(op.b, offb, b.dl, b.d, k);
  g.sync();
}

CATN_DN void op_swap(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc L = E.site(op.a), Rr = E.site(op.b);
  const int d0 = L.dl, d2 = L.d, dd = L.dr, d1 = Rr.d, d3 = Rr.dr;
  if (Rr.dl != dd) { E.status = STATUS_SHAPE; return; }
  const int m = d0 * d1, n = d2 * d3;
  // randomized SVD for large thetas (src/mps_node.jl:154-156)
  const bool use_rsvd = op.f && ((m > op.g && n > op.g) || m > op.h || n > op.h);
  double* M = E.sc.alloc((int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const double* tl = E.ptr(L);
  const double* tr = E.ptr(Rr);
  // M[(i + d0*a), (j + d2*b)] = sum_k tl[i, j, k] tr[k, a, b]: the GEMM
  // (tl as (d0 d2) x dd) * (tr as dd x (d1 d3)) written through the index map
  const int64_t tth = now_ticks();
  if (E.st->f32) {
    // FP32 storage: split-BF16 MFMA products straight from the FP32 arena
    IdxTheta o;
    o.d0 = d0; o.d1 = d1; o.d2 = d2; o.m = m;
    if (E.st->f32 == 2)
      wg_gemm_bf16<3>(g, d0 * d2, d1 * d3, dd, E.arena_f() + L.off, (int64_t)d0 * d2,
                      E.arena_f() + Rr.off, dd, M, o);
    else
      wg_gemm_bf16<2>(g, d0 * d2, d1 * d3, dd, E.arena_f() + L.off, (int64_t)d0 * d2,
                      E.arena_f() + Rr.off, dd, M, o);
  } else if ((int64_t)m * n >= 2048 && dd >= 8 && g.lds_cap >= 2 * kTK * kTM + 64) {
    IdxTheta o;
    o.d0 = d0; o.d1 = d1; o.d2 = d2; o.m = m;
    wg_gemm_t(g, d0 * d2, d1 * d3, dd, 1.0, tl, (int64_t)d0 * d2, false, tr, dd, false, 0.0, M, o);
  } else
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int64_t row = idx % m, col = idx / m;
    const int i = (int)(row % d0), a = (int)(row / d0);
    const int j = (int)(col % d2), b = (int)(col / d2);
    const double* x = tl + i + (int64_t)d0 * j;
    const double* y = tr + (int64_t)dd * (a + (int64_t)d1 * b);
    double s = 0.0;
    for (int k = 0; k < dd; ++k) s += x[(int64_t)d0 * d2 * k] * y[k];
    M[idx] = s;
  }
  g.sync();
  sub_tick(g, 7, tth);
  Split sp;
  if (use_rsvd) {
    if (!swap_rsvd(E, M, m, n, op.c != 0, op.d, sp)) return;
  } else if (!wg_split(E, M, m, n, m, op.c != 0, 0, op.x, op.d, sp)) {
    return;
  }
  const int k = sp.k;
  if (op.e) E.acc(&E.st->err, sp.disc);
  const int64_t offl = E.alloc((int64_t)m * k);
  const int64_t offr = E.alloc((int64_t)k * n);
  if (E.status) return;
  wg_copy(g, E.out(offl), sp.L, (int64_t)m * k);
  wg_copy(g, E.out(offr), sp.R, (int64_t)k * n);
  E.work(2.0 * m * n * (double)dd,
         (double)d0 * d2 * dd + (double)dd * d1 * d3 + (double)m * k + (double)k * n);
  E.rec_begin(OP_SWAP); E.rec_in(L); E.rec_in(Rr);
  E.rec_out(offl, d0, d1, k); E.rec_out(offr, k, d2, d3);
  E.rec_set(k, sp.qr, op.c, op.d, op.x);
  E.rec_commit();
  E.set_site(op.a, offl, d0, d1, k);
  E.set_site(op.b, offr, k, d2, d3);
  g.sync();
}

CATN_DN void op_permute321(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a);
  const int64_t tot = (int64_t)a.dl * a.d * a.dr;
  const int64_t off = E.alloc(tot);
  if (E.status) return;
  const double* s = E.ptr(a);
  double* dst = E.out(off);
  // dst[z, y, x] = s[x, y, z]; dst dims (dr, d, dl)
  for (int64_t idx = g.tid; idx < tot; idx += g.nthr) {
    const int z = (int)(idx % a.dr);
    const int64_t rest = idx / a.dr;
    const int y = (int)(rest % a.d), x = (int)(rest / a.d);
    dst[idx] = s[x + (int64_t)a.dl * (y + (int64_t)a.d * z)];
  }
  E.work(0.0, 2.0 * tot);
  E.rec_begin(OP_PERMUTE321); E.rec_in(a); E.rec_out(off, a.dr, a.d, a.dl);
  E.rec_commit();
  E.set_site(op.a, off, a.dr, a.d, a.dl);
  g.sync();
}

// a column-major (dl, d, dr) block is also (dl, D, d * dr / D), D = op.d: descriptor
// only (no tape record: the data, and so its adjoint, stay at the same offset)
CATN_DN void op_reshape(Exec& E, const Op& op) {
  const SiteDesc a = E.site(op.a);
  const int64_t dd = (int64_t)a.d * a.dr;
  if (op.d <= 0 || dd % op.d != 0) { E.status = STATUS_SHAPE; return; }
  E.set_site(op.a, a.off, a.dl, op.d, (int)(dd / op.d));
  E.g.sync();
}

CATN_DN void op_eat_a(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a), b = E.site(op.b);
  const int64_t n = (int64_t)a.dl * a.d * a.dr;
  if (n != (int64_t)b.dl * b.d * b.dr) { E.status = STATUS_SHAPE; return; }
  const double* x = E.ptr(a);
  const double* y = E.ptr(b);
  double part = 0.0;
  for (int64_t i = g.tid; i < n; i += g.nthr) part += x[i] * y[i];
  const double r = wg_sum(g, part);
  const double ar = fabs(r);
  E.work(2.0 * n, 2.0 * n);
  if (ar > op.x) {
    E.acc(&E.st->lnZ, log(ar));
    if (g.tid == 0) E.st->psi = E.st->psi * (r / ar);
  }
  E.rec_begin(OP_EAT_A); E.rec_in(a); E.rec_in(b);
  E.rec_set(0, 0, ar > op.x ? 1 : 0, 0, r);
  E.rec_commit();
  if (g.tid == 0) E.st->sites[op.a].live = 0;
  g.sync();
}

CATN_D double site_norm(const WG& g, const double* x, int64_t n, int method) {
  if (method == 1) {
    double part = 0.0;
    for (int64_t i = g.tid; i < n; i += g.nthr) part += x[i] * x[i];
    return sqrt(wg_sum(g, part));
  } else if (method == 2) {
    double part = 0.0;
    for (int64_t i = g.tid; i < n; i += g.nthr) part = fmax(part, fabs(x[i]));
    return wg_max(g, part);
  }
  return 1.0;
}

// Normalise site `slot` per norm_method into a new tensor version (the
// previous version stays for the reverse sweep); log(norm) is added to lnZ
// unless norm <= cutoff (src/mps_node.jl:506-527, :558-574).
CATN_D void scale_normalize(Exec& E, int slot, const SiteDesc& s, int method, double cutoff) {
  const WG& g = E.g;
  const double* x = E.ptr(s);
  const int64_t n = (int64_t)s.dl * s.d * s.dr;
  const double nrm = site_norm(g, x, n, method);
  E.work(3.0 * n, 2.0 * n);
  const int64_t off = E.alloc(n);
  if (E.status) return;
  double* y = E.out(off);
  const bool counted = nrm > cutoff;   // src/mps_node.jl:522-524, :569-571
  const double inv = counted ? 1.0 / nrm : 1.0;
  for (int64_t i = g.tid; i < n; i += g.nthr) y[i] = x[i] * inv;
  if (counted) E.acc(&E.st->lnZ, log(nrm));
  E.rec_begin(OP_NORMALIZE); E.rec_in(s); E.rec_out(off, s.dl, s.d, s.dr);
  E.rec_set(0, 0, counted ? 1 : 0, method, nrm);
  E.rec_commit();
  E.set_site(slot, off, s.dl, s.d, s.dr);
  g.sync();
}

CATN_DN void op_eat_b(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc p = E.site(op.a), t = E.site(op.b), j = E.site(op.c);
  const int D = t.dl, d = t.d;
  if (t.dr != 1 || j.d != d || p.dr != D) { E.status = STATUS_SHAPE; return; }
  double* mat = E.sc.alloc(D);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const double* tp = E.ptr(t);
  const double* jp = E.ptr(j);
  for (int i = g.tid; i < D; i += g.nthr) {
    double s = 0.0;
    for (int q = 0; q < d; ++q) s += tp[i + (int64_t)D * q] * jp[q];
    mat[i] = s;
  }
  g.sync();
  const int64_t rows = (int64_t)p.dl * p.d;
  const int64_t off = E.alloc(rows);
  if (E.status) return;
  double* np_ = E.out(off);
  const double* pp = E.ptr(p);
  for (int64_t r = g.tid; r < rows; r += g.nthr) {
    double s = 0.0;
    for (int q = 0; q < D; ++q) s += pp[r + rows * q] * mat[q];
    np_[r] = s;
  }
  E.work(2.0 * D * d + 2.0 * rows * D, (double)D * d + d + rows * D + rows);
  E.rec_begin(OP_EAT_B); E.rec_in(p); E.rec_in(t); E.rec_in(j); E.rec_out(off, p.dl, p.d, 1);
  E.rec_commit();
  E.set_site(op.a, off, p.dl, p.d, 1);
  if (g.tid == 0) { E.st->sites[op.b].live = 0; E.st->sites[op.c].live = 0; }
  g.sync();
}

CATN_DN void op_eat_c(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc p = E.site(op.a), t = E.site(op.b), h = E.site(op.c);
  const int Dl = t.dl, d = t.d, Dr = h.dr;
  if (t.dr != 1 || h.dl != 1 || h.d != d || p.dr != Dl) { E.status = STATUS_SHAPE; return; }
  double* mat = E.sc.alloc((int64_t)Dl * Dr);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  wg_mm(g, Dl, Dr, d, E.ptr(t), Dl, E.ptr(h), d, mat, Dl);
  g.sync();
  const int64_t rows = (int64_t)p.dl * p.d;
  const int64_t off = E.alloc(rows * Dr);
  if (E.status) return;
  wg_mm(g, (int)rows, Dr, Dl, E.ptr(p), rows, mat, Dl, E.out(off), rows);
  E.work(2.0 * Dl * Dr * d + 2.0 * rows * Dl * Dr,
         (double)Dl * d + (double)d * Dr + rows * Dl + rows * Dr);
  E.rec_begin(OP_EAT_C); E.rec_in(p); E.rec_in(t); E.rec_in(h); E.rec_out(off, p.dl, p.d, Dr);
  E.rec_commit();
  E.set_site(op.a, off, p.dl, p.d, Dr);
  if (g.tid == 0) { E.st->sites[op.b].live = 0; E.st->sites[op.c].live = 0; }
  g.sync();
}

CATN_DN void op_normalize(Exec& E, const Op& op) {
  SiteDesc s = E.site(op.a);
  scale_normalize(E, op.a, s, op.d, op.x);
}

CATN_DN void op_merge_fuse(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a), b = E.site(op.b);
  if (a.dr != b.dl) { E.status = STATUS_SHAPE; return; }
  const int m = a.dl * a.d, n = b.d * b.dr, k = a.dr;
  const int64_t off = E.alloc((int64_t)m * n);
  if (E.status) return;
  wg_mm(g, m, n, k, E.ptr(a), m, E.ptr(b), k, E.out(off), m);
  E.work(2.0 * m * n * (double)k, (double)m * k + (double)k * n + (double)m * n);
  E.rec_begin(OP_MERGE_FUSE); E.rec_in(a); E.rec_in(b); E.rec_out(off, a.dl, a.d * b.d, b.dr);
  E.rec_commit();
  E.set_site(op.a, off, a.dl, a.d * b.d, b.dr);
  if (g.tid == 0) E.st->sites[op.b].live = 0;
  g.sync();
}

// thin QR of a tall m x n (m > n) matrix held in Wk (overwritten):
// Q (m x n, ld m) and R (n x n, ld n) into fresh scratch
CATN_DN bool wg_thin_qr(const WG& g, Scratch& sc, double* Wk, int m, int n, double** Qo, double** Ro) {
  double* tau = sc.alloc(n);
  double* R = sc.alloc((int64_t)n * n);
  double* Q = sc.alloc((int64_t)m * n);
  double* Tall = sc.alloc((int64_t)kQBMax * n);
  if (sc.fail) return false;
  const bool blocked = wg_geqrf(g, sc, Wk, m, n, m, tau, Tall);
  if (sc.fail) return false;
  for (int64_t idx = g.tid; idx < (int64_t)n * n; idx += g.nthr) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    R[idx] = (i <= j) ? Wk[i + (int64_t)j * m] : 0.0;
  }
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int i = (int)(idx % m), j = (int)(idx / m);
    Q[idx] = (i == j) ? 1.0 : 0.0;
  }
  g.sync();
  wg_apply_q_blocked(g, sc, Wk, m, n, m, tau, blocked ? Tall : nullptr, Q, n, m, true);
  *Qo = Q;
  *Ro = R;
  return true;
}

CATN_DN void op_cut(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc A = E.site(op.a), B = E.site(op.b);
  const int dal = A.dl, d = A.d, dar = A.dr, dbl = B.dl, dbr = B.dr;
  if (B.d != d) { E.status = STATUS_SHAPE; return; }
  const int mi = dal * dar, mj = dbl * dbr;
  double* Mi = E.sc.alloc((int64_t)mi * d);
  double* Mj = E.sc.alloc((int64_t)mj * d);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  // mat[a + dal*c, b] = A[a, b, c]   (permutedims(A,(1,3,2)) reshaped)
  const double* ap = E.ptr(A);
  const double* bp = E.ptr(B);
  for (int64_t idx = g.tid; idx < (int64_t)mi * d; idx += g.nthr) {
    const int64_t row = idx % mi;
    const int b = (int)(idx / mi);
    const int a = (int)(row % dal), c = (int)(row / dal);
    Mi[idx] = ap[a + (int64_t)dal * (b + (int64_t)d * c)];
  }
  for (int64_t idx = g.tid; idx < (int64_t)mj * d; idx += g.nthr) {
    const int64_t row = idx % mj;
    const int b = (int)(idx / mj);
    const int a = (int)(row % dbl), c = (int)(row / dbl);
    Mj[idx] = bp[a + (int64_t)dbl * (b + (int64_t)d * c)];
  }
  g.sync();
  const bool opt = op.f != 0;
  const bool fl = opt && mi > d, fr = opt && mj > d;
  double *Qi = nullptr, *Ri = Mi, *Qj = nullptr, *Rj = Mj;
  int ri_rows = mi, rj_rows = mj;
  if (fl) { if (!wg_thin_qr(g, E.sc, Mi, mi, d, &Qi, &Ri)) { E.status = STATUS_SCRATCH_FULL; return; } ri_rows = d; }
  if (fr) { if (!wg_thin_qr(g, E.sc, Mj, mj, d, &Qj, &Rj)) { E.status = STATUS_SCRATCH_FULL; return; } rj_rows = d; }
  // merged = ri * rj^T  (ri_rows x rj_rows), non-conjugating
  double* Mg = E.sc.alloc((int64_t)ri_rows * rj_rows);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const int ldri = fl ? d : mi, ldrj = fr ? d : mj;
  // tiled (LDS-staged) when large enough, else one lane per entry; both sum
  // q = 0 .. d-1 in order
  wg_gemm_any(g, ri_rows, rj_rows, d, 1.0, Ri, ldri, false, Rj, ldrj, true, 0.0, Mg, ri_rows);
  SvdResult R;
  int sw = 0;
  if (wg_svd(g, E.sc, Mg, ri_rows, rj_rows, ri_rows, R, &sw) < 0) {
    E.status = E.sc.fail ? STATUS_SCRATCH_FULL : STATUS_NOCONV; return;
  }
  E.bucket = size_bucket(R.r);
  const double cut = fmax(op.x, E.noise_floor() * fro_of(R.S, R.r));
  const int nk = keep_tsvd(R.S, R.r, cut, 1 << 30);
  int myd = nk < op.d ? nk : op.d;
  if (myd == 0) myd = 1;
  E.acc(&E.st->err, tail_sum(R.S, myd, R.r));
  // mati' = U sqrt(S) (ri_rows x myd), matj' = V sqrt(S) (rj_rows x myd)
  double* Ui = E.sc.alloc((int64_t)ri_rows * myd);
  double* Uj = E.sc.alloc((int64_t)rj_rows * myd);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  for (int64_t idx = g.tid; idx < (int64_t)ri_rows * myd; idx += g.nthr) {
    const int p = (int)(idx / ri_rows);
    Ui[idx] = R.U[idx] * sqrt(R.S[p]);
  }
  for (int64_t idx = g.tid; idx < (int64_t)rj_rows * myd; idx += g.nthr) {
    const int p = (int)(idx / rj_rows);
    Uj[idx] = R.V[idx] * sqrt(R.S[p]);
  }
  g.sync();
  double* Fi = Ui;
  double* Fj = Uj;
  if (fl) {
    Fi = E.sc.alloc((int64_t)mi * myd);
    if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
    wg_mm(g, mi, myd, d, Qi, mi, Ui, d, Fi, mi);
  }
  if (fr) {
    Fj = E.sc.alloc((int64_t)mj * myd);
    if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
    wg_mm(g, mj, myd, d, Qj, mj, Uj, d, Fj, mj);
  }
  g.sync();
  const int64_t offa = E.alloc((int64_t)dal * myd * dar);
  const int64_t offb = E.alloc((int64_t)dbl * myd * dbr);
  if (E.status) return;
  double* na = E.out(offa);
  double* nb = E.out(offb);
  // A'[a, p, c] = Fi[a + dal*c, p]
  for (int64_t idx = g.tid; idx < (int64_t)dal * myd * dar; idx += g.nthr) {
    const int a = (int)(idx % dal);
    const int64_t rest = idx / dal;
    const int p = (int)(rest % myd), c = (int)(rest / myd);
    na[idx] = Fi[a + (int64_t)dal * c + (int64_t)mi * p];
  }
  for (int64_t idx = g.tid; idx < (int64_t)dbl * myd * dbr; idx += g.nthr) {
    const int a = (int)(idx % dbl);
    const int64_t rest = idx / dbl;
    const int p = (int)(rest % myd), c = (int)(rest / myd);
    nb[idx] = Fj[a + (int64_t)dbl * c + (int64_t)mj * p];
  }
  E.work((fl ? 2.0 * qr_flops(mi, d) + 2.0 * mi * d * (double)myd : 0.0) +
             (fr ? 2.0 * qr_flops(mj, d) + 2.0 * mj * d * (double)myd : 0.0) +
             2.0 * ri_rows * rj_rows * (double)d + svd_flops(ri_rows, rj_rows),
         (double)mi * d + (double)mj * d + (double)mi * myd + (double)mj * myd);
  E.rec_begin(OP_CUT); E.rec_in(A); E.rec_in(B);
  E.rec_out(offa, dal, myd, dar); E.rec_out(offb, dbl, myd, dbr);
  E.rec_set(myd, 0, op.f, 0, op.x);
  E.rec_commit();
  E.set_site(op.a, offa, dal, myd, dar);
  E.set_site(op.b, offb, dbl, myd, dbr);
  if (g.tid == 0) {
    if (op.e >= 0 && op.e < E.st->nresults) E.st->results[op.e] = myd;
    E.st->last_cut = myd;
  }
  g.sync();
}

// compress! / compress_opt! two-site step (src/mps_node.jl:363-372, :407-443)
CATN_DN void op_compress(Exec& E, const Op& op, bool opt) {
  const WG& g = E.g;
  SiteDesc A = E.site(op.a), B = E.site(op.b);
  const int d0 = A.dl, d1 = A.d, dd = A.dr, d2 = B.d, d3 = B.dr;
  if (B.dl != dd) { E.status = STATUS_SHAPE; return; }
  const int m = d0 * d1, n = d2 * d3;
  double *Ql = nullptr, *Qr = nullptr;
  bool fl = false, fr = false;
  int rl = m, rr = n;
  double* M = nullptr;
  if (!opt) {
    M = E.sc.alloc((int64_t)m * n);
    if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
    wg_mm(g, m, n, dd, E.ptr(A), m, E.ptr(B), dd, M, m);
    g.sync();
  } else {
    // matl (m x dd), matr^T (n x dd)
    double* Ml = E.sc.alloc((int64_t)m * dd);
    double* Mr = E.sc.alloc((int64_t)n * dd);
    if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
    wg_copy(g, Ml, E.ptr(A), (int64_t)m * dd);
    wg_transpose(g, Mr, E.ptr(B), dd, n);
    g.sync();
    double* Rl = Ml;
    double* Rr = Mr;
    fl = m > dd;
    fr = dd < n;
    if (fl) { if (!wg_thin_qr(g, E.sc, Ml, m, dd, &Ql, &Rl)) { E.status = STATUS_SCRATCH_FULL; return; } rl = dd; }
    if (fr) { if (!wg_thin_qr(g, E.sc, Mr, n, dd, &Qr, &Rr)) { E.status = STATUS_SCRATCH_FULL; return; } rr = dd; }
    // mat = Rl * Rr^T   (rl x rr)
    M = E.sc.alloc((int64_t)rl * rr);
    if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
    wg_gemm_any(g, rl, rr, dd, 1.0, Rl, rl, false, Rr, rr, true, 0.0, M, rl);
    E.work((fl ? 2.0 * qr_flops(m, dd) : 0.0) + (fr ? 2.0 * qr_flops(n, dd) : 0.0) +
               2.0 * rl * rr * (double)dd, 0.0);
  }
  Split sp;
  if (!wg_split(E, M, rl, rr, rl, false, 0, op.x, op.d, sp)) return;
  const int k = sp.k;
  // left = Ql * (U S)  (m x k); right = (V') Qr^T  (k x n)
  double* Gl = sp.L;
  double* Gr = sp.R;
  if (fl) {
    Gl = E.sc.alloc((int64_t)m * k);
    if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
    wg_mm(g, m, k, dd, Ql, m, sp.L, dd, Gl, m);
  }
  if (fr) {
    Gr = E.sc.alloc((int64_t)k * n);
    if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
    wg_gemm_any(g, k, n, dd, 1.0, sp.R, k, false, Qr, n, true, 0.0, Gr, k);
  }
  g.sync();
  const int64_t offa = E.alloc((int64_t)m * k);
  const int64_t offb = E.alloc((int64_t)k * n);
  if (E.status) return;
  wg_copy(g, E.out(offa), Gl, (int64_t)m * k);
  wg_copy(g, E.out(offb), Gr, (int64_t)k * n);
  E.work(opt ? (fl ? 2.0 * m * dd * (double)k : 0.0) + (fr ? 2.0 * n * dd * (double)k : 0.0)
             : 2.0 * m * n * (double)dd,
         (double)m * dd + (double)dd * n + (double)m * k + (double)k * n);
  E.rec_begin(opt ? OP_COMPRESS_OPT : OP_COMPRESS); E.rec_commit();
  E.set_site(op.a, offa, d0, d1, k);
  E.set_site(op.b, offb, k, d2, d3);
  g.sync();
}

CATN_DN void op_raw_split(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc s = E.site(op.a);   // (dleft, rest, 1)
  const int dleft = s.dl, rest = s.d, di = op.c;
  if (rest % di != 0) { E.status = STATUS_SHAPE; return; }
  const int m = dleft * di, n = rest / di;
  Split sp;
  if (!wg_split(E, E.ptr(s), m, n, m, true, 0, op.x, op.d, sp)) return;
  const int k = sp.k;
  const int64_t offu = E.alloc((int64_t)m * k);
  const int64_t offr = E.alloc((int64_t)k * n);
  if (E.status) return;
  wg_copy(g, E.out(offu), sp.L, (int64_t)m * k);
  wg_copy(g, E.out(offr), sp.R, (int64_t)k * n);
  E.work(0.0, (double)m * n + (double)m * k + (double)k * n);
  E.rec_begin(OP_RAW_SPLIT); E.rec_in(s);
  E.rec_out(offu, dleft, di, k); E.rec_out(offr, k, n, 1);
  E.rec_set(k, sp.qr, 1, op.d, op.x);
  E.rec_commit();
  E.set_site(op.b, offu, dleft, di, k);
  E.set_site(op.a, offr, k, n, 1);
  g.sync();
}

CATN_DN void op_scalar_lognorm(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc s = E.site(op.a);
  const int64_t n = (int64_t)s.dl * s.d * s.dr;
  const double* x = E.ptr(s);
  double part = 0.0;
  for (int64_t i = g.tid; i < n; i += g.nthr) part += x[i];
  const double z = wg_sum(g, part);
  E.acc(&E.st->lntail, log(fabs(z)));
  E.rec_begin(OP_SCALAR_LOGNORM); E.rec_in(s); E.rec_set(0, 0, 0, 0, z);
  E.rec_commit();
  if (g.tid == 0) E.st->sign = E.st->sign * (z > 0.0 ? 1.0 : (z < 0.0 ? -1.0 : 0.0));
  g.sync();
}

CATN_DN void op_tsvd(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc s = E.site(op.a);
  const int m = s.dl * s.d, n = s.dr;
  if (op.g == 1) {   // rsvd(A, k = op.d, oversample = op.h, power = op.i)
    double *U, *S, *V;
    const int kk = wg_rsvd(g, E.sc, E.ptr(s), m, n, m, op.d, op.h, op.i, (uint64_t)E.st->seed,
                           0, &U, &S, &V);
    if (kk < 0) { E.status = E.sc.fail ? STATUS_SCRATCH_FULL : STATUS_NOCONV; return; }
    const int64_t offu = E.alloc((int64_t)m * kk);
    const int64_t offv = E.alloc((int64_t)n * kk);
    const int64_t offs = E.alloc(kk);
    if (E.status) return;
    wg_copy(g, E.out(offu), U, (int64_t)m * kk);
    wg_copy(g, E.out(offv), V, (int64_t)n * kk);
    wg_copy(g, E.out(offs), S, kk);
    E.set_site(op.b, offu, m, 1, kk);
    E.set_site(op.c, offv, n, 1, kk);
    E.set_site(op.f, offs, kk, 1, 1);
    if (g.tid == 0) { E.st->results[op.e] = kk; E.st->dresults[op.e] = 0.0; }
    g.sync();
    return;
  }
  SvdResult R;
  int sw = 0;
  if (wg_svd(g, E.sc, E.ptr(s), m, n, m, R, &sw) < 0) {
    E.status = E.sc.fail ? STATUS_SCRATCH_FULL : STATUS_NOCONV; return;
  }
  E.bucket = size_bucket(R.r);
  const int k = keep_tsvd(R.S, R.r, op.x, op.d);
  const int64_t offu = E.alloc((int64_t)m * k);
  const int64_t offv = E.alloc((int64_t)n * k);
  const int64_t offs = E.alloc(R.r);
  if (E.status) return;
  wg_copy(g, E.out(offu), R.U, (int64_t)m * k);
  wg_copy(g, E.out(offv), R.V, (int64_t)n * k);
  wg_copy(g, E.out(offs), R.S, R.r);
  E.set_site(op.b, offu, m, 1, k);
  E.set_site(op.c, offv, n, 1, k);
  E.set_site(op.f, offs, R.r, 1, 1);
  if (g.tid == 0) {
    E.st->results[op.e] = k;
    E.st->dresults[op.e] = tail_sum(R.S, k, R.r);
  }
  g.sync();
}

}  // namespace catn

#include "zengine.h"

namespace catn {

// Execute ops [begin, end) of one instance. Returns the status.
CATN_DN int run_program(const WG& g, InstState* st, const Op* ops, int begin, int end) {
  Exec E;
  E.g = g;
  E.st = st;
  E.top = st->arena_top;
  E.status = st->status;
  E.flops = 0.0;
  E.bytes = 0.0;
  E.tape_len = st->tape_len;
  E.g.svd_fault = st->fault & 1;
  E.g.jac_tol = st->f32 ? 0x1p-26 : 0.0;
  if (g.tid == 0 && g.lds_cap > st->lds_max) st->lds_max = g.lds_cap;
  int64_t shwm = 0;
  for (int i = begin; i < end && E.status == STATUS_OK; ++i) {
    const Op op = ops[i];
    const int64_t t0 = now_ticks();
    E.bucket = -1;
    E.sc.base = st->scratch;
    E.sc.cap = st->scratch_cap;
    E.sc.top = 0;
    E.sc.fail = false;
    CATN_EMU_OP_BEGIN(g, st);
    E.snapshot(op);
    E.nconv = 0;
    E.nstage = 0;
    if (st->f32) {
      // FP64 copies of the op's input tensors (slots a, b; c for the eat! folds)
      const int nin = (op.code == OP_EAT_B || op.code == OP_EAT_C) ? 3
                      : (op.code == OP_PERMUTE321 || op.code == OP_NORMALIZE ||
                         op.code == OP_RAW_SPLIT || op.code == OP_SCALAR_LOGNORM ||
                         op.code == OP_TSVD) ? 1
                      : op.code == OP_RESHAPE ? 0 : 2;
      for (int q = 0; q < nin && E.status == STATUS_OK; ++q) {
        const SiteDesc d = E.op_desc[q];
        if (E.op_slot[q] >= 0) E.convert_in(d.off, (int64_t)d.dl * d.d * d.dr);
      }
      g.sync();
      if (E.status != STATUS_OK) {
        if (g.tid == 0) st->fail_op = i;
        break;
      }
    }
    if (st->cplx) zrun_op(E, op);
    else switch (op.code) {
      case OP_CANO_R: op_cano_r(E, op); break;
      case OP_CANO_L: op_cano_l(E, op); break;
      case OP_SWAP: op_swap(E, op); break;
      case OP_PERMUTE321: op_permute321(E, op); break;
      case OP_EAT_A: op_eat_a(E, op); break;
      case OP_EAT_B: op_eat_b(E, op); break;
      case OP_EAT_C: op_eat_c(E, op); break;
      case OP_NORMALIZE: op_normalize(E, op); break;
      case OP_MERGE_FUSE: op_merge_fuse(E, op); break;
      case OP_CUT: op_cut(E, op); break;
      case OP_COMPRESS_OPT: op_compress(E, op, true); break;
      case OP_COMPRESS: op_compress(E, op, false); break;
      case OP_RAW_SPLIT: op_raw_split(E, op); break;
      case OP_SCALAR_LOGNORM: op_scalar_lognorm(E, op); break;
      case OP_TSVD: op_tsvd(E, op); break;
      case OP_RESHAPE: op_reshape(E, op); break;
      default: break;
    }
    if (E.sc.fail && E.status == STATUS_OK) E.status = STATUS_SCRATCH_FULL;
    if (E.sc.top > shwm) shwm = E.sc.top;
    if (E.status != STATUS_OK && g.tid == 0) st->fail_op = i;
    g.sync();
    if (E.status == STATUS_OK) E.flush_out();
    if (g.tid == 0) {
      const int64_t dt = now_ticks() - t0;
      const int c = op.code & 15;
      count_add(st->op_ticks + c, dt);
      count_add(st->op_count + c, 1);
      if (E.bucket >= 0) {
        count_add(st->svd_ticks + E.bucket, dt);
        count_add(st->svd_count + E.bucket, 1);
      }
    }
  }
  if (g.tid == 0) {
    st->arena_top = E.top;
    st->status = E.status;
    st->tape_len = E.tape_len;
    st->alg_flops += E.flops;
    st->alg_bytes += E.bytes;
    if (shwm > st->scratch_hwm) st->scratch_hwm = shwm;
  }
  g.sync();
  return E.status;
}

}  // namespace catn

#include "adjoint.h"
