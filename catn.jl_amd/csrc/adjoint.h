// Reverse sweep of the contraction engine: the hand-written adjoint of
// contraction!'s lnZ with respect to every tensor version of a recorded run.
//
// The forward run (InstState::record = 1) appends one TapeRec per executed op
// and never overwrites a tensor version (the arena is append-only), so the
// reverse sweep reads every op's inputs where the forward left them. The
// adjoint arena `adj` is indexed like the arena. The sweep walks the tape
// backwards; each record adds the derivative of its own lnZ term (log-norms of
// normalisations, log|r| of a scalar eat, log|z| of the final scalar nodes) and
// pulls the adjoints of its outputs back onto its inputs. Splits (cano_to!
// steps, swap! and the Dmax cut) are replayed from their stored input with the
// forward's own factorisation path and kept count (a deterministic recompute:
// the same kernels on the same operand), and differentiated as
//   * certified full-rank steps (Householder QR, no truncation):
//       A = Q R,  Abar = (Qbar + Q copyltu(R Rbar^T - Qbar^T Q)) R^{-T};
//   * SVD steps, truncated to k of r = min(m, n) singular triplets: the full
//     thin-SVD adjoint with the adjoints of triplets k+1..r set to zero, which
//     is the exact derivative of the truncated map,
//       Abar = U G V^T + Ubar S^-1 V^T + U ((Vbar - V V^T Vbar) S^-1)^T,
//       G_ij = F_ij (J_ij - J_ji) s_j + d_ij Sbar_i + s_i F_ij (K_ij - K_ji)
//              - J_ij / s_j,   J = U^T Ubar, K = V^T Vbar,   F_ij = 1/(s_j^2 - s_i^2).
//     Pairs of kept triplets (i, j < k): lnZ is invariant under an orthogonal
//     gauge O on the new bond (L -> L O, R -> O^T R; the cut's factors both
//     -> F O), which makes L^T Lbar - Rbar R^T symmetric; substituting that
//     identity cancels (s_j^2 - s_i^2) exactly and leaves a finite G_ij with
//     no F (three forms below: U|SV', US|V', U sqrt(S)|V sqrt(S)). This is
//     exact for degenerate kept multiplets, which the +-J models produce all
//     the time. Pairs across the truncation (i < k <= j) keep F with
//     Lorentzian broadening, the only place it is needed:
//       F_ij = (s_j^2 - s_i^2) / ((s_j^2 - s_i^2)^2 + eps^2),
//       1/s -> s / (s^2 + eps'),  eps = 1e-12 s_max^2, eps' = (1e-12 s_max)^2.
// The Dmax cut splits the product P = Mi Mj^T; its adjoint is kept factored
// (Mibar = Pbar Mj, Mjbar = Pbar^T Mi without forming the mi x mj Pbar).
#pragma once

namespace catn {

constexpr double kBroaden = 1e-12;

CATN_D void wg_axpy(const WG& g, double* y, const double* x, double a, int64_t n) {
  for (int64_t i = g.tid; i < n; i += g.nthr) y[i] += a * x[i];
}

// largest |a - b| over n elements (every lane gets the value)
CATN_D double wg_maxdiff(const WG& g, const double* a, const double* b, int64_t n) {
  double part = 0.0;
  for (int64_t i = g.tid; i < n; i += g.nthr) part = fmax(part, fabs(a[i] - b[i]));
  return wg_max(g, part);
}

CATN_D double broad_sinv(double s, double e2) { return s / (s * s + e2); }

// Singular triplets that take part in the adjoint: all kept ones plus the
// discarded ones above 1e-13 s_max. Below that a singular value is treated as
// exactly zero: its vectors (numerical noise, or the unit-vector completion
// of an exactly zero column) drop out, and the complement projections
// (I - U U^T), (I - V V^T) over the remaining columns carry its exact-zero
// limit.
CATN_D int adjoint_rank(const double* S, int r, int k) {
  const double tau = 1e-13 * S[0];
  int re = k;
  while (re < r && S[re] > tau) ++re;
  return re;
}

// How the factors of a truncated SVD were handed on (the kept-pair form of G).
enum GaugeForm : int {
  GAUGE_U_SV = 1,      // L = U, R = S V'           (iso_left split)
  GAUGE_US_V = 2,      // L = U S, R = V'
  GAUGE_SQRT = 3,      // F_i = U sqrt(S), F_j = V sqrt(S)   (Dmax cut)
};

// Truncated-SVD adjoint core. U (m x r, ldu), S (r), V (n x r, ldv) is the
// full thin SVD; Ub (m x k, ld m), Sb (k), Vb (n x k, ld n) are the adjoints
// of the kept triplets. Produces A1 (m x r, ld m) and A2 (n x k, ld n) with
//   Abar = A1 V^T + U[:, :k] A2^T.
CATN_DN bool svd_adjoint(const WG& g, Scratch& sc, const double* U, int64_t ldu, const double* S,
                         const double* V, int64_t ldv, int m, int n, int r, int k,
                         const double* Ub, const double* Sb, const double* Vb, double* A1,
                         double* A2, int form) {
  double* J = sc.alloc((int64_t)r * k);
  double* K = sc.alloc((int64_t)r * k);
  double* G = sc.alloc((int64_t)r * r);
  if (sc.fail) return false;
  wg_gemm_any(g, r, k, m, 1.0, U, ldu, true, Ub, m, false, 0.0, J, r);
  wg_gemm_any(g, r, k, n, 1.0, V, ldv, true, Vb, n, false, 0.0, K, r);
  const double s0 = S[0];
  const double epsF = fmax(kBroaden * s0 * s0, 1e-150);
  const double eta = epsF * epsF;
  const double e2 = fmax((kBroaden * s0) * (kBroaden * s0), 1e-300);
  for (int64_t idx = g.tid; idx < (int64_t)r * r; idx += g.nthr) {
    const int i = (int)(idx % r), j = (int)(idx / r);
    const double si = S[i], sj = S[j];
    const double jij = j < k ? J[i + (int64_t)r * j] : 0.0;
    const double jji = i < k ? J[j + (int64_t)r * i] : 0.0;
    const double kij = j < k ? K[i + (int64_t)r * j] : 0.0;
    const double kji = i < k ? K[j + (int64_t)r * i] : 0.0;
    double v = 0.0;
    if (i != j && i < k && j < k) {
      const double isi = broad_sinv(si, e2), isj = broad_sinv(sj, e2);
      if (form == GAUGE_U_SV) v = kji * isi;
      else if (form == GAUGE_US_V) v = jij * isj;
      else v = (kij * si * isj - jji) / (si + sj) + jij * isj;
    } else if (i != j) {
      const double d = sj * sj - si * si;
      const double F = d / (d * d + eta);
      v = F * (jij - jji) * sj + si * F * (kij - kji);
    } else if (i < k) {
      v = Sb[i];
    }
    v -= jij * broad_sinv(sj, e2);
    G[idx] = v;
  }
  g.sync();
  // A1 = U G + Ub S^-1 (first k columns)
  wg_gemm_any(g, m, r, r, 1.0, U, ldu, false, G, r, false, 0.0, A1, m);
  for (int64_t idx = g.tid; idx < (int64_t)m * k; idx += g.nthr) {
    const int p = (int)(idx / m);
    A1[idx] += Ub[idx] * broad_sinv(S[p], e2);
  }
  // A2 = (Vb - V K) S^-1
  for (int64_t idx = g.tid; idx < (int64_t)n * k; idx += g.nthr) A2[idx] = Vb[idx];
  g.sync();
  wg_gemm_any(g, n, k, r, -1.0, V, ldv, false, K, r, false, 1.0, A2, n);
  for (int64_t idx = g.tid; idx < (int64_t)n * k; idx += g.nthr) {
    const int p = (int)(idx / n);
    A2[idx] *= broad_sinv(S[p], e2);
  }
  g.sync();
  return true;
}

// QR adjoint: A = Q R (Q: M x N orthonormal columns, ld ldq; R: N x N upper
// triangular, ld ldr); Qb (M x N, ld ldqb); only the upper triangle of Rb
// (N x N, ld ldrb) is used. Ab (M x N, ld M) is written.
CATN_DN bool qr_adjoint(const WG& g, Scratch& sc, const double* Q, int64_t ldq, const double* R,
                        int64_t ldr, const double* Qb, int64_t ldqb, const double* Rb,
                        int64_t ldrb, int M, int N, double* Ab) {
  double* Ru = sc.alloc((int64_t)N * N);
  double* Mx = sc.alloc((int64_t)N * N);
  double* X = sc.alloc((int64_t)N * N);
  double* B = sc.alloc((int64_t)M * N);
  if (sc.fail) return false;
  for (int64_t idx = g.tid; idx < (int64_t)N * N; idx += g.nthr) {
    const int i = (int)(idx % N), j = (int)(idx / N);
    Ru[idx] = i <= j ? Rb[i + ldrb * (int64_t)j] : 0.0;
    X[idx] = 0.0;
  }
  for (int64_t idx = g.tid; idx < (int64_t)M * N; idx += g.nthr) {
    const int i = (int)(idx % M), j = (int)(idx / M);
    B[idx] = Qb[i + ldqb * (int64_t)j];
  }
  g.sync();
  // Mx = R Ru^T - Qb^T Q
  wg_gemm_any(g, N, N, N, 1.0, R, ldr, false, Ru, N, true, 0.0, Mx, N);
  wg_gemm_any(g, N, N, M, -1.0, Qb, ldqb, true, Q, ldq, false, 1.0, Mx, N);
  // Ru <- copyltu(Mx): lower triangle mirrored onto the upper
  for (int64_t idx = g.tid; idx < (int64_t)N * N; idx += g.nthr) {
    const int i = (int)(idx % N), j = (int)(idx / N);
    Ru[idx] = i >= j ? Mx[idx] : Mx[j + (int64_t)N * i];
  }
  g.sync();
  wg_gemm_any(g, M, N, N, 1.0, Q, ldq, false, Ru, N, false, 1.0, B, M);
  if (!(wg_inv_fro_bound(g, R, N, ldr, X) > 0.0)) return false;
  wg_gemm_any(g, M, N, N, 1.0, B, M, false, X, N, true, 0.0, Ab, M);
  return true;
}

// Replay of a forward split (wg_split) with the recorded path and kept count:
// the factors needed by the adjoint, plus the forward outputs L (m x k) and
// R (k x n) as wg_split produced them.
struct SplitReplay {
  int m, n, k, r, path;
  bool iso_left;
  Factor F;       // path 1
  SvdResult S;    // path 0
  double* L;
  double* R;
};

CATN_DN bool split_replay(Exec& E, const double* A, int m, int n, int64_t lda, bool iso_left,
                          int rule, double cutoff, int maxdim, int path, int k, SplitReplay& t) {
  const WG& g = E.g;
  t.m = m; t.n = n; t.k = k; t.path = path; t.iso_left = iso_left;
  t.r = m < n ? m : n;
  if (path == 1) {
    const int maxd = rule == 1 ? (1 << 30) : maxdim;
    // the recorded step was certified on this same input: no second bound
    if (!wg_factor_fullrank(g, E.sc, A, m, n, lda, cutoff, maxd, iso_left, t.F, E.noise_floor(),
                            true) ||
        t.F.r != k) {
      E.status = E.sc.fail ? STATUS_SCRATCH_FULL : STATUS_REPLAY;
      return false;
    }
    t.L = t.F.L;
    t.R = t.F.Rt;
    return true;
  }
  if (wg_svd(g, E.sc, A, m, n, lda, t.S, nullptr) < 0 || k > t.S.r) {
    E.status = E.sc.fail ? STATUS_SCRATCH_FULL : STATUS_REPLAY;
    return false;
  }
  t.L = E.sc.alloc((int64_t)m * k);
  t.R = E.sc.alloc((int64_t)k * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return false; }
  const SvdResult& S = t.S;
  for (int64_t idx = g.tid; idx < (int64_t)m * k; idx += g.nthr) {
    const int p = (int)(idx / m);
    t.L[idx] = iso_left ? S.U[idx] : S.U[idx] * S.S[p];
  }
  for (int64_t idx = g.tid; idx < (int64_t)k * n; idx += g.nthr) {
    const int p = (int)(idx % k);
    const int64_t c = idx / k;
    const double v = S.V[c + (int64_t)n * p];
    t.R[idx] = iso_left ? S.S[p] * v : v;
  }
  g.sync();
  return true;
}

// Abar (m x n, ld m, written) from Lb (m x k, ld m) and Rb (k x n, ld k).
CATN_DN bool split_adjoint(Exec& E, const SplitReplay& t, const double* Lb, const double* Rb,
                           double* Ab) {
  const WG& g = E.g;
  const int m = t.m, n = t.n, k = t.k;
  Scratch& sc = E.sc;
  if (t.path == 1) {
    const bool trans = m < n;
    if (t.iso_left && trans) {            // L = I, R = A
      for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) Ab[idx] = Rb[idx];
      g.sync();
      return true;
    }
    if (!t.iso_left && !trans) {          // L = A, R = I
      for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) Ab[idx] = Lb[idx];
      g.sync();
      return true;
    }
    if (t.iso_left) {                     // A = Q R: L = Q (m x n), R (n x n)
      if (!qr_adjoint(g, sc, t.L, m, t.R, k, Lb, m, Rb, k, m, n, Ab)) {
        E.status = sc.fail ? STATUS_SCRATCH_FULL : STATUS_REPLAY;
        return false;
      }
      return true;
    }
    // A^T = Q' R' with Q' = R^T (n x m), R' = L^T (m x m)
    double* Qt = sc.alloc((int64_t)n * m);
    double* Rt = sc.alloc((int64_t)m * m);
    double* Qbt = sc.alloc((int64_t)n * m);
    double* Rbt = sc.alloc((int64_t)m * m);
    double* Abt = sc.alloc((int64_t)n * m);
    if (sc.fail) { E.status = STATUS_SCRATCH_FULL; return false; }
    wg_transpose(g, Qt, t.R, m, n);
    wg_transpose(g, Rt, t.L, m, m);
    wg_transpose(g, Qbt, Rb, m, n);
    wg_transpose(g, Rbt, Lb, m, m);
    g.sync();
    if (!qr_adjoint(g, sc, Qt, n, Rt, m, Qbt, n, Rbt, m, n, m, Abt)) {
      E.status = sc.fail ? STATUS_SCRATCH_FULL : STATUS_REPLAY;
      return false;
    }
    wg_transpose(g, Ab, Abt, n, m);
    g.sync();
    return true;
  }
  const SvdResult& S = t.S;
  const int r = adjoint_rank(S.S, S.r, k);
  double* Ub = sc.alloc((int64_t)m * k);
  double* Vb = sc.alloc((int64_t)n * k);
  double* Sb = sc.alloc(k);
  double* A1 = sc.alloc((int64_t)m * r);
  double* A2 = sc.alloc((int64_t)n * k);
  if (sc.fail) { E.status = STATUS_SCRATCH_FULL; return false; }
  if (t.iso_left) {   // L = U_k, R = S_k V_k^T
    for (int64_t idx = g.tid; idx < (int64_t)m * k; idx += g.nthr) Ub[idx] = Lb[idx];
    for (int64_t idx = g.tid; idx < (int64_t)n * k; idx += g.nthr) {
      const int c = (int)(idx % n), p = (int)(idx / n);
      Vb[idx] = S.S[p] * Rb[p + (int64_t)k * c];
    }
    for (int p = g.tid; p < k; p += g.nthr) {
      double s = 0.0;
      for (int c = 0; c < n; ++c) s += Rb[p + (int64_t)k * c] * S.V[c + (int64_t)n * p];
      Sb[p] = s;
    }
  } else {            // L = U_k S_k, R = V_k^T
    for (int64_t idx = g.tid; idx < (int64_t)m * k; idx += g.nthr) {
      const int p = (int)(idx / m);
      Ub[idx] = S.S[p] * Lb[idx];
    }
    for (int64_t idx = g.tid; idx < (int64_t)n * k; idx += g.nthr) {
      const int c = (int)(idx % n), p = (int)(idx / n);
      Vb[idx] = Rb[p + (int64_t)k * c];
    }
    for (int p = g.tid; p < k; p += g.nthr) {
      double s = 0.0;
      for (int i = 0; i < m; ++i) s += Lb[i + (int64_t)m * p] * S.U[i + (int64_t)m * p];
      Sb[p] = s;
    }
  }
  g.sync();
  if (!svd_adjoint(g, sc, S.U, m, S.S, S.V, n, m, n, r, k, Ub, Sb, Vb, A1, A2,
                   t.iso_left ? GAUGE_U_SV : GAUGE_US_V)) {
    E.status = STATUS_SCRATCH_FULL;
    return false;
  }
  wg_gemm_any(g, m, n, r, 1.0, A1, m, false, S.V, n, true, 0.0, Ab, m);
  wg_gemm_any(g, m, n, k, 1.0, S.U, m, false, A2, n, true, 1.0, Ab, m);
  return true;
}

// ---------------------------------------------------------------------------
// per-record reverse steps
// ---------------------------------------------------------------------------
struct BwdCtx {
  Exec* E;
  double* adj;
  // FP64 view of a recorded tensor version (converted before the record runs
  // when the network stores FP32)
  CATN_D const double* x(const TapeSite& s) const { return E->view(s.off); }
  CATN_D double* a(const TapeSite& s) const { return adj + s.off; }
};

CATN_D int64_t tsize(const TapeSite& s) { return (int64_t)s.dl * s.d * s.dr; }

CATN_D void note_replay(Exec& E, double dev) {
  if (E.g.tid == 0 && dev > E.st->replay_dev) E.st->replay_dev = dev;
}

// cano_to! right step (src/mps_node.jl:77-88)
CATN_DN void bwd_cano_r(const BwdCtx& c, const TapeRec& rc) {
  Exec& E = *c.E;
  const WG& g = E.g;
  const TapeSite &a = rc.in[0], &b = rc.in[1], &oa = rc.out[0], &ob = rc.out[1];
  const int m = a.dl * a.d, n = a.dr, k = rc.k;
  const int nb = b.d * b.dr;
  SplitReplay t;
  if (!split_replay(E, c.x(a), m, n, m, true, 1, rc.val, 0, rc.path, k, t)) return;
  note_replay(E, wg_maxdiff(g, t.L, c.x(oa), (int64_t)m * k));
  double* Rb = E.sc.alloc((int64_t)k * n);
  double* Ab = E.sc.alloc((int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const double* bbar = c.a(ob);
  // b' = R b:  Rbar = b'bar b^T,  bbar += R^T b'bar
  wg_gemm_any(g, k, n, nb, 1.0, bbar, k, false, c.x(b), n, true, 0.0, Rb, k);
  wg_gemm_any(g, n, nb, k, 1.0, t.R, k, true, bbar, k, false, 1.0, c.a(b), n);
  if (!split_adjoint(E, t, c.a(oa), Rb, Ab)) return;
  wg_axpy(g, c.a(a), Ab, 1.0, (int64_t)m * n);
  g.sync();
}

// cano_to! left step (src/mps_node.jl:89-101)
CATN_DN void bwd_cano_l(const BwdCtx& c, const TapeRec& rc) {
  Exec& E = *c.E;
  const WG& g = E.g;
  const TapeSite &a = rc.in[0], &b = rc.in[1], &oa = rc.out[0], &ob = rc.out[1];
  const int dl = a.dl, d = a.d, dr = a.dr;
  const int m = d * dr, n = dl, k = rc.k;
  const int rb = b.dl * b.d;
  double* Mt = E.sc.alloc((int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const double* ap = c.x(a);
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int64_t row = idx % m;
    const int xx = (int)(idx / m);
    const int y = (int)(row % d), z = (int)(row / d);
    Mt[idx] = ap[xx + (int64_t)dl * (y + (int64_t)d * z)];
  }
  g.sync();
  SplitReplay t;
  if (!split_replay(E, Mt, m, n, m, true, 1, rc.val, 0, rc.path, k, t)) return;
  double* Lb = E.sc.alloc((int64_t)m * k);
  double* Rb = E.sc.alloc((int64_t)k * n);
  double* Ab = E.sc.alloc((int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  // a'[p, row] = L[row, p]
  const double* abar = c.a(oa);
  const double* aout = c.x(oa);
  double dev = 0.0;
  for (int64_t idx = g.tid; idx < (int64_t)m * k; idx += g.nthr) {
    const int64_t row = idx % m;
    const int p = (int)(idx / m);
    Lb[idx] = abar[p + (int64_t)k * row];
    dev = fmax(dev, fabs(t.L[idx] - aout[p + (int64_t)k * row]));
  }
  note_replay(E, wg_max(g, dev));
  // b'[u, p] = sum_c b[u, c] R[p, c]: Rbar = b'bar^T b, bbar += b'bar R
  const double* bbar = c.a(ob);
  wg_gemm_any(g, k, n, rb, 1.0, bbar, rb, true, c.x(b), rb, false, 0.0, Rb, k);
  wg_gemm_any(g, rb, n, k, 1.0, bbar, rb, false, t.R, k, false, 1.0, c.a(b), rb);
  if (!split_adjoint(E, t, Lb, Rb, Ab)) return;
  double* ab = c.a(a);
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int64_t row = idx % m;
    const int xx = (int)(idx / m);
    const int y = (int)(row % d), z = (int)(row / d);
    ab[xx + (int64_t)dl * (y + (int64_t)d * z)] += Ab[idx];
  }
  g.sync();
}

// raw2mps step (src/mps_node.jl:31-50): the remaining (dleft, rest) block,
// as a (dleft di) x (rest / di) matrix, split U | S V^T by tsvd (chi, cutoff);
// U is the new site, S V^T the rest for the next step
CATN_DN void bwd_raw_split(const BwdCtx& c, const TapeRec& rc) {
  Exec& E = *c.E;
  const WG& g = E.g;
  const TapeSite &a = rc.in[0], &ou = rc.out[0], &orr = rc.out[1];
  const int m = ou.dl * ou.d, n = orr.d, k = rc.k;
  SplitReplay t;
  if (!split_replay(E, c.x(a), m, n, m, true, 0, rc.val, rc.aux0, rc.path, k, t)) return;
  note_replay(E, fmax(wg_maxdiff(g, t.L, c.x(ou), (int64_t)m * k),
                      wg_maxdiff(g, t.R, c.x(orr), (int64_t)k * n)));
  double* Ab = E.sc.alloc((int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  if (!split_adjoint(E, t, c.a(ou), c.a(orr), Ab)) return;
  wg_axpy(g, c.a(a), Ab, 1.0, (int64_t)m * n);
  g.sync();
}

// swap! (src/mps_node.jl:126-178): theta = tl * tr, truncated split
CATN_DN void bwd_swap(const BwdCtx& c, const TapeRec& rc) {
  Exec& E = *c.E;
  const WG& g = E.g;
  if (rc.path == 2) {   // randomized SVD: no reverse step
    E.status = STATUS_UNSUPPORTED;
    return;
  }
  const TapeSite &L = rc.in[0], &Rr = rc.in[1], &ol = rc.out[0], &orr = rc.out[1];
  const int d0 = L.dl, d2 = L.d, dd = L.dr, d1 = Rr.d, d3 = Rr.dr;
  const int m = d0 * d1, n = d2 * d3, k = rc.k;
  double* M = E.sc.alloc((int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const double* tl = c.x(L);
  const double* tr = c.x(Rr);
  if (E.st->f32) {   // the forward's split-BF16 theta, recomputed bit for bit
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
  } else {
    for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
      const int64_t row = idx % m, col = idx / m;
      const int i = (int)(row % d0), a = (int)(row / d0);
      const int j = (int)(col % d2), b = (int)(col / d2);
      const double* xx = tl + i + (int64_t)d0 * j;
      const double* y = tr + (int64_t)dd * (a + (int64_t)d1 * b);
      double s = 0.0;
      for (int q = 0; q < dd; ++q) s += xx[(int64_t)d0 * d2 * q] * y[q];
      M[idx] = s;
    }
    g.sync();
  }
  SplitReplay t;
  if (!split_replay(E, M, m, n, m, rc.flag != 0, 0, rc.val, rc.aux0, rc.path, k, t)) return;
  note_replay(E, fmax(wg_maxdiff(g, t.L, c.x(ol), (int64_t)m * k),
                      wg_maxdiff(g, t.R, c.x(orr), (int64_t)k * n)));
  double* Mb = E.sc.alloc((int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  if (!split_adjoint(E, t, c.a(ol), c.a(orr), Mb)) return;
  // Y[(i + d0 j), (a + d1 b)] = Mbar[(i + d0 a), (j + d2 b)]   (reuses M)
  double* Y = M;
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int64_t row = idx % m, col = idx / m;
    const int i = (int)(row % d0), a = (int)(row / d0);
    const int j = (int)(col % d2), b = (int)(col / d2);
    Y[(i + (int64_t)d0 * j) + (int64_t)d0 * d2 * (a + (int64_t)d1 * b)] = Mb[idx];
  }
  g.sync();
  const int P = d0 * d2, Qn = d1 * d3;
  wg_gemm_any(g, P, dd, Qn, 1.0, Y, P, false, tr, dd, true, 1.0, c.a(L), P);
  wg_gemm_any(g, dd, Qn, P, 1.0, tl, P, true, Y, P, false, 1.0, c.a(Rr), dd);
}

CATN_DN void bwd_permute321(const BwdCtx& c, const TapeRec& rc) {
  const WG& g = c.E->g;
  const TapeSite &a = rc.in[0], &o = rc.out[0];
  const double* yb = c.a(o);
  double* xb = c.a(a);
  const int64_t tot = tsize(a);
  for (int64_t idx = g.tid; idx < tot; idx += g.nthr) {
    const int xx = (int)(idx % a.dl);
    const int64_t rest = idx / a.dl;
    const int y = (int)(rest % a.d), z = (int)(rest / a.d);
    xb[idx] += yb[z + (int64_t)a.dr * (y + (int64_t)a.d * xx)];
  }
  g.sync();
}

// eat! case (a): lnZ += log|<a, b>|
CATN_DN void bwd_eat_a(const BwdCtx& c, const TapeRec& rc) {
  const WG& g = c.E->g;
  if (!rc.flag) return;
  const TapeSite &a = rc.in[0], &b = rc.in[1];
  const double inv = 1.0 / rc.val;
  const int64_t n = tsize(a);
  wg_axpy(g, c.a(a), c.x(b), inv, n);
  wg_axpy(g, c.a(b), c.x(a), inv, n);
  g.sync();
}

CATN_DN void bwd_normalize(const BwdCtx& c, const TapeRec& rc) {
  const WG& g = c.E->g;
  const TapeSite &xs = rc.in[0], &ys = rc.out[0];
  const int64_t n = tsize(xs);
  const double* yb = c.a(ys);
  double* xb = c.a(xs);
  const double nu = rc.val;
  const int method = rc.aux0;
  if (!rc.flag || method == 0) {      // y = x (not normalised, or norm 1)
    wg_axpy(g, xb, yb, 1.0, n);
    g.sync();
    return;
  }
  if (method == 1) {                  // y = x / ||x||, lnZ += log ||x||
    const double* y = c.x(ys);
    double part = 0.0;
    for (int64_t i = g.tid; i < n; i += g.nthr) part += y[i] * yb[i];
    const double dot = wg_sum(g, part);
    const double inv = 1.0 / nu;
    for (int64_t i = g.tid; i < n; i += g.nthr) xb[i] += (yb[i] - y[i] * dot + y[i]) * inv;
    g.sync();
    return;
  }
  // method 2: y = x / |x_m| with m the first index of the largest |x|
  const double* x = c.x(xs);
  double part = 0.0, best = -1.0;
  for (int64_t i = g.tid; i < n; i += g.nthr) {
    part += x[i] * yb[i];
    if (fabs(x[i]) == nu && best < 0.0) best = (double)i;
  }
  const double xy = wg_sum(g, part);
  const double neg = wg_max(g, best < 0.0 ? -1e300 : -best);
  const int64_t im = (int64_t)(-neg);
  const double inv = 1.0 / nu;
  for (int64_t i = g.tid; i < n; i += g.nthr) {
    double v = yb[i] * inv;
    if (i == im) v += (x[i] > 0.0 ? 1.0 : -1.0) * (1.0 - xy * inv) * inv;
    xb[i] += v;
  }
  g.sync();
}

// eat! case (b) fold: out[r] = sum_q p[r, q] (t j)[q]
CATN_DN void bwd_eat_b(const BwdCtx& c, const TapeRec& rc) {
  Exec& E = *c.E;
  const WG& g = E.g;
  const TapeSite &p = rc.in[0], &t = rc.in[1], &j = rc.in[2], &o = rc.out[0];
  const int D = t.dl, d = t.d;
  const int64_t rows = (int64_t)p.dl * p.d;
  double* mat = E.sc.alloc(D);
  double* matb = E.sc.alloc(D);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const double* tp = c.x(t);
  const double* jp = c.x(j);
  const double* pp = c.x(p);
  const double* yb = c.a(o);
  for (int i = g.tid; i < D; i += g.nthr) {
    double s = 0.0;
    for (int q = 0; q < d; ++q) s += tp[i + (int64_t)D * q] * jp[q];
    mat[i] = s;
  }
  g.sync();
  // pbar[r, q] += ybar[r] mat[q]
  double* pb = c.a(p);
  for (int64_t idx = g.tid; idx < rows * D; idx += g.nthr) {
    const int64_t r = idx % rows;
    const int q = (int)(idx / rows);
    pb[idx] += yb[r] * mat[q];
  }
  // matbar[q] = sum_r p[r, q] ybar[r]
  for (int q = 0; q < D; ++q) {
    double part = 0.0;
    for (int64_t r = g.tid; r < rows; r += g.nthr) part += pp[r + rows * q] * yb[r];
    const double s = wg_sum(g, part);
    if (g.tid == 0) matb[q] = s;
  }
  g.sync();
  double* tb = c.a(t);
  for (int64_t idx = g.tid; idx < (int64_t)D * d; idx += g.nthr) {
    const int i = (int)(idx % D), q = (int)(idx / D);
    tb[idx] += matb[i] * jp[q];
  }
  double* jb = c.a(j);
  for (int q = g.tid; q < d; q += g.nthr) {
    double s = 0.0;
    for (int i = 0; i < D; ++i) s += tp[i + (int64_t)D * q] * matb[i];
    jb[q] += s;
  }
  g.sync();
}

// eat! case (c) fold: out = p (t h)
CATN_DN void bwd_eat_c(const BwdCtx& c, const TapeRec& rc) {
  Exec& E = *c.E;
  const WG& g = E.g;
  const TapeSite &p = rc.in[0], &t = rc.in[1], &h = rc.in[2], &o = rc.out[0];
  const int Dl = t.dl, d = t.d, Dr = h.dr;
  const int rows = p.dl * p.d;
  double* mat = E.sc.alloc((int64_t)Dl * Dr);
  double* matb = E.sc.alloc((int64_t)Dl * Dr);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  wg_gemm_any(g, Dl, Dr, d, 1.0, c.x(t), Dl, false, c.x(h), d, false, 0.0, mat, Dl);
  const double* yb = c.a(o);
  wg_gemm_any(g, rows, Dl, Dr, 1.0, yb, rows, false, mat, Dl, true, 1.0, c.a(p), rows);
  wg_gemm_any(g, Dl, Dr, rows, 1.0, c.x(p), rows, true, yb, rows, false, 0.0, matb, Dl);
  wg_gemm_any(g, Dl, d, Dr, 1.0, matb, Dl, false, c.x(h), d, true, 1.0, c.a(t), Dl);
  wg_gemm_any(g, d, Dr, Dl, 1.0, c.x(t), Dl, true, matb, Dl, false, 1.0, c.a(h), d);
}

// merge! fuse: out (m x n) = a (m x k) b (k x n)
CATN_DN void bwd_merge_fuse(const BwdCtx& c, const TapeRec& rc) {
  const WG& g = c.E->g;
  const TapeSite &a = rc.in[0], &b = rc.in[1], &o = rc.out[0];
  const int m = a.dl * a.d, k = a.dr, n = b.d * b.dr;
  const double* yb = c.a(o);
  wg_gemm_any(g, m, k, n, 1.0, yb, m, false, c.x(b), k, true, 1.0, c.a(a), m);
  wg_gemm_any(g, k, n, m, 1.0, c.x(a), m, true, yb, m, false, 1.0, c.a(b), k);
}

// cut_bondim_opt! / cut_bondim! (src/tensor_network.jl:342-477): split of
// P = Mi Mj^T (through the QR factors when tall), kept factored.
CATN_DN void bwd_cut(const BwdCtx& c, const TapeRec& rc) {
  Exec& E = *c.E;
  const WG& g = E.g;
  const TapeSite &A = rc.in[0], &B = rc.in[1], &oA = rc.out[0], &oB = rc.out[1];
  const int dal = A.dl, d = A.d, dar = A.dr, dbl = B.dl, dbr = B.dr;
  const int mi = dal * dar, mj = dbl * dbr, k = rc.k;
  Scratch& sc = E.sc;
  double* Mi = sc.alloc((int64_t)mi * d);
  double* Mj = sc.alloc((int64_t)mj * d);
  double* Wi = sc.alloc((int64_t)mi * d);
  double* Wj = sc.alloc((int64_t)mj * d);
  if (sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const double* ap = c.x(A);
  const double* bp = c.x(B);
  for (int64_t idx = g.tid; idx < (int64_t)mi * d; idx += g.nthr) {
    const int64_t row = idx % mi;
    const int b = (int)(idx / mi);
    const int a = (int)(row % dal), cc = (int)(row / dal);
    Mi[idx] = Wi[idx] = ap[a + (int64_t)dal * (b + (int64_t)d * cc)];
  }
  for (int64_t idx = g.tid; idx < (int64_t)mj * d; idx += g.nthr) {
    const int64_t row = idx % mj;
    const int b = (int)(idx / mj);
    const int a = (int)(row % dbl), cc = (int)(row / dbl);
    Mj[idx] = Wj[idx] = bp[a + (int64_t)dbl * (b + (int64_t)d * cc)];
  }
  g.sync();
  const bool opt = rc.flag != 0;
  const bool fl = opt && mi > d, fr = opt && mj > d;
  double *Qi = nullptr, *Ri = Wi, *Qj = nullptr, *Rj = Wj;
  int ri_rows = mi, rj_rows = mj;
  if (fl) { if (!wg_thin_qr(g, sc, Wi, mi, d, &Qi, &Ri)) { E.status = STATUS_SCRATCH_FULL; return; } ri_rows = d; }
  if (fr) { if (!wg_thin_qr(g, sc, Wj, mj, d, &Qj, &Rj)) { E.status = STATUS_SCRATCH_FULL; return; } rj_rows = d; }
  double* Mg = sc.alloc((int64_t)ri_rows * rj_rows);
  if (sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const int ldri = fl ? d : mi, ldrj = fr ? d : mj;
  // the forward op_cut's product, same call
  wg_gemm_any(g, ri_rows, rj_rows, d, 1.0, Ri, ldri, false, Rj, ldrj, true, 0.0, Mg, ri_rows);
  SvdResult S;
  if (wg_svd(g, sc, Mg, ri_rows, rj_rows, ri_rows, S, nullptr) < 0 || k > S.r) {
    E.status = sc.fail ? STATUS_SCRATCH_FULL : STATUS_REPLAY;
    return;
  }
  const int r = adjoint_rank(S.S, S.r, k);
  // U_P = Qi U (mi x r), V_P = Qj V (mj x r)
  double* UP = S.U;
  double* VP = S.V;
  if (fl) {
    UP = sc.alloc((int64_t)mi * r);
    if (sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
    wg_gemm_any(g, mi, r, d, 1.0, Qi, mi, false, S.U, d, false, 0.0, UP, mi);
  }
  if (fr) {
    VP = sc.alloc((int64_t)mj * r);
    if (sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
    wg_gemm_any(g, mj, r, d, 1.0, Qj, mj, false, S.V, d, false, 0.0, VP, mj);
  }
  double* Ub = sc.alloc((int64_t)mi * k);
  double* Vb = sc.alloc((int64_t)mj * k);
  double* Sb = sc.alloc(k);
  double* A1 = sc.alloc((int64_t)mi * r);
  double* A2 = sc.alloc((int64_t)mj * k);
  double* Ci = sc.alloc((int64_t)r * d);
  double* Cj = sc.alloc((int64_t)r * d);
  double* X = sc.alloc((int64_t)k * d);
  double* Y = sc.alloc((int64_t)r * d);
  double* Mib = sc.alloc((int64_t)mi * d);
  double* Mjb = sc.alloc((int64_t)mj * d);
  if (sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  // Fi[a + dal c, p] = A'[a, p, c];  Ubar = Fibar sqrt(S), Vbar = Fjbar sqrt(S)
  const double* Fa = c.a(oA);
  const double* Fb = c.a(oB);
  const double* Fo = c.x(oA);
  double dev = 0.0;
  for (int64_t idx = g.tid; idx < (int64_t)mi * k; idx += g.nthr) {
    const int64_t row = idx % mi;
    const int p = (int)(idx / mi);
    const int a = (int)(row % dal), cc = (int)(row / dal);
    const int64_t src = a + (int64_t)dal * (p + (int64_t)k * cc);
    Ub[idx] = Fa[src] * sqrt(S.S[p]);
    dev = fmax(dev, fabs(UP[idx] * sqrt(S.S[p]) - Fo[src]));
  }
  for (int64_t idx = g.tid; idx < (int64_t)mj * k; idx += g.nthr) {
    const int64_t row = idx % mj;
    const int p = (int)(idx / mj);
    const int a = (int)(row % dbl), cc = (int)(row / dbl);
    Vb[idx] = Fb[a + (int64_t)dbl * (p + (int64_t)k * cc)] * sqrt(S.S[p]);
  }
  note_replay(E, wg_max(g, dev));
  // Sbar_p = (Fibar_p . UP_p + Fjbar_p . VP_p) / (2 sqrt(s_p))
  for (int p = g.tid; p < k; p += g.nthr) {
    double s = 0.0;
    for (int64_t i = 0; i < mi; ++i) s += Ub[i + (int64_t)mi * p] * UP[i + (int64_t)mi * p];
    for (int64_t i = 0; i < mj; ++i) s += Vb[i + (int64_t)mj * p] * VP[i + (int64_t)mj * p];
    Sb[p] = S.S[p] > 0.0 ? s / (2.0 * S.S[p]) : 0.0;
  }
  g.sync();
  if (!svd_adjoint(g, sc, UP, mi, S.S, VP, mj, mi, mj, r, k, Ub, Sb, Vb, A1, A2, GAUGE_SQRT)) {
    E.status = STATUS_SCRATCH_FULL;
    return;
  }
  // Ci = U_P^T Mi, Cj = V_P^T Mj (r x d)
  wg_gemm_any(g, r, d, mi, 1.0, UP, mi, true, Mi, mi, false, 0.0, Ci, r);
  wg_gemm_any(g, r, d, mj, 1.0, VP, mj, true, Mj, mj, false, 0.0, Cj, r);
  // Mibar = A1 Cj + U_P[:, :k] (A2^T Mj)
  wg_gemm_any(g, k, d, mj, 1.0, A2, mj, true, Mj, mj, false, 0.0, X, k);
  wg_gemm_any(g, mi, d, r, 1.0, A1, mi, false, Cj, r, false, 0.0, Mib, mi);
  wg_gemm_any(g, mi, d, k, 1.0, UP, mi, false, X, k, false, 1.0, Mib, mi);
  // Mjbar = V_P (A1^T Mi) + A2 Ci[:k]
  wg_gemm_any(g, r, d, mi, 1.0, A1, mi, true, Mi, mi, false, 0.0, Y, r);
  wg_gemm_any(g, mj, d, r, 1.0, VP, mj, false, Y, r, false, 0.0, Mjb, mj);
  wg_gemm_any(g, mj, d, k, 1.0, A2, mj, false, Ci, r, false, 1.0, Mjb, mj);
  double* Ab = c.a(A);
  double* Bb = c.a(B);
  for (int64_t idx = g.tid; idx < (int64_t)mi * d; idx += g.nthr) {
    const int64_t row = idx % mi;
    const int b = (int)(idx / mi);
    const int a = (int)(row % dal), cc = (int)(row / dal);
    Ab[a + (int64_t)dal * (b + (int64_t)d * cc)] += Mib[idx];
  }
  for (int64_t idx = g.tid; idx < (int64_t)mj * d; idx += g.nthr) {
    const int64_t row = idx % mj;
    const int b = (int)(idx / mj);
    const int a = (int)(row % dbl), cc = (int)(row / dbl);
    Bb[a + (int64_t)dbl * (b + (int64_t)d * cc)] += Mjb[idx];
  }
  g.sync();
}

// network_lognorm term log|sum(x)| (src/mps_node.jl:319-328)
CATN_DN void bwd_scalar_lognorm(const BwdCtx& c, const TapeRec& rc) {
  const WG& g = c.E->g;
  const TapeSite& s = rc.in[0];
  double* xb = c.a(s);
  const double inv = 1.0 / rc.val;
  for (int64_t i = g.tid; i < tsize(s); i += g.nthr) xb[i] += inv;
  g.sync();
}

// Reverse sweep over the whole tape of one instance. The adjoint arena is
// zeroed over [0, arena_top) first; afterwards adj holds d lnZ / d x for every
// tensor version x, in particular for the network's initial site tensors.
CATN_DN int run_backward(const WG& g, InstState* st) {
  Exec E;
  E.g = g;
  E.st = st;
  E.top = st->arena_top;
  E.status = STATUS_OK;
  E.flops = 0.0;
  E.bytes = 0.0;
  E.tape_len = st->tape_len;
  E.rp = nullptr;
  E.g.svd_fault = st->fault & 1;   // replays take the forward's SVD path
  E.g.jac_tol = st->f32 ? 0x1p-26 : 0.0;
  wg_fill(g, st->adj, 0.0, st->arena_top);
  if (g.tid == 0) st->replay_dev = 0.0;
  g.sync();
  BwdCtx c;
  c.E = &E;
  c.adj = st->adj;
  int64_t shwm = 0;
  for (int ti = st->tape_len - 1; ti >= 0 && E.status == STATUS_OK; --ti) {
    CATN_EMU_OP_BEGIN(g, st);
    const TapeRec rc = st->tape[ti];
    const int64_t t0 = now_ticks();
    // replay with the forward launch's LDS budget: the same kernels then take
    // the same code paths and reproduce the forward factors exactly
    if (rc.aux1 > g.lds_cap) { E.status = STATUS_REPLAY; if (g.tid == 0) st->fail_op = ti; break; }
    E.g.lds_cap = rc.aux1;
    E.bucket = -1;
    E.sc.base = st->scratch;
    E.sc.cap = st->scratch_cap;
    E.sc.top = 0;
    E.sc.fail = false;
    E.nconv = 0;
    E.nstage = 0;
    if (st->f32) {   // FP64 copies of the record's tensor versions
      for (int q = 0; q < rc.nin && E.status == STATUS_OK; ++q)
        E.convert_in(rc.in[q].off, tsize(rc.in[q]));
      for (int q = 0; q < rc.nout && E.status == STATUS_OK; ++q)
        E.convert_in(rc.out[q].off, tsize(rc.out[q]));
      g.sync();
      if (E.status != STATUS_OK) {
        if (g.tid == 0) st->fail_op = ti;
        break;
      }
    }
    switch (rc.code) {
      case OP_CANO_R: bwd_cano_r(c, rc); break;
      case OP_CANO_L: bwd_cano_l(c, rc); break;
      case OP_SWAP: bwd_swap(c, rc); break;
      case OP_PERMUTE321: bwd_permute321(c, rc); break;
      case OP_EAT_A: bwd_eat_a(c, rc); break;
      case OP_EAT_B: bwd_eat_b(c, rc); break;
      case OP_EAT_C: bwd_eat_c(c, rc); break;
      case OP_NORMALIZE: bwd_normalize(c, rc); break;
      case OP_MERGE_FUSE: bwd_merge_fuse(c, rc); break;
      case OP_CUT: bwd_cut(c, rc); break;
      case OP_SCALAR_LOGNORM: bwd_scalar_lognorm(c, rc); break;
      case OP_RAW_SPLIT: bwd_raw_split(c, rc); break;
      default: E.status = STATUS_UNSUPPORTED; break;
    }
    if (E.sc.fail && E.status == STATUS_OK) E.status = STATUS_SCRATCH_FULL;
    if (E.sc.top > shwm) shwm = E.sc.top;
    if (E.status != STATUS_OK && g.tid == 0) st->fail_op = ti;
    g.sync();
    if (g.tid == 0) st->bwd_ticks[rc.code & 15] += now_ticks() - t0;
  }
  if (g.tid == 0) {
    st->status = E.status;
    if (shwm > st->scratch_hwm) st->scratch_hwm = shwm;
  }
  g.sync();
  return E.status;
}

}  // namespace catn
