// Complex element types (ComplexF64, ComplexF32 storage) of the device
// interpreter.
//
// A complex network (InstState::cplx = 1) stores every site tensor as
// interleaved (re, im) values: element offset e of a site lives at values
// 2e, 2e + 1 of the arena (doubles, or floats with ComplexF32 storage, where
// each op reads FP64 copies of its inputs and its outputs are rounded into
// the arena after it, as for the real FP32 storage), so descriptors, slot
// tables, programs and the planner are the ones of the real engine. Each opcode has a complex
// counterpart here with the reference's semantics for complex T:
//   * every factorisation is a truncated SVD A = U S V^H (one-sided Jacobi on
//     the columns, complex rotations), handed on as L R = A with the real
//     engine's splits: U | S V^H, U S | V^H, and for the Dmax cut
//     U sqrt(S) | sqrt(S) V^H, whose transpose is the reference's
//     conj(V) sqrt(S) (src/tensor_network.jl:381, :466);
//   * contractions of bonds are non-conjugating products (ein"..." and `*` in
//     src/mps_node.jl), including eat!'s leaf dot sum(vi .* vj)
//     (src/mps_node.jl:475);
//   * phases are complex: eat! case (a) multiplies psi by r / |r|, and a
//     final scalar node multiplies the sign by z / |z| (Julia's sign(z)).
// cut_bondim_opt!, compress_opt! and swap!'s QR reductions are not taken: the
// SVD of the full product has the same singular values and spans the same
// kept subspaces, so kept counts, log-norms and the represented tensors are
// the reference's; only the unitary gauge of a new bond can differ, and no
// later step depends on it. The reverse sweep is real-only (complex networks
// run the plain forward).
#pragma once

namespace catn {

struct zd {
  double re, im;
};

CATN_D zd zmk(double r, double i) {
  zd z;
  z.re = r;
  z.im = i;
  return z;
}
CATN_D zd zmul(zd a, zd b) { return zmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
CATN_D zd zconj(zd a) { return zmk(a.re, -a.im); }
CATN_D zd zscale(zd a, double s) { return zmk(a.re * s, a.im * s); }
CATN_D double zabs2(zd a) { return a.re * a.re + a.im * a.im; }
CATN_D double zabs(zd a) { return hypot(a.re, a.im); }

CATN_D zd* zalloc(Scratch& sc, int64_t n) { return (zd*)sc.alloc(2 * n); }
// an op's input tensor, and where it writes the output allocated at `off`:
// the arena itself, or with ComplexF32 storage (f32) the FP64 copies and
// staging buffers of the op (Exec::convert_in, Exec::alloc)
CATN_D zd* zsite(Exec& E, const SiteDesc& s) {
  return E.st->f32 ? (zd*)E.view(s.off) : (zd*)E.st->arena + s.off;
}
CATN_D zd* zsite_at(Exec& E, int64_t off) {
  return E.st->f32 ? (zd*)E.out(off) : (zd*)E.st->arena + off;
}

// C (m x n, ldc) = op(A) op(B) (+ C when acc), op = transpose when ta / tb,
// never conjugated; one lane per entry of C, summed over q in order.
CATN_D void zgemm(const WG& g, int m, int n, int k, const zd* A, int64_t lda, bool ta, const zd* B,
                  int64_t ldb, bool tb, bool acc, zd* C, int64_t ldc) {
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int i = (int)(idx % m), j = (int)(idx / m);
    double sr = 0.0, si = 0.0;
    for (int q = 0; q < k; ++q) {
      const zd a = ta ? A[q + lda * i] : A[i + lda * q];
      const zd b = tb ? B[j + ldb * q] : B[q + ldb * j];
      sr += a.re * b.re - a.im * b.im;
      si += a.re * b.im + a.im * b.re;
    }
    zd* c = C + i + ldc * j;
    if (acc) {
      c->re += sr;
      c->im += si;
    } else {
      c->re = sr;
      c->im = si;
    }
  }
  g.sync();
}

struct ZSvd {
  int r;
  zd* U;      // m x r (ld m)
  double* S;  // r, descending
  zd* V;      // n x r (ld n); A = U diag(S) V^H
};

// Thin SVD of A (m x n, lda) by one-sided (Hestenes) Jacobi on the columns of
// W = A, or of W = A^H when m < n (then A = V' S U'^H). Column pairs of one
// round-robin round are disjoint and go to aligned lane groups; a pair (p, q)
// with Gram entries alpha = |w_p|^2, beta = |w_q|^2, gamma = w_p^H w_q is
// rotated by the real Jacobi rotation of (w_p, e^{-i arg gamma} w_q), whose
// Gram matrix [[alpha, |gamma|], [|gamma|, beta]] is real:
//   w_p' = c w_p - s e^{-i phi} w_q,  w_q' = s w_p + c e^{-i phi} w_q,
// the same unitary 2 x 2 transform applied to V's columns, so W V^H = A
// throughout. Skip tests as in the real Jacobi (|gamma| <= 4 eps sqrt(m)
// sqrt(alpha beta), or the FP32-storage floor g.jac_tol; columns below
// 1e-36 ||A||_F^2). Returns r = min(m, n), or -1 (scratch exhausted, or 60
// sweeps without convergence).
CATN_DN int zsvd(const WG& g, Scratch& sc, const zd* A, int m, int n, int64_t lda, ZSvd& out) {
  const bool tr = m < n;
  const int tm = tr ? n : m, tn = tr ? m : n;
  zd* W = zalloc(sc, (int64_t)tm * tn);
  zd* V = zalloc(sc, (int64_t)tn * tn);
  zd* U = zalloc(sc, (int64_t)tm * tn);
  zd* Vs = zalloc(sc, (int64_t)tn * tn);
  double* Ss = sc.alloc(tn);
  double* key = sc.alloc(tn);
  int* rk = (int*)sc.alloc(tn);
  if (sc.fail) return -1;
  double part = 0.0;
  for (int64_t idx = g.tid; idx < (int64_t)tm * tn; idx += g.nthr) {
    const int i = (int)(idx % tm), j = (int)(idx / tm);
    const zd v = tr ? zconj(A[j + lda * i]) : A[i + lda * j];
    W[idx] = v;
    part += zabs2(v);
  }
  for (int64_t idx = g.tid; idx < (int64_t)tn * tn; idx += g.nthr)
    V[idx] = zmk((idx % tn) == (idx / tn) ? 1.0 : 0.0, 0.0);
  const double fro2 = wg_sum(g, part);   // its barrier also publishes W and V
  const double thr = 1e-36 * fro2;
  const double tol = fmax(4.0 * 2.220446049250313e-16 * sqrt((double)(tm > 1 ? tm : 1)), g.jac_tol);
  const int N = tn + (tn & 1), npairs = N / 2;
  // lanes per column pair: the workgroup split over the pairs, and at least
  // the global Jacobi's lanes per pair (g.jac_lanes) so that a long pair is
  // read by consecutive lanes (coalesced) and the round's pairs go in passes
  int L = pow2floor(g.nthr / (npairs > 0 ? npairs : 1));
  if (L < g.jac_lanes) L = pow2floor(g.jac_lanes < g.nthr ? g.jac_lanes : g.nthr);
  if (L > 64) L = 64;
  const int ngr = g.nthr / L, grp = g.tid / L, lane = g.tid % L;
  int sw = tn < 2 ? 0 : -1;
  for (int sweep = 1; sweep <= 60 && sw < 0; ++sweep) {
    double rot = 0.0;
    for (int r = 0; r < N - 1; ++r) {
      for (int kp = grp; kp < npairs; kp += ngr) {   // uniform within a lane group
        int p, q;
        if (kp == 0) { p = r; q = N - 1; }
        else { p = (r + kp) % (N - 1); q = (r - kp + N - 1) % (N - 1); }
        if (p > q) { const int t = p; p = q; q = t; }
        if (q >= tn) continue;
        zd* wp = W + (int64_t)tm * p;
        zd* wq = W + (int64_t)tm * q;
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int i = lane; i < tm; i += L) {
          const zd x = wp[i], y = wq[i];
          al += zabs2(x);
          be += zabs2(y);
          gr += x.re * y.re + x.im * y.im;   // conj(x) y
          gi += x.re * y.im - x.im * y.re;
        }
        al = grp_sum(al, L);
        be = grp_sum(be, L);
        gr = grp_sum(gr, L);
        gi = grp_sum(gi, L);
        if (al <= thr || be <= thr) continue;
        const double ga = hypot(gr, gi);
        if (ga <= tol * sqrt(al) * sqrt(be)) continue;
        rot = 1.0;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        const zd ph = zmk(gr / ga, -gi / ga);   // e^{-i arg gamma}
        for (int i = lane; i < tm; i += L) {
          const zd x = wp[i], y = zmul(ph, wq[i]);
          wp[i] = zmk(c * x.re - s * y.re, c * x.im - s * y.im);
          wq[i] = zmk(s * x.re + c * y.re, s * x.im + c * y.im);
        }
        zd* vp = V + (int64_t)tn * p;
        zd* vq = V + (int64_t)tn * q;
        for (int i = lane; i < tn; i += L) {
          const zd x = vp[i], y = zmul(ph, vq[i]);
          vp[i] = zmk(c * x.re - s * y.re, c * x.im - s * y.im);
          vq[i] = zmk(s * x.re + c * y.re, s * x.im + c * y.im);
        }
      }
      g.sync();   // the next round pairs the columns differently
    }
    if (wg_max(g, rot) == 0.0) sw = sweep;
  }
  // singular values = column norms; descending order, stable in the column
  // index; U = W / s (a zero column keeps the unit vector e_rank)
  for (int j = g.tid; j < tn; j += g.nthr) {
    double ss = 0.0;
    for (int i = 0; i < tm; ++i) ss += zabs2(W[i + (int64_t)tm * j]);
    key[j] = sqrt(ss);
  }
  g.sync();
  for (int j = g.tid; j < tn; j += g.nthr) {
    const double sj = key[j];
    int rr = 0;
    for (int i = 0; i < tn; ++i) rr += (key[i] > sj) || (key[i] == sj && i < j);
    rk[j] = rr;
    Ss[rr] = sj;
  }
  g.sync();
  for (int64_t idx = g.tid; idx < (int64_t)tm * tn; idx += g.nthr) {
    const int i = (int)(idx % tm), j = (int)(idx / tm);
    const int rr = rk[j];
    const double sj = key[j];
    U[i + (int64_t)tm * rr] = sj > 0.0 ? zscale(W[idx], 1.0 / sj) : zmk(i == rr ? 1.0 : 0.0, 0.0);
  }
  for (int64_t idx = g.tid; idx < (int64_t)tn * tn; idx += g.nthr) {
    const int i = (int)(idx % tn), j = (int)(idx / tn);
    Vs[i + (int64_t)tn * rk[j]] = V[idx];
  }
  g.sync();
  out.r = tn;
  out.S = Ss;
  out.U = tr ? Vs : U;   // m x r (ld m)
  out.V = tr ? U : Vs;   // n x r (ld n)
  return sw < 0 ? -1 : tn;
}

// How the truncated factors are handed on: L R ~= A with
enum ZSplitForm : int {
  ZSPLIT_U_SV = 0,     // L = U,         R = S V^H          (isometry left)
  ZSPLIT_US_V = 1,     // L = U S,       R = V^H            (isometry rows right)
  ZSPLIT_SQRT = 2,     // L = U sqrt(S), R = sqrt(S) V^H    (Dmax cut)
};

struct ZSplit {
  int k;
  zd* L;      // m x k (ld m)
  zd* R;      // k x n (ld k)
  double disc;
};

// Truncated SVD split of A (m x n, lda): rule 0 keeps min(count(S > cutoff),
// maxdim), at least 1 (tsvd, src/linalg_utils.jl:24-25); rule 1 is the
// cano_to! rule (src/mps_node.jl:79-82). disc = sum of the dropped S.
CATN_DN bool zsplit(Exec& E, const zd* A, int m, int n, int64_t lda, int form, int rule,
                    double cutoff, int maxdim, ZSplit& sp) {
  const WG& g = E.g;
  const int r0 = m < n ? m : n;
  E.bucket = size_bucket(r0);
  ZSvd R;
  if (zsvd(g, E.sc, A, m, n, lda, R) < 0) {
    E.status = E.sc.fail ? STATUS_SCRATCH_FULL : STATUS_NOCONV;
    return false;
  }
  E.work(4.0 * svd_flops(m, n), 0.0);
  // ComplexF32 storage: singular values below the storage's noise floor
  // count as zero, as in the real engine's FP32 storage
  const double noise = E.noise_floor();
  if (noise > 0.0) {
    double f2 = 0.0;
    for (int i = 0; i < R.r; ++i) f2 += R.S[i] * R.S[i];
    cutoff = fmax(cutoff, noise * sqrt(f2));
  }
  const int k = rule == 1 ? keep_cano(R.S, R.r, cutoff) : keep_tsvd(R.S, R.r, cutoff, maxdim);
  zd* L = zalloc(E.sc, (int64_t)m * k);
  zd* Rm = zalloc(E.sc, (int64_t)k * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return false; }
  for (int64_t idx = g.tid; idx < (int64_t)m * k; idx += g.nthr) {
    const int p = (int)(idx / m);
    const double f = form == ZSPLIT_U_SV ? 1.0 : (form == ZSPLIT_US_V ? R.S[p] : sqrt(R.S[p]));
    L[idx] = zscale(R.U[idx], f);
  }
  for (int64_t idx = g.tid; idx < (int64_t)k * n; idx += g.nthr) {
    const int p = (int)(idx % k);
    const int64_t c = idx / k;
    const double f = form == ZSPLIT_U_SV ? R.S[p] : (form == ZSPLIT_US_V ? 1.0 : sqrt(R.S[p]));
    Rm[idx] = zscale(zconj(R.V[c + (int64_t)n * p]), f);
  }
  g.sync();
  sp.k = k;
  sp.L = L;
  sp.R = Rm;
  sp.disc = tail_sum(R.S, k, R.r);
  return true;
}

CATN_D void zcopy(const WG& g, zd* dst, const zd* src, int64_t n) {
  for (int64_t i = g.tid; i < n; i += g.nthr) dst[i] = src[i];
}

// cano_to! right step (src/mps_node.jl:77-87)
CATN_DN void zop_cano_r(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a), b = E.site(op.b);
  const int m = a.dl * a.d, n = a.dr;
  if (b.dl != n) { E.status = STATUS_SHAPE; return; }
  ZSplit sp;
  if (!zsplit(E, zsite(E, a), m, n, m, ZSPLIT_U_SV, 1, op.x, 0, sp)) return;
  const int k = sp.k;
  const int64_t nb = (int64_t)b.d * b.dr;
  const int64_t offa = E.alloc((int64_t)m * k);
  const int64_t offb = E.alloc((int64_t)k * nb);
  if (E.status) return;
  zcopy(g, zsite_at(E, offa), sp.L, (int64_t)m * k);
  zgemm(g, k, (int)nb, n, sp.R, k, false, zsite(E, b), n, false, false, zsite_at(E, offb), k);
  E.work(8.0 * k * n * (double)nb, 2.0 * ((double)m * n + (double)n * nb + (double)m * k + (double)k * nb));
  E.set_site(op.a, offa, a.dl, a.d, k);
  E.set_site(op.b, offb, k, b.d, b.dr);
  g.sync();
}

// cano_to! left step (src/mps_node.jl:90-100): SVD of permutedims(a, (2,3,1))
// as (d dr) x dl; a <- permutedims(U, (3,1,2)), b <- b * (S V^H)^T
CATN_DN void zop_cano_l(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a), b = E.site(op.b);
  const int dl = a.dl, d = a.d, dr = a.dr;
  if (b.dr != dl) { E.status = STATUS_SHAPE; return; }
  const int m = d * dr, n = dl;
  zd* Mt = zalloc(E.sc, (int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const zd* ap = zsite(E, a);
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int64_t row = idx % m;
    const int x = (int)(idx / m);
    const int y = (int)(row % d), z = (int)(row / d);
    Mt[idx] = ap[x + (int64_t)dl * (y + (int64_t)d * z)];
  }
  g.sync();
  ZSplit sp;
  if (!zsplit(E, Mt, m, n, m, ZSPLIT_U_SV, 1, op.x, 0, sp)) return;
  const int k = sp.k;
  const int64_t rb = (int64_t)b.dl * b.d;
  const int64_t offa = E.alloc((int64_t)k * m);
  const int64_t offb = E.alloc(rb * k);
  if (E.status) return;
  zd* na = zsite_at(E, offa);
  for (int64_t idx = g.tid; idx < (int64_t)k * m; idx += g.nthr) {
    const int p = (int)(idx % k);
    const int64_t row = idx / k;
    na[idx] = sp.L[row + (int64_t)m * p];
  }
  zgemm(g, (int)rb, k, n, zsite(E, b), rb, false, sp.R, k, true, false, zsite_at(E, offb), rb);
  E.work(8.0 * (double)rb * n * k, 2.0 * ((double)m * n + (double)rb * n + (double)k * m + (double)rb * k));
  E.set_site(op.a, offa, k, d, dr);
  E.set_site(op.b, offb, b.dl, b.d, k);
  g.sync();
}

// swap! numerics (src/mps_node.jl:144-170): theta = ein"ijk,kab->iajb"(tl,
// tr) as (d0 d1) x (d2 d3), tsvd(cutoff, chi); going right U | S V^H, going
// left U S | V^H.
CATN_DN void zop_swap(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc L = E.site(op.a), Rr = E.site(op.b);
  const int d0 = L.dl, d2 = L.d, dd = L.dr, d1 = Rr.d, d3 = Rr.dr;
  if (Rr.dl != dd) { E.status = STATUS_SHAPE; return; }
  const int m = d0 * d1, n = d2 * d3;
  // the randomized-SVD branch (src/mps_node.jl:154-155) is real-only here
  if (op.f && ((m > op.g && n > op.g) || m > op.h || n > op.h)) { E.status = STATUS_UNSUPPORTED; return; }
  zd* M = zalloc(E.sc, (int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const zd* tl = zsite(E, L);
  const zd* tr = zsite(E, Rr);
  for (int64_t idx = g.tid; idx < (int64_t)m * n; idx += g.nthr) {
    const int64_t row = idx % m, col = idx / m;
    const int i = (int)(row % d0), a = (int)(row / d0);
    const int j = (int)(col % d2), b = (int)(col / d2);
    const zd* x = tl + i + (int64_t)d0 * j;
    const zd* y = tr + (int64_t)dd * (a + (int64_t)d1 * b);
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < dd; ++k) {
      const zd u = x[(int64_t)d0 * d2 * k], v = y[k];
      sr += u.re * v.re - u.im * v.im;
      si += u.re * v.im + u.im * v.re;
    }
    M[idx] = zmk(sr, si);
  }
  g.sync();
  ZSplit sp;
  if (!zsplit(E, M, m, n, m, op.c != 0 ? ZSPLIT_U_SV : ZSPLIT_US_V, 0, op.x, op.d, sp)) return;
  const int k = sp.k;
  if (op.e) E.acc(&E.st->err, sp.disc);
  const int64_t offl = E.alloc((int64_t)m * k);
  const int64_t offr = E.alloc((int64_t)k * n);
  if (E.status) return;
  zcopy(g, zsite_at(E, offl), sp.L, (int64_t)m * k);
  zcopy(g, zsite_at(E, offr), sp.R, (int64_t)k * n);
  E.work(8.0 * m * n * (double)dd,
         2.0 * ((double)d0 * d2 * dd + (double)dd * d1 * d3 + (double)m * k + (double)k * n));
  E.set_site(op.a, offl, d0, d1, k);
  E.set_site(op.b, offr, k, d2, d3);
  g.sync();
}

// reverse! per site: permutedims(t, (3,2,1)) (src/mps_node.jl:240)
CATN_DN void zop_permute321(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a);
  const int64_t tot = (int64_t)a.dl * a.d * a.dr;
  const int64_t off = E.alloc(tot);
  if (E.status) return;
  const zd* s = zsite(E, a);
  zd* dst = zsite_at(E, off);
  for (int64_t idx = g.tid; idx < tot; idx += g.nthr) {
    const int z = (int)(idx % a.dr);
    const int64_t rest = idx / a.dr;
    const int y = (int)(rest % a.d), x = (int)(rest / a.d);
    dst[idx] = s[x + (int64_t)a.dl * (y + (int64_t)a.d * z)];
  }
  E.work(0.0, 4.0 * tot);
  E.set_site(op.a, off, a.dr, a.d, a.dl);
  g.sync();
}

// eat! case (a): r = sum(vi .* vj), non-conjugating (src/mps_node.jl:469-489);
// lnZ += log|r|, psi *= r / |r| unless |r| <= cutoff
CATN_DN void zop_eat_a(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a), b = E.site(op.b);
  const int64_t n = (int64_t)a.dl * a.d * a.dr;
  if (n != (int64_t)b.dl * b.d * b.dr) { E.status = STATUS_SHAPE; return; }
  const zd* x = zsite(E, a);
  const zd* y = zsite(E, b);
  double pr = 0.0, pi = 0.0;
  for (int64_t i = g.tid; i < n; i += g.nthr) {
    pr += x[i].re * y[i].re - x[i].im * y[i].im;
    pi += x[i].re * y[i].im + x[i].im * y[i].re;
  }
  const double rr = wg_sum(g, pr);
  const double ri = wg_sum(g, pi);
  const double ar = hypot(rr, ri);
  E.work(8.0 * n, 4.0 * n);
  if (ar > op.x) {
    E.acc(&E.st->lnZ, log(ar));
    if (g.tid == 0) {
      const zd ph = zmul(zmk(E.st->psi, E.st->psi_i), zmk(rr / ar, ri / ar));
      E.st->psi = ph.re;
      E.st->psi_i = ph.im;
    }
  }
  if (g.tid == 0) E.st->sites[op.a].live = 0;
  g.sync();
}

// norm of a site per norm_method: 1 2-norm, 2 largest |x|, 0 one
CATN_D double zsite_norm(const WG& g, const zd* x, int64_t n, int method) {
  if (method == 1) {
    double part = 0.0;
    for (int64_t i = g.tid; i < n; i += g.nthr) part += zabs2(x[i]);
    return sqrt(wg_sum(g, part));
  } else if (method == 2) {
    double part = 0.0;
    for (int64_t i = g.tid; i < n; i += g.nthr) part = fmax(part, zabs(x[i]));
    return wg_max(g, part);
  }
  return 1.0;
}

// normalisation after eat! (src/mps_node.jl:506-527, :558-574)
CATN_DN void zop_normalize(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc s = E.site(op.a);
  const zd* x = zsite(E, s);
  const int64_t n = (int64_t)s.dl * s.d * s.dr;
  const double nrm = zsite_norm(g, x, n, op.d);
  E.work(6.0 * n, 4.0 * n);
  const int64_t off = E.alloc(n);
  if (E.status) return;
  zd* y = zsite_at(E, off);
  const bool counted = nrm > op.x;
  const double inv = counted ? 1.0 / nrm : 1.0;
  for (int64_t i = g.tid; i < n; i += g.nthr) y[i] = zscale(x[i], inv);
  if (counted) E.acc(&E.st->lnZ, log(nrm));
  E.set_site(op.a, off, s.dl, s.d, s.dr);
  g.sync();
}

// eat! case (b) fold (src/mps_node.jl:492-520): mat = t * j (D x 1);
// new p = ein"ijk,ka->ija"(p, mat)
CATN_DN void zop_eat_b(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc p = E.site(op.a), t = E.site(op.b), j = E.site(op.c);
  const int D = t.dl, d = t.d;
  if (t.dr != 1 || j.d != d || p.dr != D) { E.status = STATUS_SHAPE; return; }
  zd* mat = zalloc(E.sc, D);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  zgemm(g, D, 1, d, zsite(E, t), D, false, zsite(E, j), d, false, false, mat, D);
  const int64_t rows = (int64_t)p.dl * p.d;
  const int64_t off = E.alloc(rows);
  if (E.status) return;
  zgemm(g, (int)rows, 1, D, zsite(E, p), rows, false, mat, D, false, false, zsite_at(E, off), rows);
  E.work(8.0 * D * d + 8.0 * rows * D, 2.0 * ((double)D * d + d + rows * D + rows));
  E.set_site(op.a, off, p.dl, p.d, 1);
  if (g.tid == 0) { E.st->sites[op.b].live = 0; E.st->sites[op.c].live = 0; }
  g.sync();
}

// eat! case (c) fold (src/mps_node.jl:533-539): mat = t * h (Dl x Dr);
// new p = ein"ijk,ka->ija"(p, mat)
CATN_DN void zop_eat_c(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc p = E.site(op.a), t = E.site(op.b), h = E.site(op.c);
  const int Dl = t.dl, d = t.d, Dr = h.dr;
  if (t.dr != 1 || h.dl != 1 || h.d != d || p.dr != Dl) { E.status = STATUS_SHAPE; return; }
  zd* mat = zalloc(E.sc, (int64_t)Dl * Dr);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  zgemm(g, Dl, Dr, d, zsite(E, t), Dl, false, zsite(E, h), d, false, false, mat, Dl);
  const int64_t rows = (int64_t)p.dl * p.d;
  const int64_t off = E.alloc(rows * Dr);
  if (E.status) return;
  zgemm(g, (int)rows, Dr, Dl, zsite(E, p), rows, false, mat, Dl, false, false, zsite_at(E, off), rows);
  E.work(8.0 * Dl * Dr * d + 8.0 * rows * Dl * Dr,
         2.0 * ((double)Dl * d + (double)d * Dr + rows * Dl + rows * Dr));
  E.set_site(op.a, off, p.dl, p.d, Dr);
  if (g.tid == 0) { E.st->sites[op.b].live = 0; E.st->sites[op.c].live = 0; }
  g.sync();
}

// merge! fuse: ein"ijk,kab->ijab" reshaped (dl, d1 d2, dr2) (src/mps_node.jl:610-616)
CATN_DN void zop_merge_fuse(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc a = E.site(op.a), b = E.site(op.b);
  if (a.dr != b.dl) { E.status = STATUS_SHAPE; return; }
  const int m = a.dl * a.d, n = b.d * b.dr, k = a.dr;
  const int64_t off = E.alloc((int64_t)m * n);
  if (E.status) return;
  zgemm(g, m, n, k, zsite(E, a), m, false, zsite(E, b), k, false, false, zsite_at(E, off), m);
  E.work(8.0 * m * n * (double)k, 2.0 * ((double)m * k + (double)k * n + (double)m * n));
  E.set_site(op.a, off, a.dl, a.d * b.d, b.dr);
  if (g.tid == 0) E.st->sites[op.b].live = 0;
  g.sync();
}

// cut_bondim! / cut_bondim_opt! numerics (src/tensor_network.jl:342-477):
// merged = mati * transpose(matj), tsvd(cutoff), myd = min(length(S), Dmax)
// (at least 1); mati' = U sqrt(S), matj' = conj(V) sqrt(S).
CATN_DN void zop_cut(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc A = E.site(op.a), B = E.site(op.b);
  const int dal = A.dl, d = A.d, dar = A.dr, dbl = B.dl, dbr = B.dr;
  if (B.d != d) { E.status = STATUS_SHAPE; return; }
  const int mi = dal * dar, mj = dbl * dbr;
  zd* Mi = zalloc(E.sc, (int64_t)mi * d);
  zd* Mj = zalloc(E.sc, (int64_t)mj * d);
  zd* P = zalloc(E.sc, (int64_t)mi * mj);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  const zd* ap = zsite(E, A);
  const zd* bp = zsite(E, B);
  // mat[a + dal c, b] = A[a, b, c]   (permutedims(A, (1,3,2)) reshaped)
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
  zgemm(g, mi, mj, d, Mi, mi, false, Mj, mj, true, false, P, mi);
  ZSplit sp;
  if (!zsplit(E, P, mi, mj, mi, ZSPLIT_SQRT, 0, op.x, op.d, sp)) return;
  const int myd = sp.k;
  E.acc(&E.st->err, sp.disc);
  const int64_t offa = E.alloc((int64_t)dal * myd * dar);
  const int64_t offb = E.alloc((int64_t)dbl * myd * dbr);
  if (E.status) return;
  zd* na = zsite_at(E, offa);
  zd* nb = zsite_at(E, offb);
  // A'[a, p, c] = mati'[a + dal c, p]; B'[a, p, c] = matj'[a + dbl c, p] = R[p, a + dbl c]
  for (int64_t idx = g.tid; idx < (int64_t)dal * myd * dar; idx += g.nthr) {
    const int a = (int)(idx % dal);
    const int64_t rest = idx / dal;
    const int p = (int)(rest % myd), c = (int)(rest / myd);
    na[idx] = sp.L[a + (int64_t)dal * c + (int64_t)mi * p];
  }
  for (int64_t idx = g.tid; idx < (int64_t)dbl * myd * dbr; idx += g.nthr) {
    const int a = (int)(idx % dbl);
    const int64_t rest = idx / dbl;
    const int p = (int)(rest % myd), c = (int)(rest / myd);
    nb[idx] = sp.R[p + (int64_t)myd * (a + (int64_t)dbl * c)];
  }
  E.work(8.0 * mi * mj * (double)d, 2.0 * ((double)mi * d + (double)mj * d + (double)mi * myd + (double)mj * myd));
  E.set_site(op.a, offa, dal, myd, dar);
  E.set_site(op.b, offb, dbl, myd, dbr);
  if (g.tid == 0) {
    if (op.e >= 0 && op.e < E.st->nresults) E.st->results[op.e] = myd;
    E.st->last_cut = myd;
  }
  g.sync();
}

// compress! / compress_opt! two-site step (src/mps_node.jl:363-372,
// :407-443): tsvd(tl * tr; cutoff, chi), tl <- U S, tr <- V^H
CATN_DN void zop_compress(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc A = E.site(op.a), B = E.site(op.b);
  const int d0 = A.dl, d1 = A.d, dd = A.dr, d2 = B.d, d3 = B.dr;
  if (B.dl != dd) { E.status = STATUS_SHAPE; return; }
  const int m = d0 * d1, n = d2 * d3;
  zd* M = zalloc(E.sc, (int64_t)m * n);
  if (E.sc.fail) { E.status = STATUS_SCRATCH_FULL; return; }
  zgemm(g, m, n, dd, zsite(E, A), m, false, zsite(E, B), dd, false, false, M, m);
  ZSplit sp;
  if (!zsplit(E, M, m, n, m, ZSPLIT_US_V, 0, op.x, op.d, sp)) return;
  const int k = sp.k;
  const int64_t offa = E.alloc((int64_t)m * k);
  const int64_t offb = E.alloc((int64_t)k * n);
  if (E.status) return;
  zcopy(g, zsite_at(E, offa), sp.L, (int64_t)m * k);
  zcopy(g, zsite_at(E, offb), sp.R, (int64_t)k * n);
  E.work(8.0 * m * n * (double)dd, 2.0 * ((double)m * dd + (double)dd * n + (double)m * k + (double)k * n));
  E.set_site(op.a, offa, d0, d1, k);
  E.set_site(op.b, offb, k, d2, d3);
  g.sync();
}

// raw2mps split step (src/mps_node.jl:40-47): the remainder (dleft, rest, 1)
// as (dleft di) x (rest / di); U -> new site, S V^H -> remainder
CATN_DN void zop_raw_split(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc s = E.site(op.a);
  const int dleft = s.dl, rest = s.d, di = op.c;
  if (di <= 0 || rest % di != 0) { E.status = STATUS_SHAPE; return; }
  const int m = dleft * di, n = rest / di;
  ZSplit sp;
  if (!zsplit(E, zsite(E, s), m, n, m, ZSPLIT_U_SV, 0, op.x, op.d, sp)) return;
  const int k = sp.k;
  const int64_t offu = E.alloc((int64_t)m * k);
  const int64_t offr = E.alloc((int64_t)k * n);
  if (E.status) return;
  zcopy(g, zsite_at(E, offu), sp.L, (int64_t)m * k);
  zcopy(g, zsite_at(E, offr), sp.R, (int64_t)k * n);
  E.work(0.0, 2.0 * ((double)m * n + (double)m * k + (double)k * n));
  E.set_site(op.b, offu, dleft, di, k);
  E.set_site(op.a, offr, k, n, 1);
  g.sync();
}

// lognorm of a scalar node (src/mps_node.jl:319-328): z = sum(site);
// log|z| into the network_lognorm sum, sign *= sign(z) = z / |z| (0 at 0)
CATN_DN void zop_scalar_lognorm(Exec& E, const Op& op) {
  const WG& g = E.g;
  SiteDesc s = E.site(op.a);
  const int64_t n = (int64_t)s.dl * s.d * s.dr;
  const zd* x = zsite(E, s);
  double pr = 0.0, pi = 0.0;
  for (int64_t i = g.tid; i < n; i += g.nthr) {
    pr += x[i].re;
    pi += x[i].im;
  }
  const double zr = wg_sum(g, pr);
  const double zi = wg_sum(g, pi);
  const double az = hypot(zr, zi);
  E.acc(&E.st->lntail, log(az));
  if (g.tid == 0) {
    const zd sg = az > 0.0 ? zmk(zr / az, zi / az) : zmk(0.0, 0.0);
    const zd ns = zmul(zmk(E.st->sign, E.st->sign_i), sg);
    E.st->sign = ns.re;
    E.st->sign_i = ns.im;
  }
  g.sync();
}

// one op of a complex network (opcodes without a complex form: STATUS_UNSUPPORTED)
CATN_DN void zrun_op(Exec& E, const Op& op) {
  switch (op.code) {
    case OP_CANO_R: zop_cano_r(E, op); break;
    case OP_CANO_L: zop_cano_l(E, op); break;
    case OP_SWAP: zop_swap(E, op); break;
    case OP_PERMUTE321: zop_permute321(E, op); break;
    case OP_EAT_A: zop_eat_a(E, op); break;
    case OP_EAT_B: zop_eat_b(E, op); break;
    case OP_EAT_C: zop_eat_c(E, op); break;
    case OP_NORMALIZE: zop_normalize(E, op); break;
    case OP_MERGE_FUSE: zop_merge_fuse(E, op); break;
    case OP_CUT: zop_cut(E, op); break;
    case OP_COMPRESS_OPT:
    case OP_COMPRESS: zop_compress(E, op); break;
    case OP_RAW_SPLIT: zop_raw_split(E, op); break;
    case OP_SCALAR_LOGNORM: zop_scalar_lognorm(E, op); break;
    case OP_RESHAPE: op_reshape(E, op); break;
    default: E.status = STATUS_UNSUPPORTED; break;
  }
}

}  // namespace catn
