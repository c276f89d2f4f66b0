// Op records of the per-instance programs executed by the device interpreter.
//
// The host planner (planner.h) restates the integer control flow of
// contraction! (src/tensor_network.jl:556-717) and of the MPSNode primitives
// (src/mps_node.jl) and emits one record per numeric primitive. Site tensors
// are addressed by slot ids; a slot's descriptor (arena offset, dl, d, dr)
// lives in device memory because virtual bond dimensions are only known on
// the device (they come out of truncated SVDs).
#pragma once
#include <stdint.h>

namespace catn {

enum OpCode : int32_t {
  OP_NOP = 0,
  // cano_to! single step to the right: SVD of site a as (dl*d) x dr, keep
  // S > cutoff (all if none), U -> a, S*V' folded into site b
  // (src/mps_node.jl:77-87)
  OP_CANO_R = 1,
  // cano_to! single step to the left: SVD of permutedims(a,(2,3,1)) as
  // (d*dr) x dl; U -> a, (S*V')^T folded into site b (src/mps_node.jl:90-100)
  OP_CANO_L = 2,
  // swap! numerics (src/mps_node.jl:144-170): a = left slot (tl), b = right
  // slot (tr), c = 1 if going right (S folded right) else 0, d = chi,
  // e = 1 if the discarded weight counts into the returned error, f = swapopt
  OP_SWAP = 3,
  // reverse! per site: permutedims(t, (3,2,1)) (src/mps_node.jl:240)
  OP_PERMUTE321 = 4,
  // eat! case (a): both leaves, non-conjugating dot (src/mps_node.jl:469-489)
  OP_EAT_A = 5,
  // eat! case (b): a = site end-1, b = site end (contracted), c = leaf j
  // (src/mps_node.jl:492-527); norm_method in d
  OP_EAT_B = 6,
  // eat! case (c) fold: a = site end-1, b = site end of node i, c = head of j
  // (src/mps_node.jl:533-539)
  OP_EAT_C = 7,
  // normalise site a per norm_method d, accumulate log(norm) into lnZ
  // (src/mps_node.jl:558-574)
  OP_NORMALIZE = 8,
  // merge! fuse: a = ein"ijk,kab->ijab"(a,b) reshaped (dl, d1*d2, dr2)
  // (src/mps_node.jl:610-616)
  OP_MERGE_FUSE = 9,
  // cut_bondim_opt! numerics after both cano_to!: a = site of node i,
  // b = site of node j, d = Dmax, e = result index, f = 1 to use the QR
  // pre-reduction (opt) (src/tensor_network.jl:414-476)
  OP_CUT = 10,
  // compress_opt! two-site step: a = site i (left), b = site j (right),
  // d = chi (src/mps_node.jl:401-443)
  OP_COMPRESS_OPT = 11,
  // compress! two-site step (src/mps_node.jl:363-372)
  OP_COMPRESS = 12,
  // raw2mps split step: a = remainder slot viewed (dleft, rest, 1), b = new
  // site slot, c = physical dim of the split leg, d = chi
  // (src/mps_node.jl:40-47)
  OP_RAW_SPLIT = 13,
  // lognorm of a scalar node: z = sum(site a); lnZ += log|z|; sign *= sign(z)
  // (src/mps_node.jl:319-328, src/tensor_network.jl:489-501)
  OP_SCALAR_LOGNORM = 14,
  // standalone truncated SVD of the matrix held in slot a viewed as
  // (dl*d) x dr: U -> slot b (m x k), V -> slot c (n x k), S -> slot f, k ->
  // result e (src/linalg_utils.jl:15-28); d = maxdim. With g = 1 the
  // randomized SVD rsvd(A, d, h, i) instead (src/linalg_utils.jl:44-58).
  OP_TSVD = 15,
  // reinterpret site a (dl, d, dr) as (dl, D, d * dr / D), D = d field, without moving data:
  // the last ket site of braket_node, (chi, D_deg * d_phys, 1) ->
  // (chi, D_deg, d_phys) (src/braket.jl:130)
  OP_RESHAPE = 16,
};

struct Op {
  int32_t code;
  int32_t a, b, c, d, e, f;
  // OP_SWAP: g, h = the randomized-SVD size switch (both dims > g, or either
  // > h; src/mps_node.jl:154). OP_TSVD: g = 1 for rsvd (k = d), h = oversample,
  // i = power iterations (src/linalg_utils.jl:44).
  int32_t g, h, i;
  int32_t pad;
  double x;  // cutoff
};

struct SiteDesc {
  int64_t off;   // element offset into the instance arena
  int32_t dl, d, dr, live;
};

// One tensor version referenced by a tape record (the arena is append-only
// while recording, so every version stays resident for the reverse sweep).
struct TapeSite {
  int64_t off;
  int32_t dl, d, dr;
};

// Tape record of one executed op: its input and output tensor versions and
// the decisions the reverse sweep must replay (factorisation path, kept
// count, norms, phases).
struct TapeRec {
  int32_t code, nin, nout, k;
  int32_t path;    // split ops: 1 certified QR, 0 Jacobi SVD, 2 randomized SVD
  int32_t flag;    // op specific (direction, counted log term, ...)
  int32_t aux0;
  int32_t aux1;    // LDS budget (doubles) of the launch that executed the op
  TapeSite in[3];
  TapeSite out[2];
  double val, val2;
};

// Per-instance device state.
struct InstState {
  double lnZ;      // sum of log-norms (src/tensor_network.jl:657, :714)
  double err;      // sum of discarded singular values (returned `error`)
  double psi;      // product of eat! phases (src/tensor_network.jl:656)
  double sign;     // product of scalar-node signs (tn.sign, :713)
  int64_t arena_top, arena_cap, scratch_cap;
  int32_t status;  // 0 ok, else an error code (see STATUS_*)
  int32_t fail_op;
  double* arena;
  double* scratch;
  SiteDesc* sites;
  int32_t* results;    // per-op integer results (cut myd, tsvd k)
  double* dresults;    // per-op floating results (tsvd discarded)
  int32_t nsites, nresults;
  double lntail;       // network_lognorm sum, added to lnZ once at the end
  // device-side profile: wall-clock ticks and counts per opcode, and per
  // size bucket (min(m, n) of the op's main SVD: <=4,8,16,32,64,128,256,>256)
  int64_t op_ticks[16];
  int32_t op_count[16];
  int64_t svd_ticks[8];
  int32_t svd_count[8];
  // algorithmic work of the executed ops (textbook counts: GEMM 2mnk,
  // Householder QR 2mn^2-2n^3/3, thin SVD 6mn^2+20n^3; bytes = site tensors
  // read once and written once), and the scratch high-water mark (elements)
  double alg_flops, alg_bytes;
  int64_t scratch_hwm;
  // sub-step ticks: SVD path 0 QR, 1 Jacobi, 2 finish + Q apply, 3 sweeps;
  // certified-QR path 4 QR, 5 inverse bound, 6 Q formation; 7 swap theta;
  // blocked Householder 8 panel factorisations, 9 T factors + trailing
  // updates; 10 blocked Q application; 11 Jacobi sweeps in global memory
  int64_t sub_ticks[16];
  int32_t last_cut;    // kept count of the last OP_CUT (read by the planner)
  int32_t record;      // 1: write a tape record per executed op
  TapeRec* tape;       // tape records (device), tape_cap entries
  double* adj;         // adjoint arena, same indexing as `arena`
  int32_t tape_len, tape_cap;
  // reverse sweep: ticks per opcode of the replayed records, and the largest
  // deviation of a replayed factorisation from the forward's stored output
  int64_t bwd_ticks[16];
  double replay_dev;
  int32_t lds_max;     // largest LDS budget (doubles) of the launches that ran ops
  int32_t seed;        // seed of the randomized-SVD sketches (TensorNetwork seed)
  // 1 or 2: site tensors are stored in FP32 (the arena holds floats at the
  // same element offsets; the adjoint arena stays FP64), ops compute in FP64
  // and the swap theta GEMM runs as split-BF16 products accumulated in FP32,
  // with 2 (3 passes) or, for 2, 3 BF16 terms per operand (6 passes)
  int32_t f32;
  // 1: complex site tensors (ComplexF64), interleaved (re, im) doubles at
  // twice the element offset; psi and sign then carry imaginary parts
  int32_t cplx;
  double psi_i, sign_i;
  // test hook (catn_net_set_fault): bit 0 makes every SVD's primary path
  // count as not converged, so that the fallback SVD runs
  int32_t fault;
};

enum Status : int32_t {
  STATUS_OK = 0,
  STATUS_ARENA_FULL = 1,
  STATUS_SCRATCH_FULL = 2,
  STATUS_SHAPE = 3,
  STATUS_NOCONV = 4,
  STATUS_UNSUPPORTED = 5,
  STATUS_TAPE_FULL = 6,
  STATUS_REPLAY = 7,     // the reverse sweep could not replay a forward step
};

}  // namespace catn
