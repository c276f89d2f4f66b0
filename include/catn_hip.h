/*
 * libcatn_hip — C ABI of the MI355X contraction engine for CATN.jl networks.
 *
 * The reference (xuanzhaogao/CATN.jl) is pure Julia and has no FFI; its only
 * extension point is the storage type AT of TensorNetwork{T,AT} / MPSNode{T,AT}
 * (src/tensor_network.jl:1, src/mps_node.jl:1) plus Adapt.adapt_structure
 * (src/adapt.jl:6-24). This header is what a Julia backend (ccall) binds in
 * place of that dispatch; each entry point names the reference function whose
 * contract it keeps. All arrays are column-major (Julia layout); node ids and
 * leg positions are 1-based like the reference. Every call returns 0 on
 * success or a negative status; catn_last_error() gives the message.
 *
 * Element type: a network built with catn_params.eltype = 1 is complex
 * (TensorNetwork{ComplexF64}). Its tensor data arguments and outputs are then
 * interleaved (re, im) doubles, the memory layout of Julia's ComplexF64
 * arrays (pass a Ptr{ComplexF64} where `const double*` is declared), and its
 * scalar phase outputs (psi, sign, eat!'s phase) take 2 doubles (re, im).
 *
 * Threading: a context is thread-compatible, not thread-safe (one host
 * thread, one HIP stream), as the reference is single-threaded.
 */
#ifndef CATN_HIP_H
#define CATN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct catn_ctx catn_ctx;
typedef struct catn_net catn_net;

/* Keyword arguments of TensorNetwork / ising_network
 * (src/tensor_network.jl:41-54, src/ising.jl:20-40). */
typedef struct catn_params {
  int32_t Dmax;        /* < 0: exact mode */
  int32_t chi;
  double cutoff;
  int32_t norm_method; /* 0 none, 1 L2, 2 max-abs */
  int32_t select;      /* 0 min-dim, 1 min-dim+triangle, 2 sequential */
  int32_t reverse, svdopt, swapopt, compress, cut_bond;
  int32_t seed;          /* seed of the randomized-SVD sketches */
  int64_t arena_elems;   /* device arena per network (0 = default) */
  int64_t scratch_elems; /* device scratch per network (0 = default) */
  /* randomized-SVD size switch of swap! (src/mps_node.jl:154): rsvd when both
   * theta dimensions exceed rsvd_both or either exceeds rsvd_any (defaults
   * 7000 and 20000, the reference's constants) */
  int32_t rsvd_both, rsvd_any;
  /* 0: FP64 site tensors (Float64). 1: FP32 storage (the reference's
   * Float32 networks, cu(tn)): site tensors rounded to FP32 between ops, ops
   * computing in FP64, the swap theta GEMM as split-BF16 (hi/lo, 3-pass) MFMA
   * products accumulated in FP32. 2: the same with a three-term BF16 split
   * (6 passes, FP32-accurate products). Log-norms accumulate in FP64. */
  int32_t precision;
  /* catn_braket_network only: 1 accepts virtual-bond graphs with cycles (the
   * reference rejects them, src/braket.jl:405-420). When an eat! closes a
   * cycle, the two bonds to the shared neighbour are merged on both nodes, ket
   * and bra legs separately, the older bond as the fast index (merge!,
   * src/mps_node.jl:595-622). Two bonds between one pair of tensors are still
   * rejected. With braket_dmax > 0, a merged ket or bra bond wider than
   * braket_dmax is then cut as cut_bondim_opt! does. */
  int32_t loops;
  /* catn_braket_network only: the Dmax of merged bonds with loops = 1; < 0
   * (the default) cuts nothing. The reference's double layer has no Dmax, so
   * this is opt-in; catn_braket_network ignores `Dmax`. */
  int32_t braket_dmax;
  /* Element type: 0 real (Float64, or Float32 storage per `precision`), 1
   * complex: ComplexF64, or with precision 1 or 2 ComplexF32 storage (FP32
   * (re, im) pairs in the arena, every product computed in FP64; host data
   * and outputs stay interleaved doubles). Complex networks run the plain
   * forward of every entry point below except
   * ising_network (real site tensors) and the reverse sweep (real only),
   * which refuse them with an error that names the element type. */
  int32_t eltype;
} catn_params;

/* Fill `p` with the reference defaults (Dmax=32, chi=32, cutoff=1e-15, ...;
 * braket_dmax = -1, eltype = 0). */
void catn_params_default(catn_params* p);

/* Create a context on HIP device `device` (one stream). */
int catn_create(catn_ctx** ctx, int device);
int catn_destroy(catn_ctx* ctx);
const char* catn_last_error(void);
/* name of the executor backing the context ("hip-gfx950") */
const char* catn_backend(catn_ctx* ctx);

/* ising_network (src/ising.jl:20-227): n spins, edges (2*nedges ints,
 * 1-based, in list order), couplings w[nedges], fields h[n], inverse
 * temperature beta; pinning arguments pos1/val1/pos2/val2 (0 = none). */
int catn_ising_network(catn_ctx* ctx, int32_t n, int32_t nedges, const int32_t* edges,
                       const double* w, const double* h, double beta, int32_t pos1,
                       int32_t val1, int32_t pos2, int32_t val2, const catn_params* p,
                       catn_net** out);

/* ising_network for a batch, with the site values built on the device
 * (src/ising.jl:68-195, site by site). Packed instance format: instance q
 * has inst_n[q] spins and inst_nedges[q] edges; `edges` (2 ints per edge),
 * `w` and `h` hold the instances' edge lists, couplings and fields
 * concatenated in instance order. Network k is instance net_inst[k] at
 * inverse temperature net_beta[k] with pinning net_pins[4k..4k+3] = (pos1,
 * val1, pos2, val2) (NULL: none). The host builds only the integer structure
 * (sites per node in edge-list order); one kernel computes every site of the
 * batch into the networks' arenas. out receives nnets networks, each as
 * catn_ising_network would build it. No reference counterpart (the reference
 * builds one network at a time on the host). */
int catn_ising_network_batch(catn_ctx* ctx, int32_t ninst, const int32_t* inst_n,
                             const int32_t* inst_nedges, const int32_t* edges, const double* w,
                             const double* h, int32_t nnets, const int32_t* net_inst,
                             const double* net_beta, const int32_t* net_pins,
                             const catn_params* p, catn_net** out);

/* TensorNetwork(tensors, ixs; ...) (src/tensor_network.jl:41-140): ntensors
 * tensors, tensor t has ndims[t] legs with sizes dims[...] (concatenated)
 * and data (concatenated, column-major); labels[...] one int64 label per leg.
 * Labels shared by two tensors become bonds; a label seen once is an open leg. */
int catn_tensor_network(catn_ctx* ctx, int32_t ntensors, const int32_t* ndims,
                        const int64_t* dims, const double* data, const int64_t* labels,
                        const catn_params* p, catn_net** out);

/* braket_network(tensors, ixs; chi, cutoff, norm_method, svdopt, swapopt)
 * (src/braket.jl:369-496): the double-layer network <psi|psi> of a ket state
 * given like catn_tensor_network (real or complex tensors, the bra chains
 * built from conj(T); each has exactly one open,
 * physical label; the virtual-bond graph must be acyclic). catn_contraction
 * then runs contraction!(bk) (src/braket.jl:960-1078): lnZ = ln <psi|psi>. */
int catn_braket_network(catn_ctx* ctx, int32_t ntensors, const int32_t* ndims,
                        const int64_t* dims, const double* data, const int64_t* labels,
                        const catn_params* p, catn_net** out);

/* Empty network holding hand-built MPS nodes (for the primitive API below). */
int catn_net_create(catn_ctx* ctx, const catn_params* p, catn_net** out);
/* MPSNode(mps, neighbor, cano, ...) (src/mps_node.jl:17-21): nsites sites,
 * site k has dims (3k..3k+2) = (dl, d, dr) and data (concatenated). */
int catn_net_add_node(catn_net* net, int32_t id, int32_t nsites, const int32_t* dims,
                      const double* data, const int32_t* neighbor, int32_t cano);
/* MPSNode(tensor, neighbor; chi, cutoff) via raw2mps (src/mps_node.jl:23-50) */
int catn_net_add_tensor_node(catn_net* net, int32_t id, int32_t ndim, const int64_t* dims,
                             const double* data, const int32_t* neighbor);
/* Allocate the network's device state and upload its site tensors now
 * (otherwise done lazily by the first call that runs on the device). */
int catn_net_upload(catn_net* net);
/* Allocate the network's reverse-sweep memory now: a tape of tape_cap records
 * (0: the default of catn_gradient_batch) and the adjoint arena, so that a
 * later catn_gradient_batch with the same tape_cap allocates nothing. No
 * counterpart in the reference (its AD allocates inside the pullback). */
int catn_net_prepare_gradient(catn_net* net, int32_t tape_cap);
/* After catn_gradient_batch on a network built from tensors
 * (catn_tensor_network / catn_net_add_tensor_node): d lnZ / d T for its
 * tensor k (0-based, in the order added), `cap` entries (at most the
 * tensor's size) in the tensor's column-major layout, lnZ = log|Z| of the
 * full contraction, so d lnZ / d T = (dZ / dT) / Z. The reverse sweep runs
 * the raw2mps splits (src/mps_node.jl:31-50) backwards. Real networks. The
 * reference has no derivative of contraction!; this extends the reverse sweep
 * of catn_gradient_batch / catn_ising_gradient to tensor networks. */
int catn_net_tensor_adjoint(catn_net* net, int32_t k, double* out, int64_t cap);
/* Test hook of tsvd's fallback (src/linalg_utils.jl:16-21: svd() retried with
 * the QR-iteration driver on a LAPACKException). flags bit 0: every SVD of
 * this network treats its primary path (QR-preconditioned one-sided Jacobi)
 * as not converged, so the fallback (one-sided Jacobi on the operand itself,
 * up to 200 sweeps) runs, as it does on a real non-convergence. Fallbacks
 * taken are counted in catn_net_svd_profile entry 12. Real networks only. */
int catn_net_set_fault(catn_net* net, int32_t flags);
int catn_net_free(catn_net* net);

/* contraction!(tn) -> (lnZ, error, psi) (src/tensor_network.jl:556-717).
 * Consumes the network like the reference. psi: 1 double, or 2 (re, im) for
 * a complex network. */
int catn_contraction(catn_net* net, double* lnZ, double* error, double* psi);
/* Same for `count` independent networks in ONE persistent launch per planning
 * segment (one workgroup per network). Outputs are arrays of length count;
 * for complex networks psi holds 2 doubles per network. Real and complex
 * networks cannot share a batch. */
int catn_contraction_batch(catn_ctx* ctx, int32_t count, catn_net* const* nets, double* lnZ,
                           double* error, double* psi);

/* Chosen form, differentiation: contraction! of `count` networks exactly as
 * catn_contraction_batch (same lnZ, error, psi, same schedule) with a tape of
 * the executed device ops, then the hand-written reverse sweep of lnZ in one
 * launch (one workgroup per network). Afterwards every network holds
 * d lnZ / d (initial site tensors). Replaces the pinning re-contractions of
 * magnetization / correlation (src/ising.jl:263-339) with one backward pass;
 * the reference has no adjoint (no rrule is defined in src/). tape_cap:
 * tape records per network (0: 65536). Networks built by catn_tensor_network
 * (raw2mps at construction) and compress=true runs are rejected by the sweep. */
int catn_gradient_batch(catn_ctx* ctx, int32_t count, catn_net* const* nets, int32_t tape_cap,
                        double* lnZ, double* error, double* psi);
/* d lnZ / d w (nedges), d lnZ / d h (n) and d lnZ / d beta of a network built
 * by catn_ising_network, after catn_gradient_batch: the chain rule through the
 * site formulas of ising_network (src/ising.jl:20-227). Exact contraction gives
 * dw_ij = beta <s_i s_j>, dh_i = beta <s_i> (src/ising.jl:263-339). Any output
 * pointer may be NULL. */
int catn_ising_gradient(catn_net* net, double* dw, double* dh, double* dbeta);
/* d lnZ / d (initial tensor of `slot`), `cap` doubles, column-major (dl,d,dr). */
int catn_net_site_adjoint(catn_net* net, int32_t slot, double* out, int64_t cap);
/* Reverse-sweep profile: wall-clock ticks per opcode (16 entries), the
 * largest deviation of a replayed factorisation from the forward's stored
 * output, and the tape length. */
int catn_net_bwd_profile(catn_net* net, int64_t* ticks, double* replay_dev, int32_t* tape_len);
/* The first `cap` tape records, 16 ints each: opcode, kept count, path
 * (1 certified QR, 0 SVD), flag, aux0, LDS budget, #inputs, #outputs, then
 * (dl, d, dr, arena offset) of inputs 1 and 2; `val` (may be NULL) gets each
 * record's scalar (norm, cutoff, phase or final scalar). */
int catn_net_tape(catn_net* net, int32_t* rec, double* val, int32_t cap);

/* Per-network results after contraction (sign: 2 doubles for a complex
 * network, the product of the final scalar nodes' z / |z|). */
int catn_net_info(catn_net* net, double* sign, int32_t* maxdim_intermediate,
                  int32_t* num_isolated, int32_t* n, double* beta);
/* The selected edge sequence (pairs, 2*count ints); *count in: capacity. */
int catn_net_edge_log(catn_net* net, int32_t* pairs, int32_t* count);
/* Structural trace of contraction! (src/tensor_network.jl:574-708), for
 * inspection and tests; no counterpart in the reference, whose locals these
 * are. `steps` gets 8 ints per step: the selected pair (i, j), the pair after
 * the order(i) >= order(j) swap, whether reverse! was applied to node i and
 * to node j (:608-619), and idx_j_in_i, idx_i_in_j after the reversals.
 * `dups` gets 3 ints per duplicate bond met while re-pointing j's neighbours
 * (:627-650): step (0-based), k, cross. `cuts` gets the kept count of every
 * Dmax cut (cut_bondim_opt!, :399-477) in order. Each *n is in: capacity (in
 * records), out: the number available; any array may be NULL. */
int catn_net_plan_trace(catn_net* net, int32_t* steps, int32_t* nsteps, int32_t* dups,
                        int32_t* ndups, int32_t* cuts, int32_t* ncuts);
/* Device-side profile of the ops executed for this network: wall-clock ticks
 * and counts per opcode (16 entries) and per SVD size bucket (8 entries:
 * min(m,n) <= 4, 8, 16, 32, 64, 128, 256, larger); tick_hz converts ticks. */
int catn_net_profile(catn_net* net, int64_t* op_ticks, int32_t* op_count, int64_t* svd_ticks,
                     int32_t* svd_count, double* tick_hz);
/* Algorithmic work of the ops executed so far (flops by textbook counts,
 * bytes = site tensors read and written once per op) and the device scratch
 * high-water mark in elements. */
int catn_net_work(catn_net* net, double* flops, double* bytes, int64_t* scratch_hwm);
/* Number of interpreter-kernel launches on this context and their summed
 * duration in seconds (HIP events on the context's stream); reset != 0 zeroes. */
int catn_kernel_stats(catn_ctx* ctx, int64_t* launches, double* seconds, int32_t reset);
/* Sub-step ticks (16 entries): SVD path [QR, Jacobi, finish + Q apply, Jacobi
 * sweep count], certified-QR path [QR, inverse bound, Q formation], swap theta,
 * blocked Householder [panels, T factors + trailing updates], blocked Q
 * application, Jacobi sweeps on global-memory operands (count), fallback SVDs
 * taken (count); rest zero. */
int catn_net_svd_profile(catn_net* net, int64_t* sub_ticks);
/* Host time spent planning and in device round trips (launch, copies,
 * waits) since the last catn_kernel_stats reset, in seconds. */
int catn_host_stats(catn_ctx* ctx, double* plan_seconds, double* run_seconds);
/* Launches of the reverse-sweep kernel and their summed duration (HIP events). */
int catn_bwd_kernel_stats(catn_ctx* ctx, int64_t* launches, double* seconds, int32_t reset);
/* Workgroup size of the interpreter kernel: 64, 128, 256 (default), 512 or
   1024 lanes. 512 and 1024 run the wide kernels (up to 16 wavefronts, one
   workgroup per CU), for batches of fewer networks than CUs whose single ops
   are large. */
int catn_set_threads(catn_ctx* ctx, int32_t threads);
/* Launch shape of the interpreter and reverse-sweep kernels: lanes per
   workgroup (64, 128, 256, 512, 1024); LDS budget per workgroup in doubles (0 =
   automatic: 10176 for batches up to 512 networks, else 5056; otherwise
   within [2048, lds_limit]); Householder panel width (8, 16, 32; 0 = 16);
   lanes per column pair of the global-memory Jacobi (8..64; 0 = 16). The
   autotuner (catn_amd.autotune) picks it by timing a batch. No counterpart in
   the reference, whose CPU path has no launch shape; the shape changes
   rounding, not the algorithm or its decisions beyond rounding. */
int catn_set_launch(catn_ctx* ctx, int32_t threads, int32_t lds_doubles, int32_t panel,
                    int32_t jacobi_lanes);
int catn_get_launch(catn_ctx* ctx, int32_t* threads, int32_t* lds_doubles, int32_t* lds_limit,
                    int32_t* panel, int32_t* jacobi_lanes);
/* Streams of a forward catn_contraction_batch of distinct networks (1..16,
   default 4, the HIP runtime's default hardware queues per process; more
   streams than GPU_MAX_HW_QUEUES share queues and serialise). With s > 1 and
   at least 4 networks per stream, the networks are
   dealt (largest first, round robin) into s groups that run concurrently,
   each on its own stream and host thread, so that a group's planning segments
   wait only for its own networks; 1 keeps every launch batch-wide. Batches
   of fewer than 3 network structures (the pinned networks of one or two
   instances) and recorded runs take one stream. No counterpart in the
   reference; results are those of the one-stream path up to the rounding of
   the launch shape. */
int catn_set_batch_streams(catn_ctx* ctx, int32_t streams);
/* Device memory statistics: arena high-water mark and capacity (elements). */
int catn_net_mem(catn_net* net, int64_t* arena_used, int64_t* arena_cap);

/* ---- MPSNode primitives (src/mps_node.jl); node ids / positions 1-based ---- */
int catn_node_swap(catn_net* net, int32_t id, int32_t i, int32_t j, double* error); /* swap! :126 */
int catn_node_cano_to(catn_net* net, int32_t id, int32_t idx);                      /* cano_to! :72 */
int catn_node_left_canonical(catn_net* net, int32_t id);                           /* :106 */
int catn_node_move(catn_net* net, int32_t id, int32_t a, int32_t b, double* error); /* move! :186 */
int catn_node_move2tail(catn_net* net, int32_t id, int32_t idx, double* error);    /* :208 */
int catn_node_move2head(catn_net* net, int32_t id, int32_t idx, double* error);    /* :225 */
int catn_node_reverse(catn_net* net, int32_t id);                                  /* reverse! :238 */
int catn_node_merge(catn_net* net, int32_t id, int32_t j, int32_t cross, double* error); /* merge! :595 */
int catn_node_compress(catn_net* net, int32_t id, int32_t opt, double* error);     /* compress(_opt)! :356/:394 */
/* eat!(node_i, node_j, idx, idxi) -> (lognorm, error, phase) (:463-575);
 * phase: 2 doubles for a complex network */
int catn_node_eat(catn_net* net, int32_t id_i, int32_t id_j, int32_t idx, int32_t idxi,
                  double* lognorm, double* error, double* phase);
/* cut_bondim! / cut_bondim_opt! (src/tensor_network.jl:342, :399) */
int catn_cut_bondim(catn_net* net, int32_t i, int32_t idx_j_in_i, int32_t opt, double* error);
/* Node structure: order, cano, neighbours (cap entries), per-site dims (3*cap) */
int catn_node_get(catn_net* net, int32_t id, int32_t* order, int32_t* cano, int32_t* neighbor,
                  int32_t* dims, int32_t cap);
/* Download site k (1-based) data (dl*d*dr elements; cap in doubles). */
int catn_node_site(catn_net* net, int32_t id, int32_t k, double* out, int64_t cap);
/* result_tensor(tn) (src/tensor_network.jl:504-523), after catn_contraction:
 * the dense tensor of the one surviving node with an open leg (a site whose
 * physical dimension is not 1), its legs in that node's MPS order, column-major
 * with the first leg fastest; the sites are fused on the device. rank, dims[k]
 * and labels[k] (the leg's neighbour entry: the negative sentinel that
 * catn_tensor_network gave the open label, -1 for the first open label in
 * order of appearance, -2 for the next, ...) describe it; dims_cap bounds
 * dims and labels. The full open result is out * exp(lnZ) * psi. A closed
 * network gives rank 0 and out[0] = 1. More than one open-leg node is an
 * error, as in the reference. out = NULL only fills rank, dims and labels;
 * otherwise the surviving node is consumed. */
int catn_result_tensor(catn_net* net, double* out, int64_t cap, int32_t* rank, int32_t* dims,
                       int32_t* labels, int32_t dims_cap);

/* rsvd(A, k, oversample, power; rng) (src/linalg_utils.jl:44-58) on the device:
 * Gaussian sketch of p = min(n, oversample*k) columns drawn by a counter-based
 * generator from `seed` (the reference draws from Julia's RNG), `power`
 * iterations Y = thinQ(A (A' Y)), Q = thinQ(Y), B = Q' A, SVD of B, top
 * kk = min(k, p) triplets. Outputs U (m x kk), S (kk), V (n x kk). */
int catn_rsvd(catn_ctx* ctx, int32_t m, int32_t n, const double* A, int32_t k, int32_t oversample,
              int32_t power, int32_t seed, int32_t* kk, double* U, double* S, double* V);

/* tsvd(A; cutoff, maxdim) (src/linalg_utils.jl:15-28) on the device.
 * A is m x n column-major. Outputs U (m x k), S (min(m,n), full spectrum,
 * descending), V (n x k), k, and the discarded sum. */
int catn_tsvd(catn_ctx* ctx, int32_t m, int32_t n, const double* A, double cutoff,
              int32_t maxdim, int32_t* k, double* U, double* S, double* V, double* discarded);

/* catn_tsvd on device memory with no allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * dA (m x n), dU (m x k), dS (min(m, n)) and dV (n x k) are device pointers the
 * caller owns; every other buffer (instance state, op, descriptors, arena,
 * scratch) lives in the library's fixed device region, a 64 MiB file-scope
 * device array per GPU held until the process ends, which bounds the matrix
 * at about 512 x 512 (larger sizes are refused with an error that gives the
 * bytes needed). Results equal catn_tsvd's on the same matrix. Calls on one
 * GPU share the region: calls that overlap on one GPU (other host threads,
 * other contexts) race and must be serialised by the caller. */
int catn_tsvd_fixed(catn_ctx* ctx, int32_t m, int32_t n, const double* dA, double cutoff,
                    int32_t maxdim, int32_t* k, double* dU, double* dS, double* dV,
                    double* discarded);

#ifdef __cplusplus
}
#endif
#endif /* CATN_HIP_H */
