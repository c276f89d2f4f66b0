// Host planner: the integer control flow of CATN's contraction engine.
//
// Restates, line by line, the bookkeeping of src/tensor_network.jl (edge
// selection, edge_count buckets, reverse/duplicate logic, Dmax cut decision)
// and of src/mps_node.jl (swap chains, canonical-centre moves, eat!/merge!
// leg bookkeeping). Every numeric step is emitted as an Op for the device
// interpreter; physical leg dimensions are tracked here (they only change by
// merge! products and by Dmax cuts, whose kept count is read back from the
// device), virtual bond dimensions live only on the device.
//
// Documented fixed choices where the reference depends on Julia Dict order:
//   * network node iteration (select=2 scan, network_lognorm sum) is in
//     ascending node id (src/tensor_network.jl:303, :494).
#pragma once
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <array>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ops.h"

namespace catn {

struct Params {
  int Dmax = 32;
  int chi = 32;
  double cutoff = 1e-15;
  int norm_method = 1;
  int select = 1;
  bool reverse = true;
  bool svdopt = true;
  bool swapopt = true;
  bool compress = false;
  bool cut_bond = false;
  int seed = 1;
  // randomized-SVD size switch of swap! (src/mps_node.jl:154): both dims above
  // rsvd_both, or either above rsvd_any
  int rsvd_both = 7000;
  int rsvd_any = 20000;
  // double-layer networks on graphs with cycles (no reference counterpart: the
  // reference rejects them, src/braket.jl:405-420)
  bool loops = false;
};

struct PNode {
  std::vector<int> slot;   // site slot per MPS position
  std::vector<int> nb;     // neighbour id per MPS position (<0: open leg)
  std::vector<int> phys;   // physical dim per MPS position
  std::vector<char> layer; // BraKetNode only: 1 ket leg, 0 bra leg (src/braket.jl:23)
  int cano = 0;            // 1-based canonical centre (reference convention)
  int order() const { return (int)slot.size(); }
  bool operator==(const PNode& o) const {
    return cano == o.cano && slot == o.slot && nb == o.nb && phys == o.phys && layer == o.layer;
  }
};

inline bool same_params(const Params& a, const Params& b) {
  return a.Dmax == b.Dmax && a.chi == b.chi && a.cutoff == b.cutoff &&
         a.norm_method == b.norm_method && a.select == b.select && a.reverse == b.reverse &&
         a.svdopt == b.svdopt && a.swapopt == b.swapopt && a.compress == b.compress &&
         a.cut_bond == b.cut_bond && a.seed == b.seed && a.rsvd_both == b.rsvd_both &&
         a.rsvd_any == b.rsvd_any && a.loops == b.loops;
}

// 64-bit mixing of a sequence of integers (splitmix64 finaliser per word)
struct Mix {
  uint64_t h = 0x9E3779B97F4A7C15ull;
  void add(uint64_t v) {
    uint64_t z = h ^ (v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2));
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    h = z ^ (z >> 31);
  }
  void add(const std::vector<int>& v) {
    add(v.size());
    for (int x : v) add((uint64_t)(uint32_t)x);
  }
};

struct PlanError : std::runtime_error {
  explicit PlanError(const std::string& s) : std::runtime_error(s) {}
};

class Planner {
 public:
  Params p;
  std::map<int, PNode> nodes;   // ascending id
  std::map<int, std::vector<std::pair<int, int>>> edge_count;
  int maxdim_intermediate = -1;
  int nresults = 0;
  std::vector<Op> prog;
  bool acc_err = true;          // whether swap discards count into `error`
  std::vector<std::pair<int, int>> edge_log;   // selected (i, j) per step
  // structural trace of every step (for inspection and tests): 8 ints per
  // step: the selected pair, the pair after the order(i) >= order(j) swap,
  // whether node i / node j was reversed, and idx_j_in_i / idx_i_in_j after
  // the reversals; 3 ints per duplicate bond: step, k, cross; and the kept
  // count of every Dmax cut in order
  std::vector<int32_t> trace_steps, trace_dups, cut_kept;

  // ---- emission helpers ---------------------------------------------------
  void emit(int code, int a = 0, int b = 0, int c = 0, int d = 0, int e = 0, int f = 0,
            double x = 0.0, int g = 0, int h = 0, int i = 0) {
    Op o;
    o.code = code; o.a = a; o.b = b; o.c = c; o.d = d; o.e = e; o.f = f;
    o.g = g; o.h = h; o.i = i; o.pad = 0; o.x = x;
    prog.push_back(o);
  }

  PNode& node(int id) {
    auto it = nodes.find(id);
    if (it == nodes.end()) throw PlanError("no node " + std::to_string(id));
    return it->second;
  }

  // ---- src/mps_node.jl ----------------------------------------------------
  // cano_to! (src/mps_node.jl:72-104); idx 1-based, 0 = last
  void cano_to(PNode& n, int idx) {
    if (idx == 0) idx = n.order();
    if (n.cano == idx) return;
    if (n.cano < idx) {
      for (int i = n.cano; i <= idx - 1; ++i) {
        emit(OP_CANO_R, n.slot[i - 1], n.slot[i], 0, 0, 0, 0, p.cutoff);
        n.cano = i + 1;
      }
    } else {
      for (int i = n.cano; i >= idx + 1; --i) {
        emit(OP_CANO_L, n.slot[i - 1], n.slot[i - 2], 0, 0, 0, 0, p.cutoff);
        n.cano = i - 1;
      }
    }
  }

  void left_canonical(PNode& n) {   // src/mps_node.jl:106
    n.cano = 1;
    cano_to(n, n.order());
  }

  // swap! (src/mps_node.jl:126-178)
  void swap(PNode& n, int i, int j) {
    if (std::abs(i - j) != 1) throw PlanError("swap!: i and j must be consecutive indices");
    if (i < 1 || j < 1 || i > n.order() || j > n.order()) throw PlanError("swap!: indices out of range");
    if (n.cano != i && n.cano != j) {
      const int target = std::abs(n.cano - i) < std::abs(n.cano - j) ? i : j;
      cano_to(n, target);
    }
    const int lo = std::min(i, j), hi = std::max(i, j);
    emit(OP_SWAP, n.slot[lo - 1], n.slot[hi - 1], i < j ? 1 : 0, p.chi, acc_err ? 1 : 0,
         p.swapopt ? 1 : 0, p.cutoff, p.rsvd_both, p.rsvd_any);
    std::swap(n.phys[lo - 1], n.phys[hi - 1]);
    n.cano = j;
    std::swap(n.nb[i - 1], n.nb[j - 1]);
    if (!n.layer.empty()) std::swap(n.layer[i - 1], n.layer[j - 1]);   // src/braket.jl:561
  }

  void move(PNode& n, int a, int b) {   // src/mps_node.jl:186-199
    if (a == b) return;
    if (b > a) {
      for (int idx = a; idx <= b - 1; ++idx) swap(n, idx, idx + 1);
    } else {
      for (int idx = a; idx >= b + 1; --idx) swap(n, idx, idx - 1);
    }
  }

  void move2tail(PNode& n, int idx) {   // src/mps_node.jl:208-217
    const int len = n.order();
    if (idx == len) { cano_to(n, len); return; }
    for (int k = idx; k <= len - 1; ++k) swap(n, k, k + 1);
    cano_to(n, len);
  }

  void move2head(PNode& n, int idx) {   // src/mps_node.jl:225-229
    move(n, idx, 1);
    cano_to(n, 1);
  }

  void reverse(PNode& n) {   // src/mps_node.jl:238-244
    if (n.order() <= 1) return;
    std::reverse(n.slot.begin(), n.slot.end());
    std::reverse(n.nb.begin(), n.nb.end());
    std::reverse(n.phys.begin(), n.phys.end());
    for (int s : n.slot) emit(OP_PERMUTE321, s);
    n.cano = n.order() + 1 - n.cano;
  }

  static int find_neighbor(const PNode& n, int j) {   // src/mps_node.jl:257-260
    for (size_t k = 0; k < n.nb.size(); ++k)
      if (n.nb[k] == j) return (int)k + 1;
    return 0;
  }

  static void add_neighbor(PNode& n, int v, int pos) {   // src/mps_node.jl:267-274
    if (pos == 0) n.nb.push_back(v);
    else n.nb.insert(n.nb.begin() + (pos - 1), v);
  }

  static int delete_neighbor(PNode& n, int v) {   // src/mps_node.jl:282-287
    const int idx = find_neighbor(n, v);
    if (idx == 0) throw PlanError("delete_neighbor!: neighbor " + std::to_string(v) + " not found");
    n.nb.erase(n.nb.begin() + (idx - 1));
    return idx;
  }

  static double logdim(const PNode& n) {   // src/mps_node.jl:299-302
    double s = 0.0;
    for (int d : n.phys) s += log2((double)d);
    return s;
  }
  static double logdim(const PNode& n, int idx) { return log2((double)n.phys[idx - 1]); }

  // eat! (src/mps_node.jl:463-575)
  void eat(PNode& ni, PNode& nj, int idx, int idxi) {
    if (ni.order() == 1) {                       // case (a)
      if (nj.order() != 1) throw PlanError("eat!: case (a) requires a leaf nodej");
      emit(OP_EAT_A, ni.slot[0], nj.slot[0], 0, 0, 0, 0, p.cutoff);
      ni.slot.clear(); ni.phys.clear();
      ni.nb.erase(ni.nb.begin() + (idx - 1));
      return;
    }
    move2tail(ni, idx);
    if (nj.order() == 1) {                       // case (b)
      const int L = ni.order();
      emit(OP_EAT_B, ni.slot[L - 2], ni.slot[L - 1], nj.slot[0]);
      // the folded tensor is stored, then normalised (src/mps_node.jl:506-527)
      emit(OP_NORMALIZE, ni.slot[L - 2], 0, 0, p.norm_method, 0, 0, p.cutoff);
      ni.cano -= 1;
      ni.slot.pop_back(); ni.phys.pop_back();
      ni.nb.pop_back();
      return;
    }
    move2head(nj, idxi);                         // case (c)
    const int L = ni.order();
    emit(OP_EAT_C, ni.slot[L - 2], ni.slot[L - 1], nj.slot[0]);
    ni.slot.pop_back(); ni.phys.pop_back();
    ni.cano = ni.order();
    ni.slot.insert(ni.slot.end(), nj.slot.begin() + 1, nj.slot.end());
    ni.phys.insert(ni.phys.end(), nj.phys.begin() + 1, nj.phys.end());
    ni.nb.pop_back();
    ni.nb.insert(ni.nb.end(), nj.nb.begin() + 1, nj.nb.end());
    cano_to(ni, ni.order());
    emit(OP_NORMALIZE, ni.slot[ni.cano - 1], 0, 0, p.norm_method, 0, 0, p.cutoff);
  }

  // merge! (src/mps_node.jl:595-622)
  void merge(PNode& n, int j, bool cross) {
    std::vector<int> pos;
    for (size_t k = 0; k < n.nb.size(); ++k)
      if (n.nb[k] == j) pos.push_back((int)k + 1);
    if (pos.size() != 2)
      throw PlanError("merge!: expected exactly 2 sites with neighbor " + std::to_string(j));
    const int idx1 = pos[0], idx2 = pos[1];
    const int target = cross ? idx1 : idx1 + 1;
    move(n, idx2, target);
    n.nb.erase(n.nb.begin() + (target - 1));
    cano_to(n, idx1);
    emit(OP_MERGE_FUSE, n.slot[idx1 - 1], n.slot[idx1]);
    n.phys[idx1 - 1] *= n.phys[idx1];
    n.slot.erase(n.slot.begin() + idx1);
    n.phys.erase(n.phys.begin() + idx1);
    // the reference keeps `cano` as is after deleteat! (cano == idx1 here)
    cano_to(n, idx1);
  }

  // result_tensor's mps2raw (src/tensor_network.jl:517-523, src/mps_node.jl
  // mps2raw): the node's sites fused left to right into one site
  // (1, d1*d2*...*dn, 1), first leg fastest; its legs' neighbour entries go with
  // the fused site, so the caller reads them first
  void fuse_all(PNode& n) {
    while (n.order() > 1) {
      emit(OP_MERGE_FUSE, n.slot[0], n.slot[1]);
      n.phys[0] *= n.phys[1];
      n.slot.erase(n.slot.begin() + 1);
      n.phys.erase(n.phys.begin() + 1);
      n.nb.erase(n.nb.begin() + 1);
      if (!n.layer.empty()) n.layer.erase(n.layer.begin() + 1);
    }
    n.cano = 1;
  }

  // compress! / compress_opt! (src/mps_node.jl:356-447)
  void compress_node(PNode& n, bool opt) {
    if (n.order() == 0) return;
    left_canonical(n);
    for (int j = n.order(); j >= 2; --j) {
      const int i = j - 1;
      emit(opt ? OP_COMPRESS_OPT : OP_COMPRESS, n.slot[i - 1], n.slot[j - 1], 0, p.chi, 0, 0,
           p.cutoff);
    }
    n.cano = 1;
  }

  // ---- src/tensor_network.jl ----------------------------------------------
  int dim_after_merge(int i, int j) {   // :152-155 (round: ties to even)
    const PNode& a = node(i);
    const int idx = find_neighbor(a, j);
    const double v = logdim(a) + logdim(node(j)) - 2.0 * logdim(a, idx);
    return (int)nearbyint(v);
  }

  void count_add_edges(const std::vector<std::pair<int, int>>& edges) {   // :167-176
    for (auto& e : edges) edge_count[dim_after_merge(e.first, e.second)].push_back(e);
  }

  void count_add_nodes(const std::vector<int>& ids) {   // :185-197
    std::set<std::pair<int, int>> edges;
    for (int i : ids) {
      if (i <= 0) continue;
      for (int j : node(i).nb) {
        if (j <= 0) continue;
        edges.insert({std::min(i, j), std::max(i, j)});
      }
    }
    count_add_edges(std::vector<std::pair<int, int>>(edges.begin(), edges.end()));
  }

  void count_remove_nodes(const std::vector<int>& ids) {   // :213-225
    for (int id : ids) {
      if (id <= 0) continue;
      const std::vector<int> nbs = node(id).nb;
      for (int nb : nbs) {
        if (nb <= 0) continue;
        const int i = std::min(id, nb), j = std::max(id, nb);
        const int c = dim_after_merge(i, j);
        auto it = edge_count.find(c);
        if (it != edge_count.end()) {
          auto& v = it->second;
          v.erase(std::remove(v.begin(), v.end(), std::make_pair(i, j)), v.end());
        }
      }
    }
  }

  void select_edge_init() {   // :233-236
    edge_count.clear();
    std::vector<int> ids;
    for (auto& kv : nodes) ids.push_back(kv.first);
    count_add_nodes(ids);
  }

  int min_cost() {
    bool found = false;
    int best = 0;
    for (auto& kv : edge_count)
      if (!kv.second.empty()) {
        if (!found || kv.first < best) best = kv.first;
        found = true;
      }
    if (!found) throw PlanError("select: no candidate edges");
    return best;
  }

  std::pair<int, int> select_edge_min_dim() {   // :248-253
    const int c = min_cost();
    maxdim_intermediate = std::max(maxdim_intermediate, c);
    return edge_count[c][0];
  }

  std::pair<int, int> select_edge_min_dim_triangle() {   // :262-288
    const int c = min_cost();
    maxdim_intermediate = std::max(maxdim_intermediate, c);
    const auto& cands = edge_count[c];
    std::pair<int, int> best = cands[0];
    int bestc = -1;
    for (auto& pr : cands) {
      const int i = pr.first, j = pr.second;
      std::set<int> neigh_j;
      for (int v : node(j).nb)
        if (v > 0 && v != i) neigh_j.insert(v);
      int both = 0;
      for (int k : neigh_j)
        if (find_neighbor(node(k), i) > 0) ++both;
      if (both > bestc) { bestc = both; best = pr; }
    }
    return best;
  }

  std::pair<int, int> select_edge_sequentially() {   // :300-316
    bool found = false;
    int bs = 0;
    std::pair<int, int> best;
    for (auto& kv : nodes) {
      const int i = kv.first;
      for (int v : kv.second.nb) {
        if (v <= 0 || !(i < v)) continue;
        if (!found || i + v < bs) { bs = i + v; best = {i, v}; found = true; }
      }
    }
    if (!found) throw PlanError("select_edge_sequentially: no edges remain");
    return best;
  }

  bool has_real_edges() const {   // :535-542
    for (auto& kv : nodes)
      for (int v : kv.second.nb)
        if (v > 0) return true;
    return false;
  }

  // cut_bondim(_opt)! up to the device op (:342-477). Emits the canonical
  // moves and the OP_CUT; returns the result index that receives `myd`.
  int emit_cut(int i, int idx_j_in_i, bool opt) {
    PNode& ni = node(i);
    const int j = ni.nb[idx_j_in_i - 1];
    if (j <= 0)
      throw PlanError("cut_bondim_opt!: leg " + std::to_string(idx_j_in_i) + " of node " +
                      std::to_string(i) + " is an open leg (no neighbor)");
    PNode& nj = node(j);
    const int idx_i_in_j = find_neighbor(nj, i);
    if (idx_i_in_j == 0)
      throw PlanError("cut_bondim_opt!: node " + std::to_string(i) + " not found as neighbor of node " +
                      std::to_string(j));
    if (opt) {
      cano_to(ni, idx_j_in_i);
      cano_to(nj, idx_i_in_j);
    }
    const int r = nresults++;
    emit(OP_CUT, ni.slot[idx_j_in_i - 1], nj.slot[idx_i_in_j - 1], 0, p.Dmax, r, opt ? 1 : 0,
         p.cutoff);
    pending_cut_[0] = i; pending_cut_[1] = idx_j_in_i; pending_cut_[2] = j;
    pending_cut_[3] = idx_i_in_j; pending_cut_[4] = r;
    return r;
  }

  // cut_bondim_opt! (src/tensor_network.jl:399-477) of an explicit leg pair:
  // leg idx_i of node i and leg idx_k of node k carry the same bond (the
  // double layer's merged ket or bra bond, which find_neighbor cannot tell
  // apart); returns the result index that receives `myd`.
  int emit_cut_legs(int i, int idx_i, int k, int idx_k) {
    PNode& ni = node(i);
    PNode& nk = node(k);
    cano_to(ni, idx_i);
    cano_to(nk, idx_k);
    const int r = nresults++;
    emit(OP_CUT, ni.slot[idx_i - 1], nk.slot[idx_k - 1], 0, p.Dmax, r, 1, p.cutoff);
    pending_cut_[0] = i; pending_cut_[1] = idx_i; pending_cut_[2] = k;
    pending_cut_[3] = idx_k; pending_cut_[4] = r;
    return r;
  }

  // apply the device's kept count to both physical legs of the cut bond
  void apply_cut(int myd) {
    cut_kept.push_back(myd);
    node(pending_cut_[0]).phys[pending_cut_[1] - 1] = myd;
    node(pending_cut_[2]).phys[pending_cut_[3] - 1] = myd;
  }
  int pending_cut_result() const { return pending_cut_[4]; }

  // ---- contraction! as a resumable state machine (:556-717) ----------------
  // step() runs until the program needs a device result (after an OP_CUT:
  // returns NEED_CUT) or the contraction is finished (returns DONE).
  enum StepResult { NEED_CUT = 1, DONE = 2 };

  void start() {
    if (!braket && (p.select == 0 || p.select == 1) && edge_count.empty()) select_edge_init();
    state_ = ST_EDGE;
  }

  // ---- src/braket.jl: the double-layer network ----------------------------
  bool braket = false;
  std::vector<std::pair<int, int>> bk_edges;   // paired edges in contraction order

  static int find_leg(const PNode& n, int j, bool isket) {   // :287-294
    for (int k = 0; k < n.order(); ++k)
      if (n.nb[k] == j && (n.layer[k] != 0) == isket) return k + 1;
    return 0;
  }

  // fuse MPS positions q, q+1 (1-based): ein"ijk,kab->ijab" reshaped (dl, d1*d2, dr2)
  void fuse_pair(PNode& n, int q) {
    emit(OP_MERGE_FUSE, n.slot[q - 1], n.slot[q]);
    n.phys[q - 1] *= n.phys[q];
    n.slot.erase(n.slot.begin() + q);
    n.phys.erase(n.phys.begin() + q);
    n.nb.erase(n.nb.begin() + q);
    n.layer.erase(n.layer.begin() + q);
    if (n.cano > q) n.cano -= 1;
  }

  // eat! of the paired (ket, bra) legs of a bond (:638-798). After the two
  // pairs are fused, the four cases are the MPS eat! folds: A = case (a),
  // C = case (b), D = case (c) without moving j's centre to its head.
  void braket_eat(PNode& ni, PNode& nj, int j_id, int i_id) {
    const int n_i = ni.order(), n_j = nj.order();
    const int bi = find_leg(ni, j_id, false);
    if (bi <= 0) throw PlanError("eat!: bra leg to neighbor " + std::to_string(j_id) + " not found in node_i");
    move(ni, bi, n_i);
    const int ki = find_leg(ni, j_id, true);
    if (ki <= 0) throw PlanError("eat!: ket leg to neighbor " + std::to_string(j_id) + " not found in node_i");
    move(ni, ki, n_i - 1);
    cano_to(ni, n_i);
    const int kj = find_leg(nj, i_id, true);
    if (kj <= 0) throw PlanError("eat!: ket leg to neighbor " + std::to_string(i_id) + " not found in node_j");
    move(nj, kj, 1);
    const int bj = find_leg(nj, i_id, false);
    if (bj <= 0) throw PlanError("eat!: bra leg to neighbor " + std::to_string(i_id) + " not found in node_j");
    move(nj, bj, 2);
    fuse_pair(ni, n_i - 1);   // W_i (dl, Dk Db, 1)
    fuse_pair(nj, 1);         // W_j (1, Dk Db, dr)
    if (n_i == 2 && n_j == 2) {                  // case A: scalar, phase z / |z|
      emit(OP_EAT_A, ni.slot[0], nj.slot[0], 0, 0, 0, 0, p.cutoff);
      ni.slot.clear(); ni.phys.clear(); ni.nb.clear(); ni.layer.clear();
      return;
    }
    if (n_i == 2) throw PlanError("eat!: case B (order(i) < order(j)) is not reached by contraction!");
    const int L = ni.order();
    if (n_j == 2) {                              // case C: fold, normalise the folded site
      emit(OP_EAT_B, ni.slot[L - 2], ni.slot[L - 1], nj.slot[0]);
      emit(OP_NORMALIZE, ni.slot[L - 2], 0, 0, p.norm_method, 0, 0, p.cutoff);
      ni.slot.pop_back(); ni.phys.pop_back(); ni.nb.pop_back(); ni.layer.pop_back();
      ni.cano = ni.order();
      return;
    }
    emit(OP_EAT_C, ni.slot[L - 2], ni.slot[L - 1], nj.slot[0]);   // case D
    ni.slot.pop_back(); ni.phys.pop_back(); ni.nb.pop_back(); ni.layer.pop_back();
    ni.cano = ni.order();
    ni.slot.insert(ni.slot.end(), nj.slot.begin() + 1, nj.slot.end());
    ni.phys.insert(ni.phys.end(), nj.phys.begin() + 1, nj.phys.end());
    ni.nb.insert(ni.nb.end(), nj.nb.begin() + 1, nj.nb.end());
    ni.layer.insert(ni.layer.end(), nj.layer.begin() + 1, nj.layer.end());
    cano_to(ni, ni.order());
    emit(OP_NORMALIZE, ni.slot[ni.cano - 1], 0, 0, p.norm_method, 0, 0, p.cutoff);
  }

  // contraction!(bk) (:960-1078): paired edges in order, compress after
  // every eat!, then braket_lognorm (:915-950, ascending node id)
  // Loopy double layer (no reference counterpart): merge! (src/mps_node.jl:
  // 595-622) of the two legs of one layer (ket or bra) that point to `target`,
  // the layer list following the moves. cross: the leg that carried the bond
  // to `target` before the eat comes second, and is moved in front, so that
  // it is the fast index of the fused leg on both ends of the bond.
  void braket_merge(PNode& n, int target, bool isket, bool cross) {
    std::vector<int> pos;
    for (int k = 0; k < n.order(); ++k)
      if (n.nb[k] == target && (n.layer[k] != 0) == isket) pos.push_back(k + 1);
    if (pos.size() != 2)
      throw PlanError("braket merge: expected two " + std::string(isket ? "ket" : "bra") +
                      " legs to node " + std::to_string(target));
    const int idx1 = pos[0], idx2 = pos[1];
    move(n, idx2, cross ? idx1 : idx1 + 1);
    cano_to(n, idx1);
    fuse_pair(n, idx1);
    cano_to(n, idx1);
  }

  StepResult braket_step() {
    for (;;) {
      // pending cuts of merged bonds (loops with Dmax > 0): one per segment
      while (!bk_cuts_.empty()) {
        const std::array<int, 3> c = bk_cuts_.front();
        bk_cuts_.erase(bk_cuts_.begin());
        const int ai = find_leg(node(c[0]), c[1], c[2] != 0);
        const int ak = find_leg(node(c[1]), c[0], c[2] != 0);
        if (ai > 0 && ak > 0 && node(c[0]).phys[ai - 1] > p.Dmax) {
          acc_err = true;
          emit_cut_legs(c[0], ai, c[1], ak);
          return NEED_CUT;
        }
      }
      if (bk_edges.empty()) break;
      int i = bk_edges[0].first, j = bk_edges[0].second;
      if (node(j).order() > node(i).order()) std::swap(i, j);
      edge_log.push_back({i, j});
      acc_err = true;
      braket_eat(node(i), node(j), j, i);
      PNode& ni = node(i);
      compress_node(ni, p.svdopt);
      for (int d : ni.phys) maxdim_intermediate = std::max(maxdim_intermediate, d);
      std::vector<std::pair<int, int>> rest;
      // neighbours of j already bonded to i (a cycle closes): their legs to
      // j, and i's legs to them, are merged below (loops only)
      std::vector<std::pair<int, std::pair<bool, bool>>> dups;   // k, (cross ket, cross bra)
      for (const auto& e : bk_edges) {
        if ((e.first == i && e.second == j) || (e.first == j && e.second == i)) continue;
        if (e.first == j || e.second == j) {   // re-point j's bonds to i (:1020-1071)
          const int other = e.first == j ? e.second : e.first;
          PNode& nk = node(other);
          const int ki = find_leg(nk, i, true);
          if (ki > 0) {
            if (!p.loops)
              throw PlanError("braket_network: a cycle closes at node " + std::to_string(other));
            dups.push_back({other, {ki > find_leg(nk, j, true),
                                    find_leg(nk, i, false) > find_leg(nk, j, false)}});
          }
          for (int& v : nk.nb)
            if (v == j) v = i;
          const auto ne = e.first == j ? std::make_pair(i, other) : std::make_pair(other, i);
          bool seen = false;   // the bond (i, other) is already listed when a cycle closes
          for (const auto& r : rest)
            if ((r.first == i && r.second == other) || (r.first == other && r.second == i)) seen = true;
          for (const auto& r : bk_edges)
            if ((r.first == i && r.second == other) || (r.first == other && r.second == i)) seen = true;
          if (!seen) rest.push_back(ne);
        } else {
          rest.push_back(e);
        }
      }
      bk_edges.swap(rest);
      nodes.erase(j);
      for (const auto& d : dups) {
        PNode& nk = node(d.first);
        braket_merge(nk, i, true, d.second.first);
        braket_merge(nk, i, false, d.second.second);
        PNode& ni2 = node(i);
        braket_merge(ni2, d.first, true, false);
        braket_merge(ni2, d.first, false, false);
        for (int dd : nk.phys) maxdim_intermediate = std::max(maxdim_intermediate, dd);
        for (int dd : ni2.phys) maxdim_intermediate = std::max(maxdim_intermediate, dd);
        if (p.Dmax > 0) {   // the merged ket bond, then the bra bond
          bk_cuts_.push_back({i, d.first, 1});
          bk_cuts_.push_back({i, d.first, 0});
        }
      }
    }
    for (auto& kv : nodes) {
      PNode& n = kv.second;
      if (n.order() == 0) continue;
      cano_to(n, n.order());
      emit(OP_NORMALIZE, n.slot[n.cano - 1], 0, 0, p.norm_method, 0, 0, p.cutoff);
    }
    state_ = ST_DONE;
    return DONE;
  }

  // Planning is a deterministic function of the planner state (and of the
  // cut results applied between segments), so networks whose states are equal
  // get the same next segment. fingerprint() buckets states; same_state()
  // confirms equality field by field before a plan is shared.
  uint64_t fingerprint() const {
    Mix m;
    m.add((uint64_t)p.Dmax); m.add((uint64_t)p.chi); m.add((uint64_t)p.select);
    m.add((uint64_t)p.norm_method); m.add((uint64_t)state_); m.add((uint64_t)cur_i_);
    m.add((uint64_t)cur_j_); m.add((uint64_t)dup_pos_); m.add(dups_);
    m.add((uint64_t)nresults); m.add((uint64_t)(int64_t)maxdim_intermediate);
    m.add((uint64_t)braket); m.add(bk_edges.size()); m.add(bk_cuts_.size());
    for (const auto& c : bk_cuts_) { m.add((uint64_t)c[0]); m.add((uint64_t)c[1]); m.add((uint64_t)c[2]); }
    for (const auto& e : bk_edges) { m.add((uint64_t)e.first); m.add((uint64_t)e.second); }
    m.add(nodes.size());
    for (const auto& kv : nodes) {
      m.add((uint64_t)kv.first); m.add((uint64_t)kv.second.cano);
      m.add(kv.second.slot); m.add(kv.second.nb); m.add(kv.second.phys);
    }
    m.add(edge_count.size());
    for (const auto& kv : edge_count) {
      m.add((uint64_t)kv.first); m.add(kv.second.size());
      for (const auto& e : kv.second) { m.add((uint64_t)e.first); m.add((uint64_t)e.second); }
    }
    return m.h;
  }
  bool same_state(const Planner& o) const {
    return same_params(p, o.p) && state_ == o.state_ && cur_i_ == o.cur_i_ &&
           cur_j_ == o.cur_j_ && dups_ == o.dups_ && dup_pos_ == o.dup_pos_ &&
           cut_pending_ == o.cut_pending_ && nresults == o.nresults &&
           maxdim_intermediate == o.maxdim_intermediate && acc_err == o.acc_err &&
           std::equal(pending_cut_, pending_cut_ + 5, o.pending_cut_) && nodes == o.nodes &&
           edge_count == o.edge_count && edge_log == o.edge_log && prog.size() == o.prog.size() &&
           trace_steps == o.trace_steps && trace_dups == o.trace_dups && cut_kept == o.cut_kept &&
           braket == o.braket && bk_edges == o.bk_edges && bk_cuts_ == o.bk_cuts_;
  }

  StepResult step() {
    if (braket) return state_ == ST_DONE ? DONE : braket_step();
    for (;;) {
      if (state_ == ST_DUPS) {
        if (cut_pending_) {
          cut_pending_ = false;
          // result was applied by the caller via apply_cut
        }
        while (dup_pos_ < (int)dups_.size()) {
          const int k = dups_[dup_pos_++];
          PNode& ni = node(cur_i_);
          const bool save = acc_err;
          acc_err = false;                       // merge!(node_i, k) result is dropped (:662)
          merge(ni, k, false);
          acc_err = save;
          const int idx_k_in_i = find_neighbor(ni, k);
          bool cut = false;
          if (p.cut_bond) cut = true;
          else if (p.Dmax > 0 && ni.phys[idx_k_in_i - 1] > p.Dmax) cut = true;
          if (cut && p.Dmax >= 0) {             // cut_bondim_opt! is a no-op for Dmax < 0
            emit_cut(cur_i_, idx_k_in_i, true);
            cut_pending_ = true;
            return NEED_CUT;
          }
        }
        finish_edge();
        state_ = ST_EDGE;
      }
      if (state_ == ST_EDGE) {
        if (!has_real_edges()) {
          network_lognorm();
          state_ = ST_DONE;
          return DONE;
        }
        begin_edge();
        state_ = ST_DUPS;
        continue;
      }
      return DONE;
    }
  }

  // network_lognorm (:489-501): emitted as scalar lognorm ops
  void network_lognorm() {
    for (auto& kv : nodes) {
      PNode& n = kv.second;
      bool scalar = true;
      for (int d : n.phys) scalar = scalar && d == 1;
      if (!scalar) continue;
      if (n.order() == 0) continue;            // lognorm of an empty MPS = (0, 1)
      if (n.order() != 1)
        throw PlanError("lognorm: computing norm of a non-scalar MPS is not supported in contraction");
      emit(OP_SCALAR_LOGNORM, n.slot[0]);
    }
  }

 private:
  enum State { ST_INIT, ST_EDGE, ST_DUPS, ST_DONE };
  State state_ = ST_INIT;
  int cur_i_ = 0, cur_j_ = 0;
  std::vector<int> dups_;
  int dup_pos_ = 0;
  bool cut_pending_ = false;
  int pending_cut_[5] = {0, 0, 0, 0, 0};
  // loopy double layer: merged bonds (node, neighbour, ket) still to be cut
  std::vector<std::array<int, 3>> bk_cuts_;

  static int cld2(int a) { return (a + 1) / 2; }   // cld(a, 2) for a >= 0

  void begin_edge() {
    std::pair<int, int> e;
    if (p.select == 0) e = select_edge_min_dim();
    else if (p.select == 1) e = select_edge_min_dim_triangle();
    else e = select_edge_sequentially();
    int i = e.first, j = e.second;
    edge_log.push_back(e);
    if (node(j).order() > node(i).order()) std::swap(i, j);   // :585-590
    PNode& ni = node(i);
    PNode& nj = node(j);
    if (p.select == 0 || p.select == 1) {                       // :594-597
      std::vector<int> affected = {i, j};
      affected.insert(affected.end(), ni.nb.begin(), ni.nb.end());
      affected.insert(affected.end(), nj.nb.begin(), nj.nb.end());
      count_remove_nodes(affected);
    }
    std::vector<int> neigh_j = nj.nb;
    int idx_j_in_i = find_neighbor(ni, j);
    int idx_i_in_j = find_neighbor(nj, i);
    bool rev_i = false, rev_j = false;
    if (p.reverse) {                                            // :608-619
      if (idx_j_in_i < cld2((int)ni.nb.size())) {
        reverse(ni);
        idx_j_in_i = find_neighbor(ni, j);
        rev_i = true;
      }
      if (idx_i_in_j >= cld2((int)nj.nb.size())) {
        reverse(nj);
        neigh_j = nj.nb;
        idx_i_in_j = find_neighbor(nj, i);
        rev_j = true;
      }
    }
    const int32_t step_no = (int32_t)edge_log.size() - 1;
    const int32_t rec[8] = {e.first, e.second, i, j, rev_i, rev_j, idx_j_in_i, idx_i_in_j};
    trace_steps.insert(trace_steps.end(), rec, rec + 8);
    dups_.clear();
    dup_pos_ = 0;
    acc_err = true;
    for (int l = 1; l <= (int)neigh_j.size(); ++l) {            // :627-650
      if (l == idx_i_in_j) continue;
      const int k = neigh_j[l - 1];
      if (k <= 0) continue;
      const int idx_k_in_i = find_neighbor(ni, k);
      PNode& nk = node(k);
      const int idx_i_in_k = find_neighbor(nk, i);
      const int idx_j_in_k = delete_neighbor(nk, j);
      add_neighbor(nk, i, idx_j_in_k);
      if (idx_k_in_i > 0) {
        dups_.push_back(k);
        const bool cross = idx_i_in_k > idx_j_in_k;
        trace_dups.push_back(step_no);
        trace_dups.push_back(k);
        trace_dups.push_back(cross ? 1 : 0);
        merge(nk, i, cross);
      }
    }
    eat(ni, nj, idx_j_in_i, idx_i_in_j);                       // :654-657
    cur_i_ = i;
    cur_j_ = j;
  }

  void finish_edge() {
    PNode& ni = node(cur_i_);
    PNode& nj = node(cur_j_);
    nj.slot.clear(); nj.nb.clear(); nj.phys.clear();           // clear!(node_j) :684
    if (p.compress) compress_node(ni, p.svdopt);               // :688-694
    if (p.select == 0 || p.select == 1) {                       // :697-699
      std::vector<int> ids = {cur_i_};
      ids.insert(ids.end(), ni.nb.begin(), ni.nb.end());
      count_add_nodes(ids);
    }
    for (int d : ni.phys) maxdim_intermediate = std::max(maxdim_intermediate, d);   // :702-707
  }
};

}  // namespace catn
