// Host-side symbolic analysis of the fixed KKT pattern, run once per setup().
//
// Replaces the reference's Eigen::AMDOrdering + QDLDL_etree analysis
// (src/conicxx/kkt/qdldl_ldlt.cpp:22-56, regularized_ldlt.cpp:51-95) with a
// supernodal multifrontal plan: fill-reducing ordering, elimination tree,
// postorder, column counts, relaxed supernode partition, front row
// structures, level schedule and every index map the device kernels use.
#pragma once

#include <cstdint>
#include <vector>

namespace cx {

// Symmetric pattern given as the lower triangle in CSC (diagonal included).
struct LowerPattern {
  int N = 0;
  std::vector<int> colptr, rowind;
};

// Approximate minimum degree ordering on the quotient graph.
// Returns perm (new -> old).
std::vector<int> amd_order(const LowerPattern& K);

// Elimination tree of the symmetric matrix permuted by perm (new -> old).
// parent[] in permuted labels, -1 for roots.
std::vector<int> etree_of(const LowerPattern& K, const std::vector<int>& perm);

// Postorder of a forest; with `key`, the children of every node are visited in
// ascending key order.
std::vector<int> postorder(const std::vector<int>& parent, const std::vector<int>& key = {});

// Column counts of L (diagonal included) for the permuted matrix
// (Gilbert-Ng-Peyton; parent/post in permuted labels).
std::vector<int> column_counts(const LowerPattern& K, const std::vector<int>& perm,
                               const std::vector<int>& parent, const std::vector<int>& post);

struct SupernodalPlan {
  int N = 0;
  std::vector<int> perm, iperm;  // final order (new -> old) and its inverse
  // supernodes
  int ns = 0;
  std::vector<int> sn_first, sn_ncol;  // columns [first, first+ncol)
  std::vector<int> row_ptr, rows;      // front row structure, first ncol rows = own columns
  std::vector<int> sn_parent;          // -1 for roots
  std::vector<int> child_ptr, child;   // children lists (CSR)
  std::vector<int> level;              // 0 for leaves
  int nlevels = 0;
  std::vector<int> level_ptr, level_sn;  // supernodes grouped by level
  std::vector<int> col_sn;               // supernode of each permuted column
  // statistics
  double nnzL = 0;      // stored entries incl. diagonal, supernodal (relaxed) storage
  double nnzL_exact = 0;  // sum of true column counts
  double flops = 0;     // sum_j c_j (c_j - 1) over the stored (relaxed) structure
  int max_front = 0;

  int nr(int s) const { return row_ptr[s + 1] - row_ptr[s]; }
};

struct RelaxParams {
  int nrelax0 = 4, nrelax1 = 16, nrelax2 = 48;
  double zrelax0 = 0.8, zrelax1 = 0.1, zrelax2 = 0.05;
  // Width cap for amalgamation.  Wide supernodes are factored in 256-column
  // panels inside one front, so a dense separator is best kept as a single
  // supernode: splitting it into a chain makes every link re-assemble an
  // almost equally large contribution block into the next.
  int max_merge_cols = 1 << 30;
  // Thin fronts absorbed into their parent by a second, equivalent order:
  // at most thin_ncol pivot columns or at most 1/64 of the update matrix's
  // rows, an update matrix of at least thin_min_update rows covering at least
  // half of the parent's rows (thin_ncol = 0: off).
  int thin_ncol = 32, thin_min_update = 1024;
};

SupernodalPlan plan_supernodes(const LowerPattern& K, const RelaxParams& rp = RelaxParams());

}  // namespace cx
