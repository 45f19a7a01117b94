// GPU octree builder (see gt_build.hip).
#pragma once

#include <algorithm>

#include "gt_internal.h"

namespace gth {

// Builds the reference's octree for n particles on the GPU and keeps it in
// device memory: nodes in level order as host-compatible geometry rows
// (kGeomStride doubles: centre, half, com, mass, quadrupole, delta) and links
// (first_child, n_children, begin, end), the tree order and tree-ordered
// positions and masses.
class TreeBuilder {
 public:
  // Allocates everything an evaluation with these bounds needs (so that a
  // later enqueue with the same plan allocates nothing, as graph capture needs).
  void reserve(const Plan &p);
  // Enqueues the build on s with no host round trip.  The node total, the
  // depth and the "more path words needed" flag go to the frame; the host
  // reads them after the evaluation and hands the node count back with
  // set_counts.  A tree with more nodes than p.node_cap is not built.
  // Batches (p.systems > 1) take the system offsets (p.systems + 1 entries,
  // device memory).  orig (device, may be null: the identity) gives the
  // caller's index of each particle in pos; bodies of a leaf are ordered by it.
  // rec (may be null): the walk's and the evaluation's records, written by
  // the build (with each node's moments, each body with its gather into
  // tree order) instead of by prep_lists; a build for the lists also keeps
  // the inverse of its order (inverse_order) for the force scatter.
  // rec: the lists' records to write with the moments and the gather (null:
  // none); groups: the group starts to flag from the nodes' links (null: none).
  void enqueue(const double *pos, const double *mass, const Plan &p, int leaf_size,
               const unsigned long long *sys_off, unsigned long long *frame, hipStream_t s,
               const uint32_t *orig = nullptr, const ListRecords *rec = nullptr,
               const GroupFlags *groups = nullptr);
  bool wrote_records() const { return wrote_records_; }
  bool flagged_groups() const { return flagged_groups_; }  // the last build flagged group starts
  void set_counts(uint32_t n_nodes, int depth) {
    n_nodes_ = n_nodes;
    depth_ = depth;
  }
  uint32_t n() const { return n_; }
  uint32_t n_nodes() const { return n_nodes_; }
  uint32_t node_cap() const { return node_cap_; }
  int depth() const { return depth_; }
  const double *geom() const { return geom_.get(); }
  const int4 *links() const { return lnk_.get(); }
  const uint32_t *order() const { return vals_.get(); }
  uint32_t *order_mutable() { return vals_.get(); }
  // Each particle's tree position (the inverse of order).
  const uint32_t *inverse_order() const { return inv_.get(); }
  uint32_t *inverse_order_mutable() { return inv_.get(); }
  const double *tree_pos() const { return tpos_.get(); }
  const double *tree_mass() const { return tmass_.get(); }
  const double *root() const { return root_.get(); }
  // Per node: its parent (kNone for a root), its depth, and its skip (the
  // node a LIFO depth-first walk takes after the node's subtree; kNone: none).
  const uint32_t *parents() const { return parent_.get(); }
  const uint32_t *depths() const { return node_depth_.get(); }
  const uint32_t *skips() const { return skip_.get(); }
  // Batches only (null for one system): system of each tree position, and the
  // root node of each system.
  const uint32_t *tree_sid() const { return systems_ > 1 ? tsid_.get() : nullptr; }
  const uint32_t *root_nodes() const { return systems_ > 1 ? root_node_.get() : nullptr; }

 private:
  uint32_t n_ = 0, n_nodes_ = 0, node_cap_ = 0, systems_ = 1;
  bool wrote_records_ = false;
  bool flagged_groups_ = false;
  int depth_ = 0;
  DBuf<double> part_, root_, geom_, tpos_, tmass_;
  DBuf<uint64_t> keys_, keys_alt_, word0_, word1_, word2_;
  DBuf<uint32_t> vals_, vals_alt_, level_counts_, level_offsets_, nbeg_, lstart_, parent_, node_depth_,
      skip_, inv_, sid_, tsid_, root_node_;
  DBuf<uint8_t> dg_, lcp_, ldep_;
  DBuf<unsigned long long> sort_tmp_, level_scan_;
  DBuf<int4> lnk_;
};

}  // namespace gth
