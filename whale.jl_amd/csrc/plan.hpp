// Execution plan of a batch of gene families, shared by the host compiler
// (host.cpp) and the device kernels (kernels.hip).
//
// The reference evaluates one family by walking the species-tree nodes in
// wm.order and, per node, every compatible clade and every time slice
// (src/core.jl:31-37, 83-128).  The plan regroups that work for one 64-lane
// wavefront per family:
//
//   segment  a (family, node) pair whose compatible-clade set C_e is not
//            empty; its ℓ block is R_e x |C_e| values, row-major (one row =
//            all compatible clades at one slice), so row writes coalesce.
//   stage    all segments of one family at the same height of the species
//            tree (leaves first).  Segments of a stage are independent; their
//            clades form the stage's items, processed row by row in lockstep.
//            Segments that share a clade (possible only on MUL trees) are put
//            in different stages.
//   item     one (segment, compatible clade) pair.  Its split list (records)
//            refers to other items of the same stage by stage-local index, so
//            the row buffers of a stage live in LDS and every operand is an
//            LDS gather; the reverse sweep scatters through the same records.
//   root     the last stage of every family: all clades, one row, processed
//            in clade-size levels (the root term uses the root's own smaller
//            clades, src/core.jl:151-158).
#pragma once
#include <stdint.h>

namespace whale {

constexpr uint32_t kNone = 0xffffffffu;

struct NodeDev {       // per model node, index = reference id - 1
  int32_t kind;        // WHALE_NODE_*
  int32_t child0;
  int32_t child1;
  int32_t tab;         // first row of this node in the slice tables
  int32_t R;           // rows = k + 1
  int32_t wgd;         // q index or -1
  int32_t parent;      // parent index or -1
  int32_t pad;
  double leafP;        // leafℙ
  double delta;        // slice length
};

struct FamDev {
  uint32_t n_clades;
  uint32_t n_stages;     // non-root stages
  uint32_t stage0;       // first stage (absolute)
  uint32_t root_item0;   // first item of the root stage (absolute); items are stage-local from here
  uint32_t clade0;       // first clade (absolute): clade sizes, exponents
  uint32_t split0;       // first split (absolute)
  uint32_t lvl0;         // root size-level boundaries lvl[lvl0 .. lvl0 + n_lvl] (root-stage-local items)
  uint32_t n_lvl;
  uint32_t max_items;    // largest non-root stage (the root stage holds all n_clades)
  uint32_t max_recs;     // most records of any non-root stage
  uint32_t root_seg;     // absolute segment of the root
  uint32_t root_recs;    // records of the root stage (all splits of the family)
  uint32_t rec0;         // first record (absolute); the family's records are contiguous
  uint32_t n_recs;       // records of all its stages and its root
  uint64_t ell0;         // first ℓ cell (absolute)
  uint64_t ibar0;        // first ī cell (absolute)
};

struct StageDev {
  uint32_t item0;        // absolute first item
  uint32_t n_items;
  uint32_t r_max;        // largest R among its segments
  uint32_t seg0;         // absolute first segment
  uint32_t n_segs;
  uint32_t rec0;         // absolute first record (records of its items, owner by owner)
  uint32_t n_recs;
  uint32_t pad;
};

struct SegDev {
  uint32_t node;         // model node index
  uint32_t nc;           // |C_e|
  uint32_t R;            // rows stored (1 for the root)
  uint32_t item0;        // stage-local index of the segment's first item
  uint32_t ell;          // first ℓ cell relative to FamDev::ell0
  uint32_t ibar;         // first ī cell relative to FamDev::ibar0 (slot k of child k: + k*nc), kNone if none
};

// Items: structure of arrays over all items of the batch (absolute index).
//   it_seg[i]   absolute segment
//   it_gid[i]   family-local clade id
//   it_rec[i]   first record; records of item i are [it_rec[i], it_rec[i+1])
//   it_cf[i]    ℓ cell (relative to ell0) of child0's top row for this clade, or -1
//   it_cg[i]    same for child1
//   it_bar[i]   ī cell (relative to ibar0) holding d logL / d top-row value, -1 at the root
// Records (one per split of an item's clade):
//   rec_ops[r]  l1 | l2 << 16, stage-local items of the two parts
//   rec_sid[r]  family-local split id (p = split_p[split0 + sid])
//   rec_own[r]  stage-local item owning the record (records of an item are contiguous)

}  // namespace whale
