// Host-side model description, ALE ingest and plan compiler.
#pragma once
#include <stdint.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "plan.hpp"

namespace whale {

struct ModelHost {
  int n = 0;                       // nodes
  int n_species = 0;
  int cond = 0;                    // WHALE_COND_*
  int root = -1;
  int n_wgd = 0;
  int words = 1;                   // 64-bit words per species bitset
  int tab_rows = 0;
  int max_height = 0;
  std::vector<NodeDev> nodes;      // device mirror, index = id - 1
  std::vector<int> order;          // wm.order
  std::vector<int> height;         // leaves 0, else 1 + max over children
  std::vector<int> tab_off;        // n + 1
  std::vector<uint64_t> bits;      // n * words: species under each node (setclade!)
};

// One family after parse_aleobserve + CCD(ale, wm, spmap), 0-based ids.
struct CCDHost {
  std::vector<int32_t> size, leaf_species, split_off, g1, g2;
  std::vector<double> p;
};

// Returns "" on success, else the error message.
std::string parse_ale(const char* text, size_t len, const std::unordered_map<std::string, int>& spmap,
                      CCDHost& out);

struct PlanHost {
  std::vector<FamDev> fam;
  std::vector<StageDev> stage;
  std::vector<SegDev> seg;
  std::vector<uint32_t> it_seg, it_gid, it_rec;
  std::vector<int32_t> it_cf, it_cg, it_bar;
  std::vector<uint32_t> rec_ops, rec_sid;
  std::vector<uint16_t> rec_own;   // stage-local item owning each record
  std::vector<double> rec_p;       // p of each record (the FP64 plain coefficient)
  std::vector<uint16_t> clade_size;
  std::vector<double> split_p;
  std::vector<uint32_t> lvl;
  std::vector<uint32_t> cost_order; // families by decreasing cost
  uint64_t ell_cells = 0, ibar_cells = 0;
  uint32_t max_items = 0;
  uint32_t max_recs = 0;             // most records of any non-root stage
  uint32_t max_root_recs = 0;        // most records of a root stage
  // workload counts
  int64_t n_clades = 0, n_splits = 0, cells = 0, split_terms = 0, sum_compat = 0, row_steps = 0;
  // per-family compat lists in the reference layout, for ℓ download:
  //   for each node: (R_e, list of compatible clades)
  std::vector<uint32_t> fam_seg0, fam_nseg;  // segments of family f: [fam_seg0[f], +fam_nseg[f])
};

std::string build_plan(const ModelHost& m, int n_fam, const int64_t* fam_clade_off, const int32_t* clade_size,
                       const int32_t* leaf_species, const int64_t* split_off, const int32_t* g1,
                       const int32_t* g2, const double* p, PlanHost& out, int nthreads);

}  // namespace whale
