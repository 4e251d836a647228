// Device-side views of the model, the batch plan and the per-call workspace,
// and the launch wrappers the C ABI calls.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "plan.hpp"

namespace whale {

struct BatchDev {
  const FamDev* fam;
  const StageDev* stage;
  const SegDev* seg;
  const uint32_t* it_seg;
  const uint32_t* it_gid;
  const uint32_t* it_rec;
  const int32_t* it_cf;
  const int32_t* it_cg;
  const int32_t* it_bar;
  const uint32_t* rec_ops;
  const uint32_t* rec_sid;
  const uint16_t* rec_own;
  const double* rec_p;
  const uint16_t* clade_size;
  const double* split_p;
  const uint32_t* lvl;
  const uint32_t* order;   // families in processing order
  uint32_t n_fam;
};

struct ModelDev {
  const NodeDev* node;
  const double* eps;
  const double* phi;
  const double* psi;
  const double* par;       // [λ(n), μ(n), p(n), q(n_wgd), η]
  int n;
  int n_wgd;
  int root;
  int tab_rows;
};

// type-erased workspace (ell, coef, ibar hold float or double)
struct WorkRaw {
  void* ell;
  void* coef;
  int32_t* E;
  void* ibar;
  double* fam_ll;
  double* acc;
  uint32_t* ticket;
  uint32_t* flag;     // per family: 1 if the FP32 form left its range (recomputed in FP64)
  uint32_t* fb;       // fb[0] = number of such families, fb[1..] = their indices
  unsigned char* gws; // stage buffers of launches whose families do not fit LDS (LaunchCfg::gws)
  double* acc_fam = nullptr;  // deterministic gradients: per-family accumulator slots
  // per-family weights of the reverse sweep (null: 1): family f's adjoint is
  // seeded with seed[f] / L_f, and seed[n_fam] = Σ seed scales the
  // conditioning term's derivative (mixture responsibilities)
  double* seed = nullptr;
};

// gradient accumulator layout (doubles)
struct AccLayout {
  int phi, psi, eps, q, rootA, rootB, n;
  __host__ __device__ AccLayout(int tab_rows, int nodes, int nwgd) {
    phi = 0;
    psi = tab_rows;
    eps = 2 * tab_rows;
    q = eps + nodes;
    rootA = q + nwgd;
    rootB = rootA + 1;
    n = rootB + 1;
  }
};

__host__ __device__ constexpr size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// LDS of one wave, laid out afresh for every stage from its item count n and
// record count nr (s = bytes per Real).  The same functions size the launches
// on the host and carve the buffers on the device.
//   forward stage:  bA bB bF sAcc [n] | sCoef sP [nr] | sOps [nr] | sOwn [nr] (u16) |
//                   sRO [n+1] | sSegL [n] (u16) | sEi [n] | sFlag [n] (scaled forms)
//   forward root:   bA bF bB [n] | sCoef [nr] | sOps [nr] | sOwn [nr] (u16) | sRO [n+1] | sEi [n] (scaled)
//   reverse stage:  bA bB bV bG sFo [n] | sCoef [nr] | sOps [nr] | sOwn [nr] (u16) |
//                   sRO [n+1] | sSegL [n] (u16)
//   reverse root:   bA bB bV bG [n] | sCoef [nr] | sOps [nr] | sOwn [nr] (u16) | sRO [n+1]
// (root records are staged only when they fit: nr = 0 reads them from global memory)
__host__ __device__ inline size_t fwd_stage_bytes(size_t s, bool scaled, size_t n, size_t nr) {
  return align16(4 * n * s) + align16(2 * nr * s) + align16(4 * nr) + align16(2 * nr) + align16(4 * (n + 1)) +
         align16(2 * n) + (scaled ? align16(4 * n) + align16(n) : 0);
}
__host__ __device__ inline size_t fwd_root_bytes(size_t s, bool scaled, size_t n, size_t nr) {
  return align16(3 * n * s) + align16(nr * s) + align16(4 * nr) + align16(2 * nr) + align16(4 * (n + 1)) +
         (scaled ? align16(4 * n) : 0);
}
__host__ __device__ inline size_t bwd_stage_bytes(size_t s, size_t n, size_t nr) {
  return align16(5 * n * s) + align16(nr * s) + align16(4 * nr) + align16(2 * nr) + align16(4 * (n + 1)) +
         align16(2 * n);
}
__host__ __device__ inline size_t bwd_root_bytes(size_t s, size_t n, size_t nr) {
  return align16(4 * n * s) + align16(nr * s) + align16(4 * nr) + align16(2 * nr) + align16(4 * (n + 1));
}

// LDS one wave needs for family d (s bytes per value): its largest non-root
// stage (items and records bounded separately) and its root stage, root
// records staged or not
__host__ __device__ inline size_t fam_lds_bytes(const FamDev& d, size_t s, bool scaled, bool fwd, bool root_recs) {
  const size_t nr = root_recs ? d.root_recs : 0;
  const size_t a = fwd ? fwd_stage_bytes(s, scaled, d.max_items, d.max_recs) : bwd_stage_bytes(s, d.max_items, d.max_recs);
  const size_t b = fwd ? fwd_root_bytes(s, scaled, d.n_clades, nr) : bwd_root_bytes(s, d.n_clades, nr);
  return a > b ? a : b;
}

// The pooled launch (k_pool): one workgroup of kPoolWaves waves per compute
// unit; after the shared area, a header (lock and page bitmap) and a pool of
// kPage-byte pages from which each family takes what both its directions need.
constexpr int kPoolWaves = 12;
constexpr uint32_t kPage = 512;
constexpr uint32_t kPoolHdrBytes = 256;  // lock + bitmap of up to 2016 pages
// bytes a family takes from the pool: the root records are staged when that
// costs at most an eighth of the pool more
__host__ __device__ inline size_t pool_need(const FamDev& d, size_t s, bool scaled, size_t pool_bytes) {
  const size_t w0 = fam_lds_bytes(d, s, scaled, true, false), w1 = fam_lds_bytes(d, s, scaled, false, false);
  const size_t r0 = fam_lds_bytes(d, s, scaled, true, true), r1 = fam_lds_bytes(d, s, scaled, false, true);
  const size_t without = w0 > w1 ? w0 : w1, with = r0 > r1 ? r0 : r1;
  return (with <= pool_bytes && with - without <= pool_bytes / 8) ? with : without;
}

// LDS of one launch of k_forward / k_backward: every wave of a workgroup runs
// its own family in `wave_bytes` of LDS (laid out per stage); `shared_bytes`
// (slice tables, reverse-sweep accumulators, node constants) come first and
// are shared by the workgroup's waves.
struct LdsCfg {
  uint32_t wave_bytes;
  uint32_t shared_bytes;
};

// a launch: `fpw` families (waves) per workgroup, slice tables in LDS or not.
// With `gws` the waves' stage buffers (wave_bytes each) live in a global
// workspace instead of LDS, one wave per workgroup: the path of families whose
// largest stage does not fit one compute unit's LDS.
struct LaunchCfg {
  int grid = 1;
  int fpw = 1;
  int waves_per_cu = 0;  // resident waves per compute unit the LDS allows
  bool tabl = false;
  bool gws = false;
  LdsCfg L{};
  size_t lds() const { return (size_t)L.shared_bytes + (gws ? 0 : (size_t)fpw * L.wave_bytes); }
  size_t gws_bytes() const { return gws ? (size_t)grid * fpw * L.wave_bytes : 0; }
};

size_t fwd_shared_bytes(int mode, bool tabl, int tab_rows, int n, int nwgd);
size_t bwd_shared_bytes(int mode, bool tabl, int tab_rows, int acc_n, int n, int nwgd);

// forward/backward launch: `mode` is WHALE_FP64_PLAIN / WHALE_FP32_SCALED /
// WHALE_FP64_SCALED; fam_single >= 0 runs that one family (first wave of a
// 1-block grid), otherwise waves take families order[fam_begin .. fam_end)
// from the ticket.  fam_count, if not null, replaces fam_end with a count read
// on the device (the FP64 pass over families listed by the FP32 pass).
// One component of a mixture launch: with K components the grid is K times
// the launch's, and workgroup w works for component w % K with that
// component's model (tables), workspace (tickets included), and, when not
// null, its own family order and count (the FP64 pass over families its
// FP32 pass flagged).
struct MixComp {
  ModelDev M;
  WorkRaw W;
  const uint32_t* order = nullptr;
  const uint32_t* count = nullptr;
};
constexpr int kMaxMix = 8;

hipError_t launch_forward(int mode, bool store_all, const LaunchCfg& C, hipStream_t st, const BatchDev& B,
                          const ModelDev& M, const WorkRaw& W, int32_t fam_single, uint32_t fam_begin,
                          uint32_t fam_end, const uint32_t* fam_count = nullptr, const MixComp* X = nullptr,
                          int K = 0);
hipError_t launch_backward(int mode, const LaunchCfg& C, hipStream_t st, const BatchDev& B, const ModelDev& M,
                           const WorkRaw& W, int32_t fam_single, uint32_t fam_begin, uint32_t fam_end,
                           const uint32_t* fam_count = nullptr, const MixComp* X = nullptr, int K = 0);
// per family i: ℓ_ij = fam_ll_j[i] + log prior_j − c_j over the K components
// (c_j the component's conditioning term), r_ij = exp(ℓ_ij − logsumexp_j ℓ_ij)
// into seed_j[i] and Σ_i r_ij into seed_j[n_fam]; out[0] = Σ_i logsumexp_j ℓ_ij
// (−inf if not finite); resp (n_fam × K, may be null) = r
hipError_t launch_mix(hipStream_t st, const MixComp* X, int K, const double* log_prior, uint32_t n_fam, int cond,
                      double* out, double* resp);
// per-wave slots of the fused forward + reverse launch: each holds one
// family's stored rows (from byte 0), ī (from ibar_off) and record
// coefficients (from coef_off)
struct FusedSlots {
  unsigned char* base = nullptr;
  uint64_t slot_bytes = 0;
  uint64_t ibar_off = 0;
  uint64_t coef_off = 0;
};
hipError_t launch_fused(int mode, const LaunchCfg& C, hipStream_t st, const BatchDev& B, const ModelDev& M,
                        const WorkRaw& W, uint32_t fam_begin, uint32_t fam_end, const uint32_t* fam_count, int grad,
                        const FusedSlots& SL);
// the pooled launch over families order[fam_begin .. fam_end): C.grid
// workgroups of kPoolWaves waves, C.L.shared_bytes of shared LDS (the reverse
// sweep's layout, tables in LDS) and C.L.wave_bytes of pool (header included)
hipError_t launch_pool(int mode, bool store_all, const LaunchCfg& C, hipStream_t st, const BatchDev& B,
                       const ModelDev& M, const WorkRaw& W, uint32_t fam_begin, uint32_t fam_end, int grad);
hipError_t launch_tables(hipStream_t st, const ModelDev& M, double* eps, double* phi, double* psi, double* drow,
                         double* dnode, const int32_t* hlevels, const int32_t* hnodes, int n_levels);
// n_eff (null: n_fam) holds the family count the conditioning term's
// derivative is scaled by (Σ of the reverse sweep's family weights)
hipError_t launch_tail(hipStream_t st, const ModelDev& M, const double* fam_ll, uint32_t n_fam, const double* acc,
                       double* ebar, const double* drow, const double* dnode, const int32_t* dlevels,
                       const int32_t* dnodes, int n_dlevels, int cond, int with_grad, double* out,
                       const double* n_eff = nullptr);
hipError_t launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, int n);
hipError_t launch_det_sum(hipStream_t st, const double* slots, uint32_t n_fam, int ncol, double* out);
hipError_t launch_fill_f64(hipStream_t st, double* p, double v, size_t n);

// What the backtracking sampler reads: the stored ℓ (FP32 or FP64 values,
// plus the FP64 region of families an FP32 call evaluated in FP64), the
// clades' exponents, each (family, node)'s segment, each segment's first
// item (its clades in increasing id), and every clade's split list.
struct BtDev {
  const FamDev* fam;
  const SegDev* seg;
  const uint32_t* it_gid;
  const uint32_t* seg_item;   // absolute first item of each segment
  const int32_t* node_seg;    // [family][node]: segment, or -1 (no compatible clade)
  const uint32_t* spl_off;    // per absolute clade: its splits [spl_off[c], spl_off[c + 1])
  const uint32_t* spl_g;      // per split: g1 | g2 << 16 (family-local clade ids)
  const double* split_p;
  const uint16_t* clade_size;
  const void* ell;
  const double* ell64;
  const int32_t* E;           // null in the plain form
  const uint8_t* fam64;       // per family: 1 if its ℓ is in ell64 (null: none)
  int ell_float;              // the main region holds FP32 values
};
hipError_t launch_backtrack(hipStream_t st, const BtDev& D, const ModelDev& M, uint32_t fam_begin, uint32_t n_fams,
                            uint32_t n_samples, uint64_t seed, uint32_t cap, int32_t* counts, int32_t* nodes,
                            int32_t* stack);

// The fixed per-GPU workspace of whale_logpdf_noalloc (a file-scope device
// array in kernels.hip) and its device address on `device`, which must be the
// current device; found once per device and kept.
constexpr size_t kScratchBytes = 64ull << 20;
constexpr int kMaxDevices = 64;
hipError_t scratch_address(int device, void** out);

}  // namespace whale
