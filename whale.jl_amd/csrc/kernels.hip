// gfx950 kernels of the amalgamated-likelihood DP.
//
//   k_tables       slice tables ε, φ, ψ of every branch (setmodel!/setslices!,
//                  src/model.jl:162-191; getα/_ϵ/_ϕ/_ψ, src/bdputil.jl:6-11)
//                  plus the per-row partial derivatives the reverse sweep needs
//   k_forward      the DP of every family (whale!/whalewgd!/within_branch!/
//                  whaleroot!, src/core.jl:83-158), one family per wavefront
//                  (several per workgroup), stage by stage, the items of a
//                  stage in lockstep row by row, row buffers in LDS
//   k_backward     reverse sweep of the same DAG: d logL / d ℓ, accumulated
//                  into d logL / d(φ, ψ, ε, q, A, B) of the slice tables
//   k_tail         table adjoint (reverse of the ε recursion, chain rule to
//                  λ, μ, q, η), conditioning (src/condition.jl:21-29) and the
//                  batch sum ℓhood(Σℓ − N·c) (src/core.jl:54,63)
//
// Precision: Real = double (plain: exactly the reference's recursion) or
// float/double with per-clade power-of-two exponents (Scaled).  In the scaled
// forms the value of clade γ is stored as v = ℓ·2^-E(γ) with one exponent per
// (family, clade), fixed the first time γ gets a non-zero value.  Because
// every term of the recursion multiplies ℓ(γ1)·ℓ(γ2) into ℓ(γ) or carries
// ℓ(γ) itself, all scaling folds into one coefficient per split,
// p·2^(E(γ1)+E(γ2)-E(γ)); rows then run on plain Real arithmetic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "dev.hpp"
#include "../../include/whale_amd.h"

namespace whale {

constexpr int kUndef = -0x40000000;  // exponent not yet fixed
constexpr int kWG = 64;              // one wave per family
// A clade's exponent is fixed at its first non-zero value, which can be far
// smaller than its later values (e.g. a large in-paralog clade grows by ~2^120
// along one branch as more slices become available for its duplications).
// New non-leaf clades therefore start at 2^-kBirth, leaving 2^(128+kBirth) of
// growth and 2^-(126-kBirth) of decay before FP32 runs out of range; a family
// that still leaves the range is recomputed in FP64 (see k_forward's flag).
constexpr int kBirth = 40;
// A clade that had no value before the current stage ("young") may still grow
// by many powers of two within this stage's branch (a large in-paralog clade
// gains a factor per slice as more slices hold its duplications).  Because it
// has no stored values elsewhere, its exponent can be raised in place: when
// its FP32 value passes 2^kBump the clade is renormalised and the rows it
// already wrote on this branch are rescaled.  The window [2^-kBirth, 2^kBump]
// is kept narrow because a split's coefficient p·2^(E1+E2-E) spans the ratio
// of the parent's value to the product of its parts' values (two windows and
// one 1/ψ): 2^(kBump + 2·kBirth + log2(1/ψ)) must stay inside FP32.  A narrow
// window also ends the stage low, with room to grow in the stages above.
constexpr int kBump = 16;
constexpr uint8_t kPending = 1;  // exponent still open (all values so far are zero)
constexpr uint8_t kYoung = 2;    // exponent was open when the stage began
constexpr uint8_t kShifted = 4;  // renormalised in this row; exponent update pending

template <typename Real>
struct WorkDev {
  Real* ell;
  Real* coef;         // per record (scaled forms)
  int32_t* E;         // per clade (scaled forms)
  Real* ibar;
  double* fam_ll;
  double* acc;        // gradient accumulators
  uint32_t* ticket;
  uint32_t* flag;
  uint32_t* fb;
  unsigned char* gws;
  double* acc_fam;    // deterministic gradients: per-family accumulator slots (null otherwise)
  double* seed;       // per-family weights of the reverse sweep (null: 1)
  __host__ __device__ explicit WorkDev(const WorkRaw& w)
      : ell((Real*)w.ell), coef((Real*)w.coef), E(w.E), ibar((Real*)w.ibar), fam_ll(w.fam_ll), acc(w.acc),
        ticket(w.ticket), flag(w.flag), fb(w.fb), gws(w.gws), acc_fam(w.acc_fam), seed(w.seed) {}
};

__device__ __forceinline__ double eps_top(const ModelDev& M, int node) {
  const NodeDev& d = M.node[node];
  return M.eps[d.tab + d.R - 1];
}

template <typename Real>
__device__ __forceinline__ Real frexp_r(Real x, int* e);
template <>
__device__ __forceinline__ float frexp_r<float>(float x, int* e) { return frexpf(x, e); }
template <>
__device__ __forceinline__ double frexp_r<double>(double x, int* e) { return frexp(x, e); }
template <typename Real>
__device__ __forceinline__ Real ldexp_r(Real x, int e);
template <>
__device__ __forceinline__ float ldexp_r<float>(float x, int e) { return ldexpf(x, e); }
template <>
__device__ __forceinline__ double ldexp_r<double>(double x, int e) { return ldexp(x, e); }

// running sum of terms m·2^k with unknown spread: keeps (sum, exponent)
template <typename Real>
struct XSum {
  Real s;
  int k;
  __device__ XSum() : s(0), k(kUndef) {}
  __device__ void add(Real m, int e) {  // adds m·2^e, m >= 0
    if (m == Real(0)) return;
    int em;
    Real mm = frexp_r(m, &em);
    e += em;
    if (k == kUndef) { s = mm; k = e; return; }
    if (e > k) { s = ldexp_r(s, k - e) + mm; k = e; }
    else s += ldexp_r(mm, e - k);
  }
};

// product p·x1·x2 with x1 = v1·2^e1, x2 = v2·2^e2 as (mantissa, exponent); v1, v2 != 0
template <typename Real>
__device__ __forceinline__ void xprod(Real p, Real v1, int e1, Real v2, int e2, Real* m, int* e) {
  int a, b;
  Real m1 = frexp_r(v1, &a), m2 = frexp_r(v2, &b);
  *m = p * m1 * m2;
  *e = e1 + e2 + a + b;
}

// ---------------------------------------------------------------------------
// k_tables: one workgroup.  The slice map over Δt is the Möbius map
//   ε ← (a + (1 − a − b) ε) / (1 − b ε)
// (the linear birth-death extinction pgf, src/model.jl:162-191); r slices
// compose to a map of the same kind, which eps_after evaluates in closed
// form.  The tops of the branches go level by level (a node's bottom value is
// a function of its children's tops), one lane per node; then every row of
// every branch is independent: ε(r−1) in closed form, then the last step and
// the row's φ, ψ and derivatives as the slice recursion writes them.
// ---------------------------------------------------------------------------
// deriv layout per table row (9 doubles): dε/dα, dε/dβ, dε/dε', dφ/dα, dφ/dβ,
// dφ/dε', dψ/dα, dψ/dβ, dψ/dε'.  Per node (4 doubles): dα/dλ, dα/dμ, dβ/dλ, dβ/dμ.
struct SliceAB {
  double a, b, dal, dam, dbl, dbm;
};

// α, β of one slice of length t (and their derivatives in λ, μ)
__device__ inline SliceAB slice_ab(double l, double u, double t) {
  SliceAB s{0, 0, 0, 0, 0, 0};
  bool close = (l == u) || (isfinite(l) && isfinite(u) && fabs(l - u) <= 1e-6);
  if (close) {
    double x = 1.0 + l * t;
    s.a = l * t / x;
    s.dal = t / (x * x);
    s.dam = 0.0;
  } else {
    double E = exp(t * (l - u));
    double N = u * (E - 1.0), D = l * E - u;
    s.a = N / D;
    double dNl = u * t * E, dNm = (E - 1.0) - u * t * E;
    double dDl = E + l * t * E, dDm = -l * t * E - 1.0;
    s.dal = (dNl * D - N * dDl) / (D * D);
    s.dam = (dNm * D - N * dDm) / (D * D);
  }
  s.b = (l / u) * s.a;
  s.dbl = s.a / u + (l / u) * s.dal;
  s.dbm = -l * s.a / (u * u) + (l / u) * s.dam;
  return s;
}

// ε after k slices from e0.  The map has the fixed points 1 and ρ = a/b
// (= μ/λ), and z = 1/(1 − ε) moves affinely: z ← κz + b/(1 − a) with
// κ = (1 − b)/(1 − a).  With δ = κ − 1, S = (κ^k − 1)/δ:
//   ε_k = (e0 + S(a − b·e0)/(1 − a)) / (κ^k + S·b(1 − e0)/(1 − a)),
// both sums of non-negative terms while e0 ≤ ρ; for e0 > ρ (a leaf whose
// sampling probability p exceeds μ/λ)
//   (ε_k − ρ)/(1 − ε_k) = κ^k (e0 − ρ)/(1 − e0)
// keeps ε_k = ρ + (excess) free of cancellation.  κ^k and S go through
// log1p/expm1 (κ is near 1 for short slices) and are scaled by κ^−k when
// κ > 1.  Accurate to a few ulp, where k dependent steps drift by k ulp.
__device__ inline double eps_after(double a, double b, int k, double e0) {
  if (k == 0 || e0 == 1.0) return e0;
  const double om = 1.0 - a;
  if (om == 0.0) return 1.0;
  const double d = (a - b) / om;
  if (a < b * e0) {
    const double rho = a / b;
    const double W = exp((double)k * log1p(d)) * (e0 - rho) / (1.0 - e0);
    return (rho + W) / (1.0 + W);
  }
  const double al = (a - b * e0) / om, be = b * (1.0 - e0) / om;
  if (d == 0.0) return (e0 + (double)k * al) / (1.0 + (double)k * be);
  const double L = (double)k * log1p(d);
  if (L <= 0.0) {
    const double S = expm1(L) / d;
    return (e0 + S * al) / (exp(L) + S * be);
  }
  const double SK = -expm1(-L) / d;  // S/κ^k
  return (e0 * exp(-L) + SK * al) / (1.0 + SK * be);
}

// one slice step from ε' = ep (the row's ε, φ, ψ)
__device__ inline void slice_step(double a, double b, double ep, double* en, double* ph, double* ps, double* den_) {
  const double den = 1.0 - b * ep;
  const double num = a + (1.0 - a - b) * ep;
  *en = num / den;
  *ph = (1.0 - a) * (1.0 - b) / (den * den);
  *ps = (1.0 - a) * (1.0 - b) * b / (den * den * den);
  *den_ = den;
}

__device__ inline double node_e0(const ModelDev& M, const NodeDev& d, int e) {
  const double* p = M.par + 2 * M.n;
  const double* q = M.par + 3 * M.n;
  if (d.kind == WHALE_NODE_WGD) {
    double ec = eps_top(M, d.child0);
    double qq = q[d.wgd];
    return qq * (ec * ec) + (1.0 - qq) * ec;
  }
  if (d.kind == WHALE_NODE_LEAF) return p[e];
  return eps_top(M, d.child0) * eps_top(M, d.child1);
}

__global__ __launch_bounds__(1024) void k_tables(ModelDev M, double* eps, double* phi, double* psi, double* drow, double* dnode,
                         const int32_t* hlevels, const int32_t* hnodes, int n_levels) {
  extern __shared__ int32_t s_tab[];  // first table row of each node, then tab_rows
  const double* lam = M.par;
  const double* mu = M.par + M.n;
  for (int e = threadIdx.x; e < M.n; e += blockDim.x) s_tab[e] = M.node[e].tab;
  if (threadIdx.x == 0) s_tab[M.n] = M.tab_rows;
  // tops, children first
  for (int L = 0; L < n_levels; ++L) {
    for (int k = hlevels[L] + threadIdx.x; k < hlevels[L + 1]; k += blockDim.x) {
      const int e = hnodes[k];
      const NodeDev& d = M.node[e];
      const int t0 = d.tab;
      const double e0 = node_e0(M, d, e);
      eps[t0] = e0;
      phi[t0] = 1.0;
      psi[t0] = 1.0;
      SliceAB s{0, 0, 0, 0, 0, 0};
      if (d.R > 1) s = slice_ab(lam[e], mu[e], d.delta);
      dnode[4 * e + 0] = s.dal;
      dnode[4 * e + 1] = s.dam;
      dnode[4 * e + 2] = s.dbl;
      dnode[4 * e + 3] = s.dbm;
      if (d.R > 1) {
        double en, ph, ps, den;
        slice_step(s.a, s.b, eps_after(s.a, s.b, d.R - 2, e0), &en, &ph, &ps, &den);
        eps[t0 + d.R - 1] = en;
      }
    }
    __syncthreads();
  }
  // every row r ≥ 1 of every branch
  for (int t = threadIdx.x; t < M.tab_rows; t += blockDim.x) {
    int lo = 0, hi = M.n;  // the node e with s_tab[e] ≤ t < s_tab[e + 1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (s_tab[mid] <= t) lo = mid;
      else hi = mid;
    }
    const int e = lo;
    const int r = t - s_tab[e];
    if (r == 0) continue;
    const NodeDev& d = M.node[e];
    const SliceAB s = slice_ab(lam[e], mu[e], d.delta);
    const double a = s.a, b = s.b;
    const double ep = eps_after(a, b, r - 1, eps[d.tab]);
    double en, ph, ps, den;
    slice_step(a, b, ep, &en, &ph, &ps, &den);
    if (r == d.R - 1)
      en = eps[t];  // the top, as the levels wrote it for the parent
    else
      eps[t] = en;
    phi[t] = ph;
    psi[t] = ps;
    double* g = drow + 9 * (size_t)t;
    double d2 = den * den, d3 = d2 * den, d4 = d3 * den;
    g[0] = (1.0 - ep) / den;
    g[1] = ep * (en - 1.0) / den;
    g[2] = ((1.0 - a - b) + b * en) / den;
    g[3] = -(1.0 - b) / d2;
    g[4] = -(1.0 - a) / d2 + (1.0 - a) * (1.0 - b) * 2.0 * ep / d3;
    g[5] = (1.0 - a) * (1.0 - b) * 2.0 * b / d3;
    g[6] = -(1.0 - b) * b / d3;
    g[7] = (1.0 - a) * (1.0 - 2.0 * b) / d3 + 3.0 * ep * (1.0 - a) * (1.0 - b) * b / d4;
    g[8] = (1.0 - a) * (1.0 - b) * b * 3.0 * b / d4;
  }
}

// ---------------------------------------------------------------------------
// wave-level reductions
// ---------------------------------------------------------------------------
template <typename Real>
__device__ __forceinline__ Real wave_sum(Real v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWG);
  return v;
}

// Sums of a value over runs of equal keys within each 16-lane row of the wave
// (records of one item sit on consecutive lanes), with DPP row shifts: the
// last lane of a run then holds the run's sum (*last), so one LDS atomic per
// run and row replaces up to 16 on the same address.  All lanes take part.
template <int OFF>
__device__ __forceinline__ float dpp_shr(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x110 + OFF, 0xf, 0xf, false));
}
template <int OFF>
__device__ __forceinline__ double dpp_shr(double v) {
  const long long x = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(x & 0xffffffffll), 0x110 + OFF, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), 0x110 + OFF, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int OFF, typename Real>
__device__ __forceinline__ void row_seg_step(Real& v, uint32_t key) {
  const Real vu = dpp_shr<OFF>(v);
  const uint32_t ku = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)key, 0x110 + OFF, 0xf, 0xf, false);
  if (ku == key) v += vu;
}
template <typename Real>
__device__ __forceinline__ Real row_seg_sum(Real v, uint32_t key, bool* last) {
  row_seg_step<1>(v, key);
  row_seg_step<2>(v, key);
  row_seg_step<4>(v, key);
  row_seg_step<8>(v, key);
  const uint32_t kn = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)key, 0x101, 0xf, 0xf, false);  // row_shl:1
  *last = kn != key || (threadIdx.x & 15) == 15;
  return v;
}

// The same sums for records held in registers across rows: which row shifts
// stay inside the lane's run, whether the lane ends it and how many shifts
// the longest run of the 64 lanes needs depend on the keys only, so they are
// found once per stage.  Each row's sum is then at most four DPP adds under
// those lane masks, and none past what the longest run needs (most items
// have one or two splits).
// The masks are held as 0/1 multipliers (one FMA per shift, no lane masks
// to keep in scalar registers); a term past the range (inf) can then spread
// a NaN within its 16 lanes, which the range checks catch like the inf.
struct RunMask {
  float s1, s2, s4, s8;  // 1 where the shift stays inside the lane's run
  uint32_t last;         // 1 where the lane ends its run
  int steps;             // shifts the longest run needs (wave-uniform)
  __device__ bool ends() const { return last != 0u; }
};
__device__ __forceinline__ RunMask run_mask(uint32_t key) {
  const bool b1 = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)key, 0x111, 0xf, 0xf, false) == key;
  const bool b2 = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)key, 0x112, 0xf, 0xf, false) == key;
  const bool b4 = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)key, 0x114, 0xf, 0xf, false) == key;
  const bool b8 = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)key, 0x118, 0xf, 0xf, false) == key;
  const uint32_t kn = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)key, 0x101, 0xf, 0xf, false);  // row_shl:1
  const bool real = key != kNone;
  RunMask m;
  m.s1 = b1 ? 1.0f : 0.0f;
  m.s2 = b2 ? 1.0f : 0.0f;
  m.s4 = b4 ? 1.0f : 0.0f;
  m.s8 = b8 ? 1.0f : 0.0f;
  m.last = (real && (kn != key || (threadIdx.x & 15) == 15)) ? 1u : 0u;
  // a run reaching back 2^k lanes also reaches back every shorter distance
  m.steps = __any(real && b8) ? 4 : __any(real && b4) ? 3 : __any(real && b2) ? 2 : __any(real && b1) ? 1 : 0;
  return m;
}
template <typename Real>
__device__ __forceinline__ Real run_sum(Real v, const RunMask& m) {
  // the count is uniform: read it into a scalar register so that the tests
  // are scalar branches, not exec-mask changes around each shift
  const int st = __builtin_amdgcn_readfirstlane(m.steps);
  if (st > 0) v = fma(dpp_shr<1>(v), (Real)m.s1, v);
  if (st > 1) v = fma(dpp_shr<2>(v), (Real)m.s2, v);
  if (st > 2) v = fma(dpp_shr<4>(v), (Real)m.s4, v);
  if (st > 3) v = fma(dpp_shr<8>(v), (Real)m.s8, v);
  return v;
}

// atomic adds without return: LDS (ds_add) or global (FP64 atomics in L2)
template <typename Real>
__device__ __forceinline__ void lds_add(Real* p, Real v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <bool Lds>
__device__ __forceinline__ void acc_add(double* p, double v) {
  if (v == 0.0) return;
  if (Lds) lds_add(p, v);
  else unsafeAtomicAdd(p, v);
}

// coefficient of a record in the scaled forms: p·2^(E1+E2-E), 0 while any exponent is open
template <typename Real>
__device__ __forceinline__ Real coef_of(Real p, int e1, int e2, int Eg) {
  if (Eg == kUndef || e1 == kUndef || e2 == kUndef) return Real(0);
  return ldexp_r(p, e1 + e2 - Eg);
}

// One wavefront evaluates one family; a workgroup holds blockDim.x / 64 such
// waves (at most kMaxFPW), each on its own family, sharing the slice tables (and, in the reverse sweep, the
// gradient accumulators) in LDS.  All synchronisation inside a family is
// therefore wave-level.  The LDS executes a wave's instructions in order, so
// between rows it is enough to keep the compiler from moving LDS accesses
// across the boundary: no s_barrier and no wait on outstanding global stores.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Where a wave reads back global memory its lanes wrote (ℓ and ī of earlier
// stages, exponents) the stores must have completed: once per stage.
__device__ __forceinline__ void wave_sync_mem() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// The synchronisation between a wave's passes over its stage buffers: those
// buffers are LDS, or (Gws) the wave's region of a global workspace for
// families whose stages do not fit one compute unit's LDS, where every pass
// must also wait for the previous pass's global stores.
template <bool Gws>
__device__ __forceinline__ void wsync() {
  if (Gws) wave_sync_mem();
  else wave_sync();
}

// φ_r, ψ_r of table row `row`: from the workgroup's LDS copy or from global memory
template <typename Real, bool TabL>
__device__ __forceinline__ void slice(const ModelDev& M, const Real* sTab, uint32_t row, Real& ph, Real& ps) {
  if (TabL) {
    typedef Real Real2 __attribute__((ext_vector_type(2)));
    const Real2 t = reinterpret_cast<const Real2*>(sTab)[row];
    ph = t.x;
    ps = t.y;
  } else {
    ph = (Real)M.phi[row];
    ps = (Real)M.psi[row];
  }
}

// ---------------------------------------------------------------------------
// the segments of a stage, one per lane
// ---------------------------------------------------------------------------
// Row 0 and the register caches need each item's segment and node.  A stage's
// segments are loaded once, lane s holding segment s, and items fetch theirs
// with a lane shuffle instead of a chain of dependent global loads.  Stages of
// more than 64 segments read them from global memory.
// row-invariant data of one item in the row loops
struct ItemRow {
  uint32_t cell;    // ℓ cell of row 0, relative to the family
  uint32_t stride;  // |C_e| of the item's segment (row pitch)
  uint32_t tab;     // first table row of the segment's node
  uint32_t R;       // rows of the segment
};

struct SegInfo {
  uint32_t ell, nc, R, tab, item0, ibar;
  int32_t node, kind, c0, c1, wgd;
};

__device__ __forceinline__ SegInfo seg_info(const BatchDev& B, const ModelDev& M, uint32_t si) {
  const SegDev sg = B.seg[si];
  const NodeDev& nd = M.node[sg.node];
  SegInfo t;
  t.ell = sg.ell;
  t.nc = sg.nc;
  t.R = sg.R;
  t.tab = (uint32_t)nd.tab;
  t.item0 = sg.item0;
  t.ibar = sg.ibar;
  t.node = (int32_t)sg.node;
  t.kind = nd.kind;
  t.c0 = nd.child0;
  t.c1 = nd.child1;
  t.wgd = nd.wgd;
  return t;
}

// A generic pointer to global memory as a global-address-space pointer: its
// loads and stores are global ones, which LDS accesses need not wait for (a
// generic access counts against both the vector-memory and the LDS
// counters, so a later wait for LDS also waits for it).  Only for pointers
// known to address global memory.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> glob(T* p) {
  return (gptr<T>)p;
}

// items (records) per lane that a stage's set-up loads at once
constexpr int kStageU = 4;

// A record every lane of the wave loaded from the same address, moved to
// scalar registers.  The kernels write global memory, so the compiler loads
// such records with vector loads and treats their fields as varying across
// lanes; tests on them (item and record counts, row counts) would then be
// exec-mask branches instead of scalar ones.
template <typename T>
__device__ __forceinline__ T uniform(const T& x) {
  static_assert(sizeof(T) % 4 == 0, "a record of whole dwords");
  uint32_t w[sizeof(T) / 4];
  __builtin_memcpy(w, &x, sizeof(T));
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) w[k] = (uint32_t)__builtin_amdgcn_readfirstlane((int)w[k]);
  T y;
  __builtin_memcpy(&y, w, sizeof(T));
  return y;
}

struct SegTab {
  SegInfo mine;
  uint32_t seg0;
  bool inreg;
  __device__ void load(const BatchDev& B, const ModelDev& M, uint32_t seg0_, uint32_t nsg, uint32_t lane) {
    seg0 = seg0_;
    inreg = nsg <= (uint32_t)kWG;
    if (inreg && lane < nsg) mine = seg_info(B, M, seg0 + lane);
  }
  // segment `sl` (stage-local); every lane of the wave must call this together
  __device__ __forceinline__ SegInfo get(const BatchDev& B, const ModelDev& M, uint32_t sl) const {
    if (!inreg) return seg_info(B, M, seg0 + sl);
    const int src = (int)sl;
    SegInfo t;
    t.ell = __shfl(mine.ell, src, kWG);
    t.nc = __shfl(mine.nc, src, kWG);
    t.R = __shfl(mine.R, src, kWG);
    t.tab = __shfl(mine.tab, src, kWG);
    t.item0 = __shfl(mine.item0, src, kWG);
    t.ibar = __shfl(mine.ibar, src, kWG);
    t.node = __shfl(mine.node, src, kWG);
    t.kind = __shfl(mine.kind, src, kWG);
    t.c0 = __shfl(mine.c0, src, kWG);
    t.c1 = __shfl(mine.c1, src, kWG);
    t.wgd = __shfl(mine.wgd, src, kWG);
    return t;
  }
  // the row-loop data of item i of segment `sl` (items past the register
  // cache); every lane of the wave must call this together
  __device__ __forceinline__ ItemRow row(const BatchDev& B, const ModelDev& M, uint32_t sl, uint32_t i) const;
  // first table row of segment `sl`'s node; every lane must call this together
  __device__ __forceinline__ uint32_t tab(const BatchDev& B, const ModelDev& M, uint32_t sl) const {
    if (!inreg) return (uint32_t)M.node[B.seg[seg0 + sl].node].tab;
    return __shfl(mine.tab, (int)sl, kWG);
  }
};

// per-node constants of the current parameters in LDS: ε at the top of each
// branch, leafℙ, and q per WGD
template <typename Real>
struct NodeC {
  Real *etop, *leafP, *q;
  __device__ NodeC(unsigned char* base, int n) {
    etop = reinterpret_cast<Real*>(base);
    leafP = etop + n;
    q = leafP + n;
  }
  __device__ void fill(const ModelDev& M, int tid, int nthreads) {
    for (int e = tid; e < M.n; e += nthreads) {
      etop[e] = (Real)eps_top(M, e);
      leafP[e] = (Real)M.node[e].leafP;
    }
    for (int w = tid; w < M.n_wgd; w += nthreads) q[w] = (Real)M.par[3 * M.n + w];
  }
};

__host__ __device__ inline size_t nodec_bytes(size_t real, int n, int nwgd) {
  return align16((2 * (size_t)n + (size_t)(nwgd > 0 ? nwgd : 1)) * real);
}

// ---------------------------------------------------------------------------
// row-invariant data of one item, cached in registers for a stage's rows
// ---------------------------------------------------------------------------
// Items per lane held in registers by the forward (kCache) and the reverse
// sweep (kCacheB): 128 per stage; items past them fetch their segment data
// by lane shuffles.  Measured: larger caches (up to 384 items) were slower,
// their unrolled per-slot code and registers cost more than they save.
constexpr int kCache = 2;
constexpr int kCacheB = 2;

__device__ __forceinline__ ItemRow item_row(const SegInfo& sg, uint32_t i) {
  ItemRow it;
  it.cell = sg.ell + (i - sg.item0);
  it.stride = sg.nc;
  it.tab = sg.tab;
  it.R = sg.R;
  return it;
}

__device__ __forceinline__ ItemRow SegTab::row(const BatchDev& B, const ModelDev& M, uint32_t sl, uint32_t i) const {
  if (!inreg) return item_row(seg_info(B, M, seg0 + sl), i);
  const int src = (int)sl;
  ItemRow it;
  it.cell = __shfl(mine.ell, src, kWG) + (i - __shfl(mine.item0, src, kWG));
  it.stride = __shfl(mine.nc, src, kWG);
  it.tab = __shfl(mine.tab, src, kWG);
  it.R = __shfl(mine.R, src, kWG);
  return it;
}

// clade-size level boundaries of a family's root stage: the first 128 held
// in two registers across the wave and read with readlane (uniform level)
struct LevelBounds {
  uint32_t lo, hi;
  __device__ LevelBounds(const BatchDev& B, const FamDev& F, uint32_t lane) {
    lo = lane <= F.n_lvl ? B.lvl[F.lvl0 + lane] : 0u;
    hi = lane + kWG <= F.n_lvl ? B.lvl[F.lvl0 + kWG + lane] : 0u;
  }
  __device__ __forceinline__ uint32_t at(const BatchDev& B, const FamDev& F, uint32_t L) const {
    if (L < (uint32_t)kWG) return (uint32_t)__builtin_amdgcn_readlane((int)lo, (int)L);
    if (L < 2u * kWG) return (uint32_t)__builtin_amdgcn_readlane((int)hi, (int)(L - kWG));
    return B.lvl[F.lvl0 + L];
  }
};

// LDS of one wave for a forward stage of n items and nr records (the layout of
// fwd_stage_bytes) and for the root (fwd_root_bytes)
template <typename Real>
struct FwdStage {
  Real *bA, *bB, *bF, *sAcc, *sCoef, *sP;
  uint32_t *sOps, *sRO;
  uint16_t *sOwn, *sSegL;
  int32_t* sEi;
  uint8_t* sFlag;
  __device__ FwdStage(unsigned char* p, size_t n, size_t nr) {
    const size_t s = sizeof(Real);
    bA = reinterpret_cast<Real*>(p);
    bB = bA + n;
    bF = bB + n;
    sAcc = bF + n;
    p += align16(4 * n * s);
    sCoef = reinterpret_cast<Real*>(p);
    sP = sCoef + nr;
    p += align16(2 * nr * s);
    sOps = reinterpret_cast<uint32_t*>(p);
    p += align16(4 * nr);
    sOwn = reinterpret_cast<uint16_t*>(p);
    p += align16(2 * nr);
    sRO = reinterpret_cast<uint32_t*>(p);
    p += align16(4 * (n + 1));
    sSegL = reinterpret_cast<uint16_t*>(p);
    p += align16(2 * n);
    sEi = reinterpret_cast<int32_t*>(p);
    p += align16(4 * n);
    sFlag = reinterpret_cast<uint8_t*>(p);
  }
};

template <typename Real>
struct FwdRoot {
  Real *bA, *bF, *bB, *sCoef;
  uint32_t *sOps, *sRO;
  uint16_t* sOwn;
  int32_t* sEi;
  __device__ FwdRoot(unsigned char* p, size_t n, size_t nr) {
    const size_t s = sizeof(Real);
    bA = reinterpret_cast<Real*>(p);
    bF = bA + n;
    bB = bF + n;
    p += align16(3 * n * s);
    sCoef = reinterpret_cast<Real*>(p);
    p += align16(nr * s);
    sOps = reinterpret_cast<uint32_t*>(p);
    p += align16(4 * nr);
    sOwn = reinterpret_cast<uint16_t*>(p);
    p += align16(2 * nr);
    sRO = reinterpret_cast<uint32_t*>(p);
    p += align16(4 * (n + 1));
    sEi = reinterpret_cast<int32_t*>(p);
  }
};

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
constexpr int kMaxFPW = 8;  // families (waves) per workgroup, at most
// resident waves per SIMD the DP kernels are compiled for (registers ≤ 512/2):
// the stage buffers in LDS hold about two waves per SIMD of most size
// classes anyway, and at three (≤ 168 registers) the reverse sweep and the
// FP64 forms spill to scratch
constexpr int kWavesPerEU = 2;

// a mixture launch's workgroup takes its component's model, workspace and,
// when given, family order and count
template <typename Real>
__device__ inline void mix_select(const MixComp& x, BatchDev& B, ModelDev& M, WorkDev<Real>& W,
                                  const uint32_t*& fam_count) {
  M = x.M;
  W = WorkDev<Real>(x.W);
  if (x.order) B.order = x.order;
  if (x.count) fam_count = x.count;
}

// The forward DP of family f by one wave: its stages (row buffers in the
// wave's region wb), then its root.  ℓ is written through `ell` (the
// family's cells, relative indices) and record coefficients through `coef`
// (absolute record indices).
template <typename Real, bool Scaled, bool StoreAll, bool TabL, bool Gws>
__device__ __forceinline__ void fwd_family(const BatchDev& B, const ModelDev& M, const WorkDev<Real>& W,
                                           const LdsCfg& L, uint32_t f, unsigned char* const wb, Real* sTab,
                                           const NodeC<Real>& NC, uint32_t lane, Real* ell_arg, Real* coef_arg,
                                           uint32_t coef_base) {
  const gptr<Real> ell = glob(ell_arg);
  const gptr<Real> coef = glob(coef_arg);
  constexpr int RC = 2;  // records per lane held in registers (128 per stage)
  const FamDev F = uniform(B.fam[f]);
  const gptr<int32_t> E = glob(W.E + F.clade0);
  if (Scaled)
    for (uint32_t c = lane; c < F.n_clades; c += kWG) E[c] = kUndef;
  // ---------------- non-root stages
  for (uint32_t s = 0; s < F.n_stages; ++s) {
    const StageDev S = uniform(B.stage[F.stage0 + s]);
    const uint32_t n = S.n_items, i0 = S.item0;
    const uint32_t rb = S.rec0, nrec = S.n_recs;
    const FwdStage<Real> T(wb, n, nrec);  // fits the wave's LDS (size class)
    Real* const bA = T.bA;
    Real* const bB = T.bB;
    Real* const bF = T.bF;
    Real* const sAcc = T.sAcc;
    Real* const sCoef = T.sCoef;
    Real* const sP = T.sP;
    uint32_t* const sOps = T.sOps;
    uint16_t* const sOwn = T.sOwn;
    uint32_t* const sRO = T.sRO;
    uint16_t* const sSegL = T.sSegL;
    int32_t* const sEi = T.sEi;
    uint8_t* const sFlag = T.sFlag;
    wave_sync_mem();  // ℓ and exponents of the stages below
    SegTab ST;
    ST.load(B, M, S.seg0, S.n_segs, lane);
    // kStageU items (records) per lane at once: all their index loads, then
    // all the loads of ℓ and exponents they name (in-bounds dummies past n),
    // then the LDS writes.  The first batches of records and items go
    // together, so a stage of up to 256 items and 256 records is set up in
    // two dependent round trips to memory.
    uint32_t op[kStageU];
    uint16_t ow[kStageU];
    double pp[kStageU];
    auto rec_load = [&](uint32_t k00) {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t k = k00 + u * kWG + lane;
        const uint32_t ka = rb + (k < nrec ? k : 0u);
        op[u] = B.rec_ops[ka];
        ow[u] = B.rec_own[ka];
        pp[u] = B.rec_p[ka];
      }
    };
    auto rec_put = [&](uint32_t k00) {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t k = k00 + u * kWG + lane;
        if (k < nrec) {
          sOps[k] = op[u];
          sOwn[k] = ow[u];
          const Real p = (Real)pp[u];
          sCoef[k] = p;  // the scaled forms rescale it after row 0
          sP[k] = p;
        }
      }
    };
    int32_t cf[kStageU], cg[kStageU];
    uint32_t sg[kStageU], rc[kStageU], gd[kStageU];
    Real vf[kStageU], vg[kStageU];
    int ev[kStageU];
    auto item_load = [&](uint32_t i00) {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        const uint32_t ia = i0 + (i < n ? i : 0u);
        cf[u] = B.it_cf[ia];
        cg[u] = B.it_cg[ia];
        sg[u] = B.it_seg[ia];
        rc[u] = B.it_rec[ia];
        gd[u] = Scaled ? B.it_gid[ia] : 0u;
      }
    };
    auto item_gather = [&]() {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        vf[u] = ell[cf[u] >= 0 ? cf[u] : 0];
        vg[u] = ell[cg[u] >= 0 ? cg[u] : 0];
        ev[u] = Scaled ? E[gd[u]] : 0;
      }
    };
    auto item_put = [&](uint32_t i00) {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        if (i < n) {
          sSegL[i] = (uint16_t)(sg[u] - S.seg0);
          bF[i] = cf[u] >= 0 ? vf[u] : Real(0);
          bB[i] = cg[u] >= 0 ? vg[u] : Real(0);  // child1 tops (bB is free until row 1)
          sAcc[i] = 0;
          sRO[i] = rc[u] - rb;
          if (Scaled) {
            sEi[i] = ev[u];
            sFlag[i] = ev[u] == kUndef ? (kPending | kYoung) : 0;
          }
        }
      }
    };
    if (nrec > 0) rec_load(0);
    item_load(0);
    item_gather();
    if (nrec > 0) rec_put(0);
    item_put(0);
    for (uint32_t i00 = kStageU * kWG; i00 < n; i00 += kStageU * kWG) {
      item_load(i00);
      item_gather();
      item_put(i00);
    }
    for (uint32_t k00 = kStageU * kWG; k00 < nrec; k00 += kStageU * kWG) {
      rec_load(k00);
      rec_put(k00);
    }
    if (lane == 0) sRO[n] = nrec;
    wsync<Gws>();
    // Row 0 of a stage without WGD segments: the split terms are summed
    // record-parallel first (into sAcc, which is zero until the row loop),
    // so the item pass below only adds the direct terms.  A clade still
    // without an exponent ("open") takes the largest exponent of its terms:
    // the first record pass raises sEi[own] to kOpen + that exponent, above
    // every real exponent; the item pass normalises and fixes it.
    constexpr int kOpen = 1 << 28;
    const bool fast0 = ST.inreg && !__any(lane < S.n_segs && ST.mine.kind == WHALE_NODE_WGD);
    if (fast0) {
      if (Scaled) {
        for (uint32_t k = lane; k < nrec; k += kWG) {
          const uint32_t own = sOwn[k];
          const int Eo = sEi[own];
          if (Eo != kUndef && Eo < kOpen / 2) continue;
          const uint32_t ops = sOps[k];
          const uint32_t a = ops & 0xffff, b = ops >> 16;
          const Real pk = sCoef[k];
          const Real fa = bF[a], fb = bF[b], ga = bB[a], gb = bB[b];
          const int ea = sEi[a], eb = sEi[b];
          int emax = kUndef;
          Real m;
          int ee, em;
          if (fa != Real(0) && gb != Real(0)) { xprod(pk, fa, ea, gb, eb, &m, &ee); (void)frexp_r(m, &em); emax = max(emax, ee + em); }
          if (ga != Real(0) && fb != Real(0)) { xprod(pk, ga, ea, fb, eb, &m, &ee); (void)frexp_r(m, &em); emax = max(emax, ee + em); }
          if (emax != kUndef) atomicMax(&sEi[own], emax + kOpen);
        }
        wsync<Gws>();
      }
      for (uint32_t base = 0; base < nrec; base += kWG) {  // uniform: the run sums use DPP
        const uint32_t k = base + lane;
        uint32_t own = kNone;
        Real t = 0;
        if (k < nrec) {
          own = sOwn[k];
          const uint32_t ops = sOps[k];
          const uint32_t a = ops & 0xffff, b = ops >> 16;
          const Real pk = sCoef[k];
          const Real fa = bF[a], fb = bF[b], ga = bB[a], gb = bB[b];
          if (!Scaled) {
            t = pk * (fa * gb + ga * fb);
          } else {
            const int Eo = sEi[own], ea = sEi[a], eb = sEi[b];
            if (Eo >= kOpen / 2) {
              const int K = Eo - kOpen;
              Real m;
              int ee;
              if (fa != Real(0) && gb != Real(0)) { xprod(pk, fa, ea, gb, eb, &m, &ee); t += ldexp_r(m, ee - K); }
              if (ga != Real(0) && fb != Real(0)) { xprod(pk, ga, ea, fb, eb, &m, &ee); t += ldexp_r(m, ee - K); }
            } else if (Eo != kUndef && ea != kUndef && eb != kUndef) {
              // fixed exponent: the term formed in FP64 (wide enough for any
              // product of the stored values), rescaled to it
              t = (Real)ldexp((double)pk * ((double)fa * gb + (double)ga * fb), ea + eb - Eo);
            }
          }
        }
        bool last;
        t = row_seg_sum(t, own, &last);
        if (last && own != kNone && t != Real(0)) lds_add(&sAcc[own], t);
      }
      wsync<Gws>();
    }
    // row 0: terms from the child branches (whale!/whalewgd!, src/core.jl:95-128)
    auto row0 = [&](const SegInfo& sg, uint32_t i) {
      const uint32_t r0 = sRO[i], r1 = sRO[i + 1];
      Real v = 0;
      if (fast0 && sg.kind != WHALE_NODE_LEAF) {
        // speciation: the split terms are in sAcc
        const Real sp = sAcc[i];
        sAcc[i] = 0;
        const Real ef = NC.etop[sg.c0], eg = NC.etop[sg.c1];
        if (!Scaled) {
          v = sp + (bF[i] * eg + bB[i] * ef);
        } else {
          const int Eg = sEi[i];
          if (Eg >= kOpen / 2) {
            int e2;
            v = ldexp_r(frexp_r(sp, &e2), -kBirth);
            sEi[i] = Eg - kOpen + e2 + kBirth;
            sFlag[i] = kYoung;
          } else if (Eg != kUndef) {
            v = (bF[i] * eg + bB[i] * ef) + sp;
          }
        }
      } else if (sg.kind == WHALE_NODE_LEAF) {
        if (r1 == r0) {  // a clade of one gene (every larger clade has a split)
          const Real lp = NC.leafP[sg.node];
          if (Scaled) {
            const int Eg = sEi[i];
            if (Eg == kUndef) {
              int ex;
              v = frexp_r(lp, &ex);
              sEi[i] = ex;
              sFlag[i] = kYoung;
            } else {
              v = ldexp_r(lp, -Eg);
            }
          } else {
            v = lp;
          }
        }
      } else if (!Scaled) {
        if (sg.kind == WHALE_NODE_WGD) {
          const Real qq = NC.q[sg.wgd];
          const Real ef = NC.etop[sg.c0];
          Real r = 0;
          for (uint32_t k = r0; k < r1; ++k) {
            const uint32_t ops = sOps[k];
            r += sCoef[k] * bF[ops & 0xffff] * bF[ops >> 16];
          }
          v = r * qq + ((Real(1) - qq) * bF[i] + (Real(2) * qq) * ef * bF[i]);
        } else {
          const Real ef = NC.etop[sg.c0], eg = NC.etop[sg.c1];
          Real sp = 0;
          for (uint32_t k = r0; k < r1; ++k) {
            const uint32_t ops = sOps[k];
            const uint32_t a = ops & 0xffff, b = ops >> 16;
            sp += sCoef[k] * (bF[a] * bB[b] + bB[a] * bF[b]);
          }
          v = sp + (bF[i] * eg + bB[i] * ef);
        }
      } else {
        // scaled: explicit exponents (the operands come from other segments)
        const bool wgd = sg.kind == WHALE_NODE_WGD;
        const Real qq = wgd ? NC.q[sg.wgd] : Real(1);
        const Real ef = NC.etop[sg.c0];
        const Real eg = wgd ? Real(0) : NC.etop[sg.c1];
        const int Eg = sEi[i];
        if (Eg != kUndef) {
          // exponent fixed: the split terms, rescaled to it, summed in FP64
          // (wide enough for any product of the stored values)
          double sacc = 0.0;
          for (uint32_t k = r0; k < r1; ++k) {
            const uint32_t ops = sOps[k];
            const uint32_t a = ops & 0xffff, b = ops >> 16;
            const int ea = sEi[a], eb = sEi[b];
            if (ea == kUndef || eb == kUndef) continue;  // parts without values yet
            const double fa = bF[a], fb = bF[b];
            const double t = wgd ? (double)qq * fa * fb : fa * (double)bB[b] + (double)bB[a] * fb;
            sacc += ldexp((double)sCoef[k] * t, ea + eb - Eg);
          }
          const Real base = wgd ? ((Real(1) - qq) * bF[i] + (Real(2) * qq) * ef * bF[i]) : (bF[i] * eg + bB[i] * ef);
          v = base + (Real)sacc;
        } else {
          // first values of this clade: explicit exponents, then its own
          XSum<Real> xs;
          for (uint32_t k = r0; k < r1; ++k) {
            const uint32_t ops = sOps[k];
            const uint32_t a = ops & 0xffff, b = ops >> 16;
            const Real pk = sCoef[k];
            const Real fa = bF[a], fb = bF[b];
            const Real ga = wgd ? fa : bB[a], gb = wgd ? fb : bB[b];
            const int ea = sEi[a], eb = sEi[b];
            Real m;
            int ex;
            if (fa != Real(0) && gb != Real(0)) { xprod(pk * qq, fa, ea, gb, eb, &m, &ex); xs.add(m, ex); }
            if (!wgd && ga != Real(0) && fb != Real(0)) { xprod(pk, ga, ea, fb, eb, &m, &ex); xs.add(m, ex); }
          }
          if (xs.k != kUndef) {
            int ex;
            v = ldexp_r(frexp_r(xs.s, &ex), -kBirth);
            sEi[i] = xs.k + ex + kBirth;
            sFlag[i] = kYoung;
          }
        }
      }
      bA[i] = v;
      if (StoreAll || sg.R == 1) ell[sg.ell + (i - sg.item0)] = v;
    };
    // row 0 of every item; the first kCache per lane also fill the register cache
    ItemRow cache[kCache];
#pragma unroll
    for (int c = 0; c < kCache; ++c) {
      if ((uint32_t)c * kWG < n) {  // uniform: segment lookups are lane shuffles
        const uint32_t i = lane + c * kWG;
        const SegInfo sg = ST.get(B, M, i < n ? sSegL[i] : 0u);
        if (i < n) {
          cache[c] = item_row(sg, i);
          row0(sg, i);
        }
      }
    }
    for (uint32_t base = kCache * kWG; base < n; base += kWG) {
      const uint32_t i = base + lane;
      const SegInfo sg = ST.get(B, M, i < n ? sSegL[i] : 0u);
      if (i < n) row0(sg, i);
    }
    wsync<Gws>();
    // scaled forms: coefficients for the exponents now fixed.  A record whose
    // clade is still open keeps p (the open clade sums explicit exponents).
    auto refresh = [&]() {
      for (uint32_t k = lane; k < nrec; k += kWG) {
        const uint32_t ops = sOps[k];
        const int Eg = sEi[sOwn[k]];
        const Real p = sP[k];
        const Real c = coef_of(p, sEi[ops & 0xffff], sEi[ops >> 16], Eg);
        sCoef[k] = Eg == kUndef ? p : c;
        if (StoreAll) coef[rb + k - coef_base] = c;
      }
    };
    // items still without values (open clades with splits): they take the
    // explicit-exponent path until their first value
    auto count_open = [&]() -> uint32_t {
      uint32_t cnt = 0;
      for (uint32_t base = 0; base < n; base += kWG) {
        const uint32_t i = base + lane;
        const bool open = i < n && (sFlag[i] & kPending) && sRO[i + 1] > sRO[i];
        cnt += (uint32_t)__popcll(__ballot(open));
      }
      return cnt;
    };
    uint32_t n_open = 0;
    if (Scaled) {
      refresh();
      wsync<Gws>();
      n_open = count_open();
    }
    // the first RC records per lane in registers: operands, owner, coefficient
    // (row-invariant until the coefficients are refreshed)
    uint32_t rops[RC], rown[RC];
    Real rcoef[RC];
    RunMask rm[RC];
    auto load_recs = [&]() {
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        const uint32_t k = lane + c * kWG;
        const bool in = k < nrec;
        rops[c] = in ? sOps[k] : 0u;
        rown[c] = in ? (uint32_t)sOwn[k] : kNone;
        rcoef[c] = in ? sCoef[k] : Real(0);
        rm[c] = run_mask(rown[c]);
      }
    };
    load_recs();
    // rows 1..R-1: within_branch! (src/core.jl:83-92)
    Real* cur = bA;
    Real* nxt = bB;
    for (uint32_t r = 1; r < S.r_max; ++r) {
      bool dirty = false;
      // split sums of every item, one record per lane (lane-balanced however
      // the records are spread over the items); the cached records' operand
      // reads are all issued before any is used
      {
        Real va[RC], vb[RC];
#pragma unroll
        for (int c = 0; c < RC; ++c) {
          if ((uint32_t)c * kWG < nrec) {
            va[c] = cur[rops[c] & 0xffff];
            vb[c] = cur[rops[c] >> 16];
          }
        }
#pragma unroll
        for (int c = 0; c < RC; ++c) {
          if ((uint32_t)c * kWG < nrec) {  // uniform: the row sums use DPP
            const Real t = run_sum(rcoef[c] * va[c] * vb[c], rm[c]);
            if (rm[c].ends() && t != Real(0)) lds_add(&sAcc[rown[c]], t);
          }
        }
      }
      for (uint32_t base = RC * kWG; base < nrec; base += kWG) {  // uniform: the row sums use DPP
        const uint32_t k = base + lane;
        uint32_t own = kNone;
        Real t = 0;
        if (k < nrec) {
          const uint32_t ops = sOps[k];
          own = sOwn[k];
          t = sCoef[k] * cur[ops & 0xffff] * cur[ops >> 16];
        }
        bool last;
        t = row_seg_sum(t, own, &last);
        if (last && own != kNone && t != Real(0)) lds_add(&sAcc[own], t);
      }
      wsync<Gws>();
      // a young clade past 2^kBump is renormalised in place (its only values are
      // this branch's rows); its exponent is raised after the row, as other items
      // of this row still read the clade's previous row at the old scale
      auto bump_check = [&](const ItemRow& it, uint32_t i, Real& v, uint8_t fl) {
        if (Scaled && sizeof(Real) == sizeof(float) && (fl & kYoung) &&
            fabsf((float)v) > (float)(1ull << kBump) && isfinite((float)v)) {
          int ex;
          (void)frexp_r(v, &ex);
          const int sh = ex + kBirth;
          v = ldexp_r(v, -sh);
          if (StoreAll)
            for (uint32_t q = 0; q < r; ++q) {
              const gptr<Real> p = ell + it.cell + (size_t)q * it.stride;
              *p = ldexp_r(*p, -sh);
            }
          bF[i] = (Real)sh;  // bF is free after row 0
          sFlag[i] = fl | kShifted;
          dirty = true;
        }
      };
      auto update = [&](const ItemRow& it, uint32_t i) {
        Real v = cur[i];
        if (r < it.R) {
          Real ph, ps;
          slice<Real, TabL>(M, sTab, it.tab + r, ph, ps);
          const uint8_t fl = Scaled ? sFlag[i] : 0;
          const uint32_t r0 = sRO[i], r1 = sRO[i + 1];
          if (!Scaled || !(fl & kPending) || r1 == r0) {
            v = ph * v + ps * sAcc[i];
            bump_check(it, i, v, fl);
          } else {
            // clade still open: sum with explicit exponents (coefficients hold p)
            XSum<Real> xs;
            for (uint32_t k = r0; k < r1; ++k) {
              const uint32_t ops = sOps[k];
              const uint32_t a = ops & 0xffff, b = ops >> 16;
              const Real va = cur[a], vb = cur[b];
              if (va == Real(0) || vb == Real(0)) continue;
              Real m;
              int ex;
              xprod(sCoef[k], va, sEi[a], vb, sEi[b], &m, &ex);
              xs.add(m, ex);
            }
            if (xs.k != kUndef) {
              int ex;
              v = ldexp_r(frexp_r(ps * xs.s, &ex), -kBirth);
              sEi[i] = xs.k + ex + kBirth;  // readers of row r-1 see 0 here and skip it
              sFlag[i] = kYoung;
              dirty = true;
            }
          }
          if (StoreAll || r == it.R - 1) ell[it.cell + (size_t)r * it.stride] = v;
        }
        nxt[i] = v;
        sAcc[i] = 0;
      };
      if (!Scaled || n_open == 0) {
        // no clade of the stage is still open: every cached item takes the
        // plain update; all slots' LDS reads are issued before any is used
        Real v0[kCache], ac[kCache], ph[kCache], ps[kCache];
        uint8_t fl[kCache];
#pragma unroll
        for (int c = 0; c < kCache; ++c) {
          if ((uint32_t)c * kWG < n) {
            const uint32_t i = lane + c * kWG;
            const uint32_t ii = i < n ? i : 0u;
            const bool act = i < n && r < cache[c].R;
            slice<Real, TabL>(M, sTab, act ? cache[c].tab + r : 0u, ph[c], ps[c]);
            v0[c] = cur[ii];
            ac[c] = sAcc[ii];
            fl[c] = (Scaled && sizeof(Real) == sizeof(float)) ? sFlag[ii] : 0;
          }
        }
#pragma unroll
        for (int c = 0; c < kCache; ++c) {
          if ((uint32_t)c * kWG < n) {
            const uint32_t i = lane + c * kWG;
            if (i < n) {
              const ItemRow& it = cache[c];
              Real v = v0[c];
              if (r < it.R) {
                v = ph[c] * v + ps[c] * ac[c];
                bump_check(it, i, v, fl[c]);
                if (StoreAll || r == it.R - 1) ell[it.cell + (size_t)r * it.stride] = v;
              }
              nxt[i] = v;
              sAcc[i] = 0;
            }
          }
        }
      } else {
#pragma unroll
        for (int c = 0; c < kCache; ++c) {
          const uint32_t i = lane + c * kWG;
          if (i < n) update(cache[c], i);
        }
      }
      for (uint32_t base = kCache * kWG; base < n; base += kWG) {  // uniform: lane shuffles
        const uint32_t i = base + lane;
        const ItemRow it = ST.row(B, M, i < n ? sSegL[i] : 0u, i);
        if (i < n) update(it, i);
      }
      wsync<Gws>();
      if (Scaled && __any(dirty)) {
        for (uint32_t i = lane; i < n; i += kWG) {
          if (sFlag[i] & kShifted) {
            sEi[i] += (int)bF[i];
            sFlag[i] &= (uint8_t)~kShifted;
          }
        }
        wsync<Gws>();
        refresh();
        wsync<Gws>();
        n_open = count_open();
        load_recs();
      }
      Real* t = cur;
      cur = nxt;
      nxt = t;
    }
    // exponents of this stage's clades, for the stages above and ℓ download
    if (Scaled)
      for (uint32_t i = lane; i < n; i += kWG) E[B.it_gid[i0 + i]] = sEi[i];
  }
  // ---------------- root: all clades, by size level (whaleroot!, src/core.jl:139-158)
  {
    wave_sync_mem();
    const uint32_t n = F.n_clades;
    const uint32_t i0 = F.root_item0;
    const SegDev sg = uniform(B.seg[F.root_seg]);
    const NodeDev& nd = M.node[M.root];
    const double eta = M.par[3 * M.n + M.n_wgd];
    const double er = eps_top(M, M.root);
    const double xi = 1.0 - (1.0 - eta) * er;
    const Real cA = (Real)((1.0 - eta) * xi / eta);
    const Real cB = (Real)(eta * (1.0 - er) / (xi * xi));
    const Real ef = NC.etop[nd.child0], eg = NC.etop[nd.child1];
    const uint32_t rb = B.it_rec[i0], nrec = B.it_rec[i0 + n] - rb;
    // records staged in LDS when they fit the wave's region
    const bool inl = fwd_root_bytes(sizeof(Real), Scaled, n, nrec) <= L.wave_bytes;
    const FwdRoot<Real> T(wb, n, inl ? nrec : 0);
    Real* const bA = T.bA;
    Real* const bF = T.bF;
    Real* const bB = T.bB;
    Real* const sCoef = T.sCoef;
    uint32_t* const sOps = T.sOps;
    uint16_t* const sOwn = T.sOwn;
    uint32_t* const sRO = T.sRO;
    int32_t* const sEi = T.sEi;
    for (uint32_t i00 = 0; i00 < n; i00 += kStageU * kWG) {
      int32_t cf[kStageU], cg[kStageU];
      uint32_t rc[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        const uint32_t ia = i0 + (i < n ? i : 0u);
        cf[u] = B.it_cf[ia];
        cg[u] = B.it_cg[ia];
        rc[u] = B.it_rec[ia];
      }
      Real vf[kStageU], vg[kStageU];
      int ev[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        vf[u] = ell[cf[u] >= 0 ? cf[u] : 0];
        vg[u] = ell[cg[u] >= 0 ? cg[u] : 0];
        ev[u] = Scaled ? E[i < n ? i : 0u] : 0;  // root items are the clades in order
      }
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        if (i < n) {
          bF[i] = cf[u] >= 0 ? vf[u] : Real(0);
          bB[i] = cg[u] >= 0 ? vg[u] : Real(0);
          sRO[i] = rc[u] - rb;
          if (Scaled) sEi[i] = ev[u];
        }
      }
    }
    if (lane == 0) sRO[n] = nrec;
    if (inl) {
      for (uint32_t k00 = 0; k00 < nrec; k00 += kStageU * kWG) {
        uint32_t op[kStageU];
        uint16_t ow[kStageU];
        double pp[kStageU];
#pragma unroll
        for (int u = 0; u < kStageU; ++u) {
          const uint32_t k = k00 + u * kWG + lane;
          const uint32_t ka = rb + (k < nrec ? k : 0u);
          op[u] = B.rec_ops[ka];
          ow[u] = B.rec_own[ka];
          pp[u] = B.rec_p[ka];
        }
#pragma unroll
        for (int u = 0; u < kStageU; ++u) {
          const uint32_t k = k00 + u * kWG + lane;
          if (k < nrec) {
            sOps[k] = op[u];
            sOwn[k] = ow[u];
            sCoef[k] = (Real)pp[u];
          }
        }
      }
    }
    const LevelBounds LB(B, F, lane);
    // bA accumulates each clade's root value: it starts at the direct term
    // cB·(F(γ)ε_g + G(γ)ε_f) (zero for a clade without a value below the
    // root), then the level passes add the split terms of its records.
    for (uint32_t i = lane; i < n; i += kWG) bA[i] = cB * (bF[i] * eg + bB[i] * ef);
    wsync<Gws>();
    // record k of a level: operands, p, owner
    auto rrec = [&](uint32_t k, uint32_t& ops, Real& p, uint32_t& own) {
      if (inl) { ops = sOps[k]; p = sCoef[k]; own = sOwn[k]; }
      else { ops = B.rec_ops[rb + k]; p = (Real)B.rec_p[rb + k]; own = B.rec_own[rb + k]; }
    };
    // Scaled forms: a clade still without an exponent at its level ("open")
    // takes the largest exponent of its split terms.  The first record pass
    // raises sEi[own] to kOpen + that exponent (above every real exponent,
    // so the open owners stay recognisable); the second adds each term
    // rescaled to its owner's exponent; the item pass normalises the open
    // clades' sums and fixes their exponents.
    constexpr int kOpen = 1 << 28;
    for (uint32_t Lv = 0; Lv < F.n_lvl; ++Lv) {
      const uint32_t a0 = LB.at(B, F, Lv), a1 = LB.at(B, F, Lv + 1);
      const uint32_t k0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)sRO[a0]);
      const uint32_t k1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)sRO[a1]);
      if (k0 == k1) continue;  // clades without splits keep their direct term
      if (Scaled) {
        for (uint32_t k = k0 + lane; k < k1; k += kWG) {
          uint32_t ops, own;
          Real pk;
          rrec(k, ops, pk, own);
          const int Eo = sEi[own];
          if (Eo != kUndef && Eo < kOpen / 2) continue;  // exponent fixed below the root
          const uint32_t x = ops & 0xffff, y = ops >> 16;
          const int ex_ = sEi[x], ey = sEi[y];
          int emax = kUndef;
          Real m;
          int ee;
          if (bA[x] != Real(0) && bA[y] != Real(0)) { xprod(pk * cA, bA[x], ex_, bA[y], ey, &m, &ee); int em; (void)frexp_r(m, &em); emax = max(emax, ee + em); }
          if (bF[x] != Real(0) && bB[y] != Real(0)) { xprod(pk * cB, bF[x], ex_, bB[y], ey, &m, &ee); int em; (void)frexp_r(m, &em); emax = max(emax, ee + em); }
          if (bB[x] != Real(0) && bF[y] != Real(0)) { xprod(pk * cB, bB[x], ex_, bF[y], ey, &m, &ee); int em; (void)frexp_r(m, &em); emax = max(emax, ee + em); }
          if (emax != kUndef) atomicMax(&sEi[own], emax + kOpen);
        }
        wsync<Gws>();
      }
      for (uint32_t base = k0; base < k1; base += kWG) {  // uniform: the run sums use DPP
        const uint32_t k = base + lane;
        uint32_t own = kNone;
        Real t = 0;
        if (k < k1) {
          uint32_t ops;
          Real pk;
          rrec(k, ops, pk, own);
          const uint32_t x = ops & 0xffff, y = ops >> 16;
          if (!Scaled) {
            t = pk * (cA * bA[x] * bA[y] + cB * (bF[x] * bB[y] + bB[x] * bF[y]));
          } else {
            const int Eo = sEi[own], ex_ = sEi[x], ey = sEi[y];
            if (Eo >= kOpen / 2) {
              // open owner: each product as (mantissa, exponent) against its largest
              const int K = Eo - kOpen;
              Real m;
              int ee;
              if (bA[x] != Real(0) && bA[y] != Real(0)) { xprod(pk * cA, bA[x], ex_, bA[y], ey, &m, &ee); t += ldexp_r(m, ee - K); }
              if (bF[x] != Real(0) && bB[y] != Real(0)) { xprod(pk * cB, bF[x], ex_, bB[y], ey, &m, &ee); t += ldexp_r(m, ee - K); }
              if (bB[x] != Real(0) && bF[y] != Real(0)) { xprod(pk * cB, bB[x], ex_, bF[y], ey, &m, &ee); t += ldexp_r(m, ee - K); }
            } else if (Eo != kUndef && ex_ != kUndef && ey != kUndef) {
              // fixed exponent: the term rescaled to it, formed in FP64 (wide
              // enough for any product of the stored values)
              const double tt = (double)cA * bA[x] * bA[y] + (double)cB * ((double)bF[x] * bB[y] + (double)bB[x] * bF[y]);
              t = (Real)ldexp((double)pk * tt, ex_ + ey - Eo);
            }
          }
        }
        bool last;
        t = row_seg_sum(t, own, &last);
        if (last && own != kNone && t != Real(0)) lds_add(&bA[own], t);
      }
      wsync<Gws>();
      if (Scaled) {
        for (uint32_t i = a0 + lane; i < a1; i += kWG) {
          const int Eo = sEi[i];
          if (Eo >= kOpen / 2) {
            int e2;
            const Real mm = frexp_r(bA[i], &e2);
            bA[i] = ldexp_r(mm, -kBirth);
            sEi[i] = Eo - kOpen + e2 + kBirth;
          }
        }
        wsync<Gws>();
      }
    }
    for (uint32_t i = lane; i < n; i += kWG) {
      ell[sg.ell + i] = bA[i];
      if (Scaled) E[i] = sEi[i];
    }
    if (lane == 0) {
      const Real Lr = bA[n - 1];
      double ll;
      if (Lr > Real(0) && isfinite((double)Lr)) {
        ll = log((double)Lr);
        if (Scaled) ll += (double)sEi[n - 1] * 0.69314718055994530942;
      } else {
        ll = -INFINITY;
        if (Scaled && sizeof(Real) == sizeof(float)) {
          // overflow, or everything underflowed: redo this family in FP64 (a true zero stays -Inf there)
          W.flag[f] = 1u;
          W.fb[1 + atomicAdd(W.fb, 1u)] = f;
        }
      }
      W.fam_ll[f] = ll;
    }
    if (Scaled && StoreAll) {
      // final coefficients of the root records, for the reverse sweep
      for (uint32_t k = lane; k < nrec; k += kWG) {
        const uint32_t ops = inl ? sOps[k] : B.rec_ops[rb + k];
        const Real p = inl ? sCoef[k] : (Real)B.rec_p[rb + k];
        const int Eg = sEi[inl ? sOwn[k] : B.rec_own[rb + k]];
        coef[rb + k - coef_base] = coef_of(p, sEi[ops & 0xffff], sEi[ops >> 16], Eg);
      }
    }
  }
}


template <typename Real, bool Scaled, bool StoreAll, bool TabL, bool Gws>
__global__ __launch_bounds__(kWG* kMaxFPW, kWavesPerEU) void k_forward(BatchDev B, ModelDev M, WorkDev<Real> W, LdsCfg L,
                                                       int32_t fam_single, uint32_t fam_begin, uint32_t fam_end,
                                                       const uint32_t* fam_count, const MixComp* X, int K) {
  if (K > 0) mix_select(X[blockIdx.x % (uint32_t)K], B, M, W, fam_count);
  if (fam_count) fam_end = fam_begin + *fam_count;
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t lane = threadIdx.x % kWG;
  const uint32_t wv = threadIdx.x / kWG;
  Real* sTab = reinterpret_cast<Real*>(smem);
  NodeC<Real> NC(smem + (TabL ? align16(2 * (size_t)M.tab_rows * sizeof(Real)) : 0), M.n);
  if (TabL) {
    for (int r = threadIdx.x; r < M.tab_rows; r += (int)blockDim.x) {
      sTab[2 * r] = (Real)M.phi[r];
      sTab[2 * r + 1] = (Real)M.psi[r];
    }
  }
  NC.fill(M, threadIdx.x, (int)blockDim.x);
  __syncthreads();
  // the wave's stage buffers: its share of the dynamic LDS, or (Gws) its
  // region of the global workspace (one wave per workgroup)
  unsigned char* const wb = Gws ? W.gws + ((size_t)blockIdx.x * (blockDim.x / kWG) + wv) * L.wave_bytes
                                : smem + L.shared_bytes + (size_t)wv * L.wave_bytes;
  bool first = true;
  for (;;) {
    uint32_t f;
    if (fam_single >= 0) {
      if (!first || wv != 0) break;
      f = (uint32_t)fam_single;
    } else {
      uint32_t k = 0;
      if (lane == 0) k = atomicAdd(W.ticket, 1u);
      k = fam_begin + __builtin_amdgcn_readfirstlane(k);
      if (k >= fam_end) break;
      f = __builtin_amdgcn_readfirstlane(B.order[k]);  // wave-uniform: family data in scalar registers
    }
    first = false;
    fwd_family<Real, Scaled, StoreAll, TabL, Gws>(B, M, W, L, f, wb, sTab, NC, lane, W.ell + B.fam[f].ell0,
                                                  W.coef, 0u);
  }
}

// ---------------------------------------------------------------------------
// backward (reverse sweep); requires a forward with StoreAll
// ---------------------------------------------------------------------------
// LDS of one wave for a reverse-sweep stage (bwd_stage_bytes) and for the
// root (bwd_root_bytes).  Shared by the workgroup: φ, ψ per table row (Real)
// and the gradient accumulators in the global layout (doubles, flushed to
// global memory once per workgroup) when TabL, then the node constants.
template <typename Real>
struct BwdStage {
  Real *bA, *bB, *bV, *bG, *sFo, *sCoef;
  uint32_t *sOps, *sRO;
  uint16_t *sOwn, *sSegL;
  __device__ BwdStage(unsigned char* p, size_t n, size_t nr) {
    const size_t s = sizeof(Real);
    bA = reinterpret_cast<Real*>(p);
    bB = bA + n;
    bV = bB + n;
    bG = bV + n;
    sFo = bG + n;
    p += align16(5 * n * s);
    sCoef = reinterpret_cast<Real*>(p);
    p += align16(nr * s);
    sOps = reinterpret_cast<uint32_t*>(p);
    p += align16(4 * nr);
    sOwn = reinterpret_cast<uint16_t*>(p);
    p += align16(2 * nr);
    sRO = reinterpret_cast<uint32_t*>(p);
    p += align16(4 * (n + 1));
    sSegL = reinterpret_cast<uint16_t*>(p);
  }
};

template <typename Real>
struct BwdRoot {
  Real *bA, *bB, *bV, *bG, *sCoef;
  uint32_t *sOps, *sRO;
  uint16_t* sOwn;
  __device__ BwdRoot(unsigned char* p, size_t n, size_t nr) {
    const size_t s = sizeof(Real);
    bA = reinterpret_cast<Real*>(p);
    bB = bA + n;
    bV = bB + n;
    bG = bV + n;
    p += align16(4 * n * s);
    sCoef = reinterpret_cast<Real*>(p);
    p += align16(nr * s);
    sOps = reinterpret_cast<uint32_t*>(p);
    p += align16(4 * nr);
    sOwn = reinterpret_cast<uint16_t*>(p);
    p += align16(2 * nr);
    sRO = reinterpret_cast<uint32_t*>(p);
  }
};

// per-item data of the row-0 step of the reverse sweep
struct ItemTop {
  uint32_t ib;      // ī cell (relative to ibar0) of this clade in the child0 slot; child1: + nc
  uint32_t nc;
  int32_t kind;
  int32_t c0, c1;   // children
  int32_t wgd;      // q index of a WGD node
};

__device__ __forceinline__ ItemTop item_top(const SegInfo& sg, uint32_t i) {
  ItemTop t;
  t.ib = sg.ibar == kNone ? kNone : sg.ibar + (i - sg.item0);
  t.nc = sg.nc;
  t.kind = sg.kind;
  t.c0 = sg.c0;
  t.c1 = sg.c1;
  t.wgd = sg.wgd;
  return t;
}

// The reverse sweep of family f by one wave, from its stored ℓ (`ell`,
// relative cells) and record coefficients (`coefs`, absolute records; the
// plain form takes p), writing ī through `ibar` (relative cells) and adding
// into the accumulators accPS (LDS, with TabL) and accX.
template <typename Real, bool Scaled, bool TabL, bool Gws>
__device__ __forceinline__ void bwd_family(const BatchDev& B, const ModelDev& M, const WorkDev<Real>& W,
                                           const LdsCfg& L, uint32_t f, unsigned char* const wb, Real* sTab,
                                           Real* const accPS, double* const accX, const NodeC<Real>& NC,
                                           uint32_t lane, const Real* ell_arg, Real* ibar_arg,
                                           const Real* coefs_arg, uint32_t coef_base) {
  const gptr<const Real> ell = glob(ell_arg);
  const gptr<Real> ibar = glob(ibar_arg);
  const gptr<const Real> coefs = glob(coefs_arg);
  constexpr int RC = 2;  // records per lane held in registers (128 per stage)
  const AccLayout AL(M.tab_rows, M.n, M.n_wgd);
  // Deterministic gradients: this family's accumulators go to its own slot
  // with plain stores (one wave per workgroup, tables and accumulators in LDS,
  // each table row flushed once per family); a fixed-order sum over the
  // families follows.  Otherwise they go to the batch sums with atomics.
  double* const accF = TabL && W.acc_fam ? W.acc_fam + (size_t)f * AL.n : nullptr;
  if (accF) {
    for (int j = (int)lane; j < AL.n; j += kWG) accF[j] = 0.0;
    wave_sync_mem();
  }
  const FamDev F = uniform(B.fam[f]);
  auto coef_at = [&](uint32_t k) -> Real { return Scaled ? coefs[k - coef_base] : (Real)B.rec_p[k]; };
  // ---------------- root
  {
    const uint32_t n = F.n_clades;
    const uint32_t i0 = F.root_item0;
    const SegDev sg = uniform(B.seg[F.root_seg]);
    const NodeDev& nd = M.node[M.root];
    const double eta = M.par[3 * M.n + M.n_wgd];
    const double er = eps_top(M, M.root);
    const double xi = 1.0 - (1.0 - eta) * er;
    const Real cA = (Real)((1.0 - eta) * xi / eta);
    const Real cB = (Real)(eta * (1.0 - er) / (xi * xi));
    const Real ef = NC.etop[nd.child0], eg = NC.etop[nd.child1];
    const uint32_t rb = B.it_rec[i0], nrec = B.it_rec[i0 + n] - rb;
    const bool inl = bwd_root_bytes(sizeof(Real), n, nrec) <= L.wave_bytes;
    const BwdRoot<Real> T(wb, n, inl ? nrec : 0);
    Real* const bA = T.bA;
    Real* const bB = T.bB;
    Real* const bV = T.bV;
    Real* const bG = T.bG;
    Real* const sCoef = T.sCoef;
    uint32_t* const sOps = T.sOps;
    uint16_t* const sOwn = T.sOwn;
    uint32_t* const sRO = T.sRO;
    const LevelBounds LB(B, F, lane);
    for (uint32_t i00 = 0; i00 < n; i00 += kStageU * kWG) {
      int32_t cf[kStageU], cg[kStageU];
      uint32_t rc[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        const uint32_t ia = i0 + (i < n ? i : 0u);
        cf[u] = B.it_cf[ia];
        cg[u] = B.it_cg[ia];
        rc[u] = B.it_rec[ia];
      }
      Real vv[kStageU], vf[kStageU], vg[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        vv[u] = ell[sg.ell + (i < n ? i : 0u)];
        vf[u] = ell[cf[u] >= 0 ? cf[u] : 0];
        vg[u] = ell[cg[u] >= 0 ? cg[u] : 0];
      }
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        if (i < n) {
          bV[i] = vv[u];
          bB[i] = cf[u] >= 0 ? vf[u] : Real(0);  // child0 tops
          bG[i] = cg[u] >= 0 ? vg[u] : Real(0);  // child1 tops
          bA[i] = 0;
          sRO[i] = rc[u] - rb;
        }
      }
    }
    if (lane == 0) sRO[n] = nrec;
    if (inl)
      for (uint32_t k00 = 0; k00 < nrec; k00 += kStageU * kWG) {
        uint32_t op[kStageU];
        uint16_t ow[kStageU];
        Real cc[kStageU];
#pragma unroll
        for (int u = 0; u < kStageU; ++u) {
          const uint32_t k = k00 + u * kWG + lane;
          const uint32_t ka = rb + (k < nrec ? k : 0u);
          op[u] = B.rec_ops[ka];
          ow[u] = B.rec_own[ka];
          cc[u] = Scaled ? coefs[ka - coef_base] : (Real)B.rec_p[ka];
        }
#pragma unroll
        for (int u = 0; u < kStageU; ++u) {
          const uint32_t k = k00 + u * kWG + lane;
          if (k < nrec) {
            sOps[k] = op[u];
            sOwn[k] = ow[u];
            sCoef[k] = cc[u];
          }
        }
      }
    wsync<Gws>();
    // record k: operands, coefficient, owning clade
    auto rec = [&](uint32_t k, uint32_t& ops, Real& c, uint32_t& own) {
      if (inl) { ops = sOps[k]; c = sCoef[k]; own = sOwn[k]; }
      else { ops = B.rec_ops[rb + k]; c = coef_at(rb + k); own = B.rec_own[rb + k]; }
    };
    const Real Lr = bV[n - 1];
    const bool ok = Lr > Real(0) && isfinite((double)Lr);
    if (ok) {
      if (lane == 0) bA[n - 1] = (W.seed ? (Real)W.seed[f] : Real(1)) / Lr;
      wsync<Gws>();
      // d logL / d ℓ of each root clade, larger clades first, scattered into
      // the parts; a level's records (those of its clades) one per lane
      for (int Lv = (int)F.n_lvl - 1; Lv >= 0; --Lv) {
        const uint32_t a0 = LB.at(B, F, (uint32_t)Lv), a1 = LB.at(B, F, (uint32_t)Lv + 1);
        const uint32_t k1 = sRO[a1];
        for (uint32_t k = sRO[a0] + lane; k < k1; k += kWG) {
          uint32_t ops, own;
          Real c;
          rec(k, ops, c, own);
          const Real t = cA * bA[own] * c;
          if (t == Real(0)) continue;
          const uint32_t x = ops & 0xffff, y = ops >> 16;
          lds_add(&bA[x], t * bV[y]);
          lds_add(&bA[y], t * bV[x]);
        }
        wsync<Gws>();
      }
      Real sA = 0, sBc = 0, sF = 0, sG = 0;
      for (uint32_t k = lane; k < nrec; k += kWG) {
        uint32_t ops, own;
        Real c;
        rec(k, ops, c, own);
        const Real g = bA[own];
        const uint32_t x = ops & 0xffff, y = ops >> 16;
        sA += g * (c * bV[x] * bV[y]);
        sBc += g * (c * (bB[x] * bG[y] + bG[x] * bB[y]));
      }
      for (uint32_t i = lane; i < n; i += kWG) {
        const Real g = bA[i];
        sBc += g * (bB[i] * eg + bG[i] * ef);
        sF += g * bB[i];
        sG += g * bG[i];
      }
      sA = wave_sum(sA);
      sBc = wave_sum(sBc);
      sF = wave_sum(sF);
      sG = wave_sum(sG);
      if (lane == 0) {
        acc_add<TabL>(accX + AL.rootA, (double)sA);
        acc_add<TabL>(accX + AL.rootB, (double)sBc);
        acc_add<TabL>(accX + AL.eps + nd.child1, (double)cB * (double)sF);
        acc_add<TabL>(accX + AL.eps + nd.child0, (double)cB * (double)sG);
      }
      // ī of the child tops: child0 slot takes ℓ of child1 tops and vice versa
      wsync<Gws>();
      for (int slot = 0; slot < 2; ++slot) {
        const Real* other = slot == 0 ? bG : bB;
        const Real eo = slot == 0 ? eg : ef;
        for (uint32_t i = lane; i < n; i += kWG) bV[i] = bA[i] * eo;
        wsync<Gws>();
        for (uint32_t k = lane; k < nrec; k += kWG) {
          uint32_t ops, own;
          Real c;
          rec(k, ops, c, own);
          const Real g = bA[own] * c;
          if (g == Real(0)) continue;
          const uint32_t x = ops & 0xffff, y = ops >> 16;
          lds_add(&bV[x], g * other[y]);
          lds_add(&bV[y], g * other[x]);
        }
        wsync<Gws>();
        for (uint32_t i = lane; i < n; i += kWG) ibar[sg.ibar + slot * n + i] = cB * bV[i];
        wsync<Gws>();
      }
    } else {
      for (uint32_t i = lane; i < 2 * n; i += kWG) ibar[sg.ibar + i] = 0;
    }
  }
  // ---------------- non-root stages, top down
  for (int s = (int)F.n_stages - 1; s >= 0; --s) {
    const StageDev S = uniform(B.stage[F.stage0 + s]);
    const uint32_t n = S.n_items, i0 = S.item0;
    const uint32_t rb = S.rec0, nrec = S.n_recs;
    const BwdStage<Real> T(wb, n, nrec);  // fits the wave's LDS (size class)
    Real* bA = T.bA;  // d logL / d ℓ, current row
    Real* bB = T.bB;  // next row (and ī sums at row 0)
    Real* const bV = T.bV;  // ℓ of the row below / child0 tops
    Real* const bG = T.bG;  // child1 tops; ψ·ā per item in the row loop
    Real* const sFo = T.sFo;  // per-item sums and factors of the row-0 step
    Real* const sCoef = T.sCoef;
    uint32_t* const sOps = T.sOps;
    uint16_t* const sOwn = T.sOwn;
    uint32_t* const sRO = T.sRO;
    uint16_t* const sSegL = T.sSegL;
    wave_sync_mem();  // ī written by the stage above
    SegTab ST;
    ST.load(B, M, S.seg0, S.n_segs, lane);
    // kStageU items (records) per lane at once: their index loads, then the
    // loads of ī they name, then the LDS writes; the first batches of
    // records and items go together (two dependent round trips to memory)
    uint32_t op[kStageU];
    uint16_t ow[kStageU];
    Real cc[kStageU];
    auto rec_load = [&](uint32_t k00) {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t k = k00 + u * kWG + lane;
        const uint32_t ka = rb + (k < nrec ? k : 0u);
        op[u] = B.rec_ops[ka];
        ow[u] = B.rec_own[ka];
        cc[u] = Scaled ? coefs[ka - coef_base] : (Real)B.rec_p[ka];
      }
    };
    auto rec_put = [&](uint32_t k00) {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t k = k00 + u * kWG + lane;
        if (k < nrec) {
          sOps[k] = op[u];
          sOwn[k] = ow[u];
          sCoef[k] = cc[u];
        }
      }
    };
    int32_t ib[kStageU];
    uint32_t rc[kStageU], sg[kStageU];
    Real vb[kStageU];
    auto item_load = [&](uint32_t i00) {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        const uint32_t ia = i0 + (i < n ? i : 0u);
        ib[u] = B.it_bar[ia];
        rc[u] = B.it_rec[ia];
        sg[u] = B.it_seg[ia];
      }
    };
    auto item_gather = [&]() {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) vb[u] = ibar[ib[u]];
    };
    auto item_put = [&](uint32_t i00) {
#pragma unroll
      for (int u = 0; u < kStageU; ++u) {
        const uint32_t i = i00 + u * kWG + lane;
        if (i < n) {
          bA[i] = vb[u];
          bV[i] = 0;
          sRO[i] = rc[u] - rb;
          sSegL[i] = (uint16_t)(sg[u] - S.seg0);
        }
      }
    };
    if (nrec > 0) rec_load(0);
    item_load(0);
    item_gather();
    if (nrec > 0) rec_put(0);
    item_put(0);
    for (uint32_t i00 = kStageU * kWG; i00 < n; i00 += kStageU * kWG) {
      item_load(i00);
      item_gather();
      item_put(i00);
    }
    for (uint32_t k00 = kStageU * kWG; k00 < nrec; k00 += kStageU * kWG) {
      rec_load(k00);
      rec_put(k00);
    }
    if (lane == 0) sRO[n] = nrec;
    wsync<Gws>();
    ItemRow cache[kCacheB];
    ItemTop tops[kCacheB];
    // ℓ of the rows below, loaded two rows ahead into two register sets
    // that take turns (rows r and r-1 use preA and preB)
    Real preA[kCacheB], preB[kCacheB];
    const int rtop = (int)S.r_max - 1;
#pragma unroll
    for (int c = 0; c < kCacheB; ++c) {
      preA[c] = 0;
      preB[c] = 0;
      if ((uint32_t)c * kWG < n) {  // uniform: segment lookups are lane shuffles
        const uint32_t i = lane + c * kWG;
        const SegInfo sg = ST.get(B, M, i < n ? sSegL[i] : 0u);
        if (i < n) {
          cache[c] = item_row(sg, i);
          tops[c] = item_top(sg, i);
          const ItemRow& it = cache[c];
          if (rtop >= 1 && rtop < (int)it.R) preA[c] = ell[it.cell + (size_t)(rtop - 1) * it.stride];
          if (rtop >= 2 && rtop - 1 < (int)it.R) preB[c] = ell[it.cell + (size_t)(rtop - 2) * it.stride];
        }
      }
    }
    // runs of items sharing a segment (table row), for the d/dφ, d/dψ sums
    RunMask im[kCacheB];
#pragma unroll
    for (int c = 0; c < kCacheB; ++c) {
      if ((uint32_t)c * kWG < n) {
        const uint32_t i = lane + c * kWG;
        im[c] = run_mask(i < n ? cache[c].tab : kNone);
      }
    }
    Real* const sT = bG;  // ψ_r·ā(r) per item, for the record pass
    // the first RC records per lane in registers: operands, owner,
    // coefficient, and the table row of the owner's segment with the runs
    // of records sharing it (the records of a segment are consecutive)
    uint32_t rops[RC], rown[RC], rtab[RC];
    Real rcoef[RC];
    RunMask rs[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      const uint32_t k = lane + c * kWG;
      const bool in = k < nrec;
      rops[c] = in ? sOps[k] : 0u;
      rown[c] = in ? (uint32_t)sOwn[k] : kNone;
      rcoef[c] = in ? sCoef[k] : Real(0);
      if ((uint32_t)c * kWG < nrec) {  // uniform: lane shuffles
        const uint32_t t = ST.tab(B, M, in ? (uint32_t)sSegL[rown[c]] : 0u);
        rtab[c] = in ? t : kNone;
        rs[c] = run_mask(rtab[c]);
      }
    }
    // d/dψ_r: Σ over a segment's records of ā(r)[owner]·c·ℓ(γ1)ℓ(γ2), one
    // atomic per run of records (records past their segment's rows have
    // ψ_r·ā = 0 and add nothing)
    auto add_psi = [&](uint32_t key, int r, Real x) {
      if (TabL) lds_add(&accPS[AL.psi + key + r], x);
      else acc_add<false>(W.acc + AL.psi + key + r, (double)x);
    };
    // Per row r (top down) two item passes and one record pass:
    //   A(r)    items: ℓ of row r-1 (prefetched), φ_r·ā(r) into the other
    //           buffer (the start of ā(r-1)) and ψ_r·ā(r) for the records
    //   REC(r)  records: the ψ part scattered into the two parts' ā(r-1),
    //           and the owner's split sum (forcing of d/dψ_r)
    //   C(r)    items: d/dφ_r, d/dψ_r per segment
    // C(r) and A(r-1) share one pass over the items.
    // A(x) for item i from ā(x) = g into dst; pv = ℓ(x-1)
    // (branch-free: past its segment's rows an item passes ā through and
    // its records' forcing is zero; pv is then a finite stale value)
    auto item_a = [&](const int x, const ItemRow& it, uint32_t i, Real g, Real pv, Real* dst) {
      const bool in = x < (int)it.R;
      Real ph, ps;
      slice<Real, TabL>(M, sTab, in ? it.tab + x : 0u, ph, ps);
      bV[i] = pv;
      dst[i] = in ? ph * g : g;
      sT[i] = in ? ps * g : Real(0);
    };
    auto refill = [&](const int x, Real (&pre)[kCacheB]) {  // ℓ(x-3) for A(x-2)
#pragma unroll
      for (int c = 0; c < kCacheB; ++c) {
        const uint32_t i = lane + c * kWG;
        if (i < n && x >= 3 && x - 2 < (int)cache[c].R) pre[c] = ell[cache[c].cell + (size_t)(x - 3) * cache[c].stride];
      }
    };
    // C(r) for item i: d/dφ_r from ā(r) = g and ℓ(r-1) = v (all lanes take part)
    auto item_c = [&](const int r, const ItemRow& it, bool on, Real g, Real v, const RunMask& rmk) {
      Real cphi = on ? g * v : Real(0);
      cphi = run_sum(cphi, rmk);
      if (rmk.ends() && on && cphi != Real(0)) {
        const uint32_t key = it.tab;
        if (TabL) lds_add(&accPS[AL.phi + key + r], cphi);
        else acc_add<false>(W.acc + AL.phi + key + r, (double)cphi);
      }
    };
    // A(rtop) on its own
    {
#pragma unroll
      for (int c = 0; c < kCacheB; ++c) {
        const uint32_t i = lane + c * kWG;
        if (i < n) item_a(rtop, cache[c], i, bA[i], preA[c], bB);
      }
      for (uint32_t base = kCacheB * kWG; base < n; base += kWG) {  // uniform: lane shuffles
        const uint32_t i = base + lane;
        const ItemRow it = ST.row(B, M, i < n ? sSegL[i] : 0u, i);
        if (i < n)
          item_a(rtop, it, i, bA[i], rtop < (int)it.R ? ell[it.cell + (size_t)(rtop - 1) * it.stride] : Real(0), bB);
      }
      refill(rtop, preA);
    }
    auto rec_pass = [&](const int r) {
      // records, one per lane; the cached records' reads are all issued
      // before any is used
      {
        Real va[RC], vb[RC], g[RC], ga[RC];
#pragma unroll
        for (int c = 0; c < RC; ++c) {
          if ((uint32_t)c * kWG < nrec) {
            const uint32_t o = rown[c] != kNone ? rown[c] : 0u;
            va[c] = bV[rops[c] & 0xffff];
            vb[c] = bV[rops[c] >> 16];
            g[c] = rown[c] != kNone ? sT[o] : Real(0);
            ga[c] = bA[o];
          }
        }
#pragma unroll
        for (int c = 0; c < RC; ++c) {
          if ((uint32_t)c * kWG < nrec) {  // uniform: the run sums use DPP
            const Real t = g[c] * rcoef[c];
            if (t != Real(0)) {
              lds_add(&bB[rops[c] & 0xffff], t * vb[c]);
              lds_add(&bB[rops[c] >> 16], t * va[c]);
            }
            const Real x = run_sum(t != Real(0) ? ga[c] * rcoef[c] * va[c] * vb[c] : Real(0), rs[c]);
            if (rs[c].ends() && x != Real(0)) add_psi(rtab[c], r, x);
          }
        }
      }
      for (uint32_t base = RC * kWG; base < nrec; base += kWG) {  // uniform: shuffles, row sums
        const uint32_t k = base + lane;
        const bool in = k < nrec;
        const uint32_t own = in ? (uint32_t)sOwn[k] : 0u;
        const uint32_t tb = ST.tab(B, M, sSegL[own]);
        const uint32_t key = in ? tb : kNone;
        Real x = 0;
        if (in) {
          const uint32_t ops = sOps[k];
          const uint32_t a = ops & 0xffff, b = ops >> 16;
          const Real c = sCoef[k];
          const Real va = bV[a], vb = bV[b];
          const Real t = sT[own] * c;
          if (t != Real(0)) {
            lds_add(&bB[a], t * vb);
            lds_add(&bB[b], t * va);
            x = bA[own] * c * va * vb;
          }
        }
        bool last;
        x = row_seg_sum(x, key, &last);
        if (last && key != kNone && x != Real(0)) add_psi(key, r, x);
      }
    };
    // REC(r), then C(r) with A(r-1); A(r-1) reads ℓ(r-2) from `pre`
    auto row_step = [&](const int r, Real (&pre)[kCacheB]) {
      wsync<Gws>();
      rec_pass(r);
      wsync<Gws>();
      const bool more = r >= 2;  // uniform
#pragma unroll
      for (int c = 0; c < kCacheB; ++c) {
        if ((uint32_t)c * kWG < n) {  // uniform: the run sums use DPP
          const uint32_t i = lane + c * kWG;
          const uint32_t ii = i < n ? i : 0u;
          const Real g = bA[ii], v = bV[ii], g1 = bB[ii];
          item_c(r, cache[c], i < n && r < (int)cache[c].R, g, v, im[c]);
          if (more && i < n) item_a(r - 1, cache[c], i, g1, pre[c], bA);
        }
      }
      for (uint32_t base = kCacheB * kWG; base < n; base += kWG) {  // uniform: lane shuffles, DPP
        const uint32_t i = base + lane;
        const uint32_t ii = i < n ? i : 0u;
        const ItemRow it = ST.row(B, M, i < n ? sSegL[i] : 0u, i);
        const Real g = bA[ii], v = bV[ii], g1 = bB[ii];
        item_c(r, it, i < n && r < (int)it.R, g, v, run_mask(i < n ? it.tab : kNone));
        if (more && i < n)
          item_a(r - 1, it, i, g1, r - 1 < (int)it.R ? ell[it.cell + (size_t)(r - 2) * it.stride] : Real(0), bA);
      }
      if (more) refill(r - 1, pre);
      Real* t = bA;
      bA = bB;
      bB = t;
    };
    for (int r = rtop; r >= 1; r -= 2) {
      row_step(r, preB);
      if (r >= 2) row_step(r - 1, preA);
    }
    wsync<Gws>();
    if (TabL) {
      // move this stage's d/dφ_r, d/dψ_r into the global FP64 sums (each LDS
      // partial then holds at most one stage's contributions per family)
      for (uint32_t sl = 0; sl < S.n_segs; ++sl) {
        uint32_t tab, R;
        if (ST.inreg) {
          tab = (uint32_t)__builtin_amdgcn_readlane((int)ST.mine.tab, (int)sl);
          R = (uint32_t)__builtin_amdgcn_readlane((int)ST.mine.R, (int)sl);
        } else {
          const SegInfo g = seg_info(B, M, S.seg0 + sl);
          tab = g.tab;
          R = g.R;
        }
        for (uint32_t r = 1 + lane; r < R; r += kWG) {
          const Real a = __hip_atomic_exchange(&accPS[AL.phi + tab + r], Real(0), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
          const Real b = __hip_atomic_exchange(&accPS[AL.psi + tab + r], Real(0), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
          if (accF) {
            accF[AL.phi + tab + r] = (double)a;
            accF[AL.psi + tab + r] = (double)b;
          } else {
            if (a != Real(0)) unsafeAtomicAdd(W.acc + AL.phi + tab + r, (double)a);
            if (b != Real(0)) unsafeAtomicAdd(W.acc + AL.psi + tab + r, (double)b);
          }
        }
      }
    }
    // row 0: adjoints of the child tops, ε and q
    // pass over all items: cached tops for the first kCacheB per lane
    auto each = [&](auto&& fn) {
#pragma unroll
      for (int c = 0; c < kCacheB; ++c) {
        const uint32_t i = lane + c * kWG;
        if (i < n) fn(tops[c], i);
      }
      for (uint32_t base = kCacheB * kWG; base < n; base += kWG) {
        const uint32_t i = base + lane;
        const SegInfo sg = ST.get(B, M, i < n ? sSegL[i] : 0u);
        if (i < n) fn(item_top(sg, i), i);
      }
    };
    // pass over all records: the first RC per lane from registers
    auto each_rec = [&](auto&& fn) {
#pragma unroll
      for (int c = 0; c < RC; ++c)
        if ((uint32_t)c * kWG < nrec && rown[c] != kNone) fn(rops[c], rown[c], rcoef[c]);
      for (uint32_t k = RC * kWG + lane; k < nrec; k += kWG) fn(sOps[k], (uint32_t)sOwn[k], sCoef[k]);
    };
    // child tops; a WGD item's one child also stands in the child1 slot, so
    // the split terms below read the "other" top from bG for both kinds
    bool any_wgd = false;
    each([&](const ItemTop& tp, uint32_t i) {
      const int32_t cf = B.it_cf[i0 + i], cg = B.it_cg[i0 + i];
      const Real v0 = cf >= 0 ? ell[cf] : Real(0);
      bV[i] = v0;
      bG[i] = tp.kind == WHALE_NODE_WGD ? v0 : (cg >= 0 ? ell[cg] : Real(0));
      sFo[i] = 0;
      any_wgd |= tp.kind == WHALE_NODE_WGD;
    });
    wsync<Gws>();
    if (__any(any_wgd)) {
      // WGD retention sums Σ c·ℓ(γ1)ℓ(γ2) over the child's tops, per owner
      each_rec([&](uint32_t ops, uint32_t own, Real c) {
        const Real t = c * bV[ops & 0xffff] * bV[ops >> 16];
        if (t != Real(0)) lds_add(&sFo[own], t);
      });
      wsync<Gws>();
    }
    // ε̄, q̄ and the direct terms of the child0 slot; then each owner's
    // factor of its split terms in that slot
    each([&](const ItemTop& tp, uint32_t i) {
      const Real g = bA[i];
      Real fac = 0;
      if (tp.kind == WHALE_NODE_SPECIATION) {
        const Real eg = NC.etop[tp.c1];
        acc_add<TabL>(accX + AL.eps + tp.c1, (double)(g * bV[i]));
        acc_add<TabL>(accX + AL.eps + tp.c0, (double)(g * bG[i]));
        bB[i] = g * eg;
        fac = g;
      } else if (tp.kind == WHALE_NODE_WGD) {
        const Real qq = NC.q[tp.wgd];
        const Real ef = NC.etop[tp.c0];
        const Real ret = sFo[i];
        acc_add<TabL>(accX + AL.q + tp.wgd, (double)(g * (ret - bV[i] + Real(2) * ef * bV[i])));
        acc_add<TabL>(accX + AL.eps + tp.c0, (double)(g * Real(2) * qq * bV[i]));
        bB[i] = g * ((Real(1) - qq) + Real(2) * qq * ef);
        fac = qq * g;
      }
      sFo[i] = fac;
    });
    wsync<Gws>();
    // split terms of the child0 slot: speciation pairs child0 parts with child1
    // tops of the other part; a WGD pairs both parts on its one child
    each_rec([&](uint32_t ops, uint32_t own, Real c) {
      const Real g = sFo[own] * c;
      if (g == Real(0)) return;
      const uint32_t a = ops & 0xffff, b = ops >> 16;
      lds_add(&bB[a], g * bG[b]);
      lds_add(&bB[b], g * bG[a]);
    });
    wsync<Gws>();
    each([&](const ItemTop& tp, uint32_t i) {
      if (tp.kind == WHALE_NODE_SPECIATION || tp.kind == WHALE_NODE_WGD) ibar[tp.ib] = bB[i];
      const bool sp = tp.kind == WHALE_NODE_SPECIATION;
      if (sp) bB[i] = bA[i] * NC.etop[tp.c0];
      sFo[i] = sp ? bA[i] : Real(0);
    });
    wsync<Gws>();
    // child1 slot of speciation nodes
    each_rec([&](uint32_t ops, uint32_t own, Real c) {
      const Real g = sFo[own] * c;
      if (g == Real(0)) return;
      const uint32_t a = ops & 0xffff, b = ops >> 16;
      lds_add(&bB[a], g * bV[b]);
      lds_add(&bB[b], g * bV[a]);
    });
    wsync<Gws>();
    each([&](const ItemTop& tp, uint32_t i) {
      if (tp.kind == WHALE_NODE_SPECIATION) ibar[tp.ib + tp.nc] = bB[i];
    });
  }
  if (accF) {
    // ε, q and the root constants of this family, from the workgroup's LDS
    wsync<Gws>();
    for (int j = AL.eps + (int)lane; j < AL.n; j += kWG) {
      accF[j] = accX[j];
      accX[j] = 0.0;
    }
    wsync<Gws>();
  }
}

template <typename Real, bool Scaled, bool TabL, bool Gws>
__global__ __launch_bounds__(kWG* kMaxFPW, kWavesPerEU) void k_backward(BatchDev B, ModelDev M, WorkDev<Real> W, LdsCfg L,
                                                        int32_t fam_single, uint32_t fam_begin, uint32_t fam_end,
                                                        const uint32_t* fam_count, const MixComp* X, int K) {
  constexpr int RC = 2;  // records per lane held in registers (128 per stage)
  if (K > 0) mix_select(X[blockIdx.x % (uint32_t)K], B, M, W, fam_count);
  if (fam_count) fam_end = fam_begin + *fam_count;
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t lane = threadIdx.x % kWG;
  const uint32_t wv = threadIdx.x / kWG;
  const AccLayout AL(M.tab_rows, M.n, M.n_wgd);
  const int T2 = 2 * M.tab_rows;
  Real* sTab = reinterpret_cast<Real*>(smem);
  const size_t tab_b = align16((size_t)T2 * sizeof(Real));
  // accumulators: with TabL, d/dφ_r and d/dψ_r in LDS at the working precision
  // (flushed into the global FP64 sums at the end of every stage), the rest
  // (ε, q, A, B) in LDS in FP64; otherwise everything goes to global memory
  Real* const accPS = reinterpret_cast<Real*>(smem + tab_b);
  double* const accX = TabL ? reinterpret_cast<double*>(smem + 2 * tab_b) - AL.eps : W.acc;
  NodeC<Real> NC(smem + (TabL ? 2 * tab_b + align16(8 * (size_t)(AL.n - AL.eps)) : 0), M.n);
  if (TabL) {
    for (int r = threadIdx.x; r < M.tab_rows; r += (int)blockDim.x) {
      sTab[2 * r] = (Real)M.phi[r];
      sTab[2 * r + 1] = (Real)M.psi[r];
    }
    for (int j = threadIdx.x; j < T2; j += (int)blockDim.x) accPS[j] = 0;
    for (int j = AL.eps + threadIdx.x; j < AL.n; j += (int)blockDim.x) accX[j] = 0.0;
  }
  NC.fill(M, threadIdx.x, (int)blockDim.x);
  __syncthreads();
  // the wave's stage buffers: its share of the dynamic LDS, or (Gws) its
  // region of the global workspace (one wave per workgroup)
  unsigned char* const wb = Gws ? W.gws + ((size_t)blockIdx.x * (blockDim.x / kWG) + wv) * L.wave_bytes
                                : smem + L.shared_bytes + (size_t)wv * L.wave_bytes;
  const Real* coefs = Scaled ? W.coef : nullptr;
  bool first = true;
  for (;;) {
    uint32_t f;
    if (fam_single >= 0) {
      if (!first || wv != 0) break;
      f = (uint32_t)fam_single;
    } else {
      uint32_t k = 0;
      if (lane == 0) k = atomicAdd(W.ticket, 1u);
      k = fam_begin + __builtin_amdgcn_readfirstlane(k);
      if (k >= fam_end) break;
      f = __builtin_amdgcn_readfirstlane(B.order[k]);  // wave-uniform: family data in scalar registers
    }
    first = false;
    if (Scaled && sizeof(Real) == sizeof(float) && W.flag[f]) continue;  // done by the FP64 pass
    bwd_family<Real, Scaled, TabL, Gws>(B, M, W, L, f, wb, sTab, accPS, accX, NC, lane, W.ell + B.fam[f].ell0,
                                        W.ibar + B.fam[f].ibar0, coefs, 0u);
  }
  if (TabL && !W.acc_fam) {
    __syncthreads();
    for (int j = threadIdx.x; j < T2; j += (int)blockDim.x) {
      const Real v = accPS[j];
      if (v != Real(0)) unsafeAtomicAdd(W.acc + j, (double)v);
    }
    for (int j = AL.eps + threadIdx.x; j < AL.n; j += (int)blockDim.x) {
      const double v = accX[j];
      if (v != 0.0) unsafeAtomicAdd(W.acc + j, v);
    }
  }
}

// ---------------------------------------------------------------------------
// tail: one workgroup.  Batch sum, conditioning, table adjoint, chain rule.
// out: [total, dλ(n), dμ(n), dq(n_wgd), dη]
// ---------------------------------------------------------------------------
// k_fused: the forward DP and (grad) the reverse sweep of each family by one
// wave, with the family's stored rows, ī and record coefficients in the
// wave's own slot of a scratch buffer.  A family's rows live only while its
// wave works on it, so ℓ takes (resident waves) x (the launch's largest
// family) instead of every cell of the batch: the evaluation under a memory
// budget for ℓ.  Shared LDS as in the reverse sweep; a wave's stage region
// fits both directions.
// ---------------------------------------------------------------------------
template <typename Real, bool Scaled, bool TabL, bool Gws>
__global__ __launch_bounds__(kWG* kMaxFPW, kWavesPerEU) void k_fused(BatchDev B, ModelDev M, WorkDev<Real> W, LdsCfg L,
                                                     uint32_t fam_begin, uint32_t fam_end, const uint32_t* fam_count,
                                                     int grad, FusedSlots SL) {
  if (fam_count) fam_end = fam_begin + *fam_count;
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t lane = threadIdx.x % kWG;
  const uint32_t wv = threadIdx.x / kWG;
  const AccLayout AL(M.tab_rows, M.n, M.n_wgd);
  const int T2 = 2 * M.tab_rows;
  Real* sTab = reinterpret_cast<Real*>(smem);
  const size_t tab_b = align16((size_t)T2 * sizeof(Real));
  Real* const accPS = reinterpret_cast<Real*>(smem + tab_b);
  double* const accX = TabL ? reinterpret_cast<double*>(smem + 2 * tab_b) - AL.eps : W.acc;
  NodeC<Real> NC(smem + (TabL ? 2 * tab_b + align16(8 * (size_t)(AL.n - AL.eps)) : 0), M.n);
  if (TabL) {
    for (int r = threadIdx.x; r < M.tab_rows; r += (int)blockDim.x) {
      sTab[2 * r] = (Real)M.phi[r];
      sTab[2 * r + 1] = (Real)M.psi[r];
    }
    for (int j = threadIdx.x; j < T2; j += (int)blockDim.x) accPS[j] = 0;
    for (int j = AL.eps + threadIdx.x; j < AL.n; j += (int)blockDim.x) accX[j] = 0.0;
  }
  NC.fill(M, threadIdx.x, (int)blockDim.x);
  __syncthreads();
  const size_t slot = (size_t)blockIdx.x * (blockDim.x / kWG) + wv;
  unsigned char* const wb = Gws ? W.gws + slot * L.wave_bytes : smem + L.shared_bytes + (size_t)wv * L.wave_bytes;
  unsigned char* const sp = SL.base + slot * SL.slot_bytes;
  Real* const ell = reinterpret_cast<Real*>(sp);
  Real* const ibar = reinterpret_cast<Real*>(sp + SL.ibar_off);
  Real* const coef = reinterpret_cast<Real*>(sp + SL.coef_off);
  for (;;) {
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(W.ticket, 1u);
    k = fam_begin + __builtin_amdgcn_readfirstlane(k);
    if (k >= fam_end) break;
    const uint32_t f = __builtin_amdgcn_readfirstlane(B.order[k]);
    const uint32_t rec0 = B.fam[f].rec0;
    wave_sync_mem();  // the slot's previous family is done with it
    fwd_family<Real, Scaled, true, TabL, Gws>(B, M, W, L, f, wb, sTab, NC, lane, ell, coef, rec0);
    if (!grad) continue;
    wave_sync_mem();  // the forward's rows, coefficients and flag are visible to the wave
    if (Scaled && sizeof(Real) == sizeof(float) &&
        __hip_atomic_load(&W.flag[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
      continue;  // done by the FP64 pass
    bwd_family<Real, Scaled, TabL, Gws>(B, M, W, L, f, wb, sTab, accPS, accX, NC, lane, ell, ibar, coef, rec0);
  }
  if (TabL && grad && !W.acc_fam) {
    __syncthreads();
    for (int j = threadIdx.x; j < T2; j += (int)blockDim.x) {
      const Real v = accPS[j];
      if (v != Real(0)) unsafeAtomicAdd(W.acc + j, (double)v);
    }
    for (int j = AL.eps + threadIdx.x; j < AL.n; j += (int)blockDim.x) {
      const double v = accX[j];
      if (v != 0.0) unsafeAtomicAdd(W.acc + j, v);
    }
  }
}

// ---------------------------------------------------------------------------
// k_pool: every family's forward and (grad) reverse sweep by one wave, with
// the family's stage buffers carved from a pool of LDS pages that the
// workgroup's waves share.  One workgroup per compute unit holds the unit's
// LDS: the shared area (tables, accumulators, node constants, as k_fused),
// then the pool.  A wave takes the next family, allocates the pages that
// family needs (contiguous, first fit), runs it and frees them, so small
// families run many to a compute unit and large ones few, in one launch over
// the whole batch and with no boundary between the directions.  Rows, ī and
// coefficients go to the batch's full storage as in the stored form.
//
// Allocation: the wave holding the pool lock takes the next ticket and waits
// (lock held) until its pages are free; frees need no lock (atomic and-not on
// the page bitmap), and a wave holding pages never waits for another wave,
// so every allocation of a family no larger than the empty pool completes.
// The waits are bounded all the same: a family that cannot be placed within
// the bound gets a NaN logL instead of a hang.
// ---------------------------------------------------------------------------
constexpr uint32_t kPoolSpin = 1u << 21;  // sleeps before an allocation gives up

struct PoolHdr {
  uint32_t lock;
  uint32_t bits[kPoolHdrBytes / 4 - 1];  // page bitmap, 1 = in use
};

// lane 0: the lock, then the next ticket and its pages.  Returns the ticket
// (≥ n_tickets when the batch is done) and the first page (kNone if the wait
// gave up).
__device__ inline uint32_t pool_take(PoolHdr* H, uint32_t npages, uint32_t* ticket, uint32_t n_tickets,
                                     const BatchDev& B, uint32_t fam_begin, size_t s, bool scaled, uint32_t pool_bytes,
                                     uint32_t* first, uint32_t* pages) {
  uint32_t spin = 0;
  for (;;) {
    uint32_t expect = 0u;
    if (__hip_atomic_compare_exchange_strong(&H->lock, &expect, 1u, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP))
      break;
    __builtin_amdgcn_s_sleep(4);
    if (++spin > kPoolSpin) {  // never expected: the lock holder's wait is bounded too
      *first = kNone;
      *pages = 0;
      return atomicAdd(ticket, 1u);
    }
  }
  const uint32_t k = atomicAdd(ticket, 1u);
  *first = kNone;
  *pages = 0;
  if (k < n_tickets) {
    const FamDev& d = B.fam[B.order[fam_begin + k]];
    const uint32_t need = (uint32_t)pool_need(d, s, scaled, pool_bytes);
    const uint32_t np = (need + kPage - 1) / kPage;
    *pages = np;
    for (spin = 0; spin < kPoolSpin; ++spin) {
      uint32_t run = 0;
      for (uint32_t p = 0; p < npages; ++p) {
        const uint32_t used = (__hip_atomic_load(&H->bits[p >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >>
                               (p & 31)) & 1u;
        run = used ? 0u : run + 1u;
        if (run == np) {
          *first = p + 1 - np;
          break;
        }
      }
      if (*first != kNone) break;
      __builtin_amdgcn_s_sleep(8);
    }
    if (*first != kNone)
      for (uint32_t p = *first; p < *first + np;) {  // mark, a word at a time
        const uint32_t w = p >> 5, lo = p & 31, hi = min(32u, lo + (*first + np - p));
        const uint32_t m = (hi == 32 ? ~0u : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
        __hip_atomic_fetch_or(&H->bits[w], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        p += hi - lo;
      }
  }
  __hip_atomic_store(&H->lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  return k;
}

__device__ inline void pool_free(PoolHdr* H, uint32_t first, uint32_t np) {
  for (uint32_t p = first; p < first + np;) {
    const uint32_t w = p >> 5, lo = p & 31, hi = min(32u, lo + (first + np - p));
    const uint32_t m = (hi == 32 ? ~0u : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
    __hip_atomic_fetch_and(&H->bits[w], ~m, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    p += hi - lo;
  }
}

template <typename Real, bool Scaled, bool StoreAll>
__global__ __launch_bounds__(kWG* kPoolWaves, kWavesPerEU) void k_pool(BatchDev B, ModelDev M, WorkDev<Real> W,
                                                                     LdsCfg L, uint32_t fam_begin, uint32_t fam_end,
                                                                     int grad) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t lane = threadIdx.x % kWG;
  const AccLayout AL(M.tab_rows, M.n, M.n_wgd);
  const int T2 = 2 * M.tab_rows;
  Real* sTab = reinterpret_cast<Real*>(smem);
  const size_t tab_b = align16((size_t)T2 * sizeof(Real));
  Real* const accPS = reinterpret_cast<Real*>(smem + tab_b);
  double* const accX = reinterpret_cast<double*>(smem + 2 * tab_b) - AL.eps;
  NodeC<Real> NC(smem + 2 * tab_b + align16(8 * (size_t)(AL.n - AL.eps)), M.n);
  PoolHdr* const H = reinterpret_cast<PoolHdr*>(smem + L.shared_bytes);
  unsigned char* const pool = smem + L.shared_bytes + kPoolHdrBytes;
  const uint32_t pool_bytes = L.wave_bytes - kPoolHdrBytes;
  const uint32_t npages = pool_bytes / kPage;
  for (int r = threadIdx.x; r < M.tab_rows; r += (int)blockDim.x) {
    sTab[2 * r] = (Real)M.phi[r];
    sTab[2 * r + 1] = (Real)M.psi[r];
  }
  for (int j = threadIdx.x; j < T2; j += (int)blockDim.x) accPS[j] = 0;
  for (int j = AL.eps + threadIdx.x; j < AL.n; j += (int)blockDim.x) accX[j] = 0.0;
  for (int j = threadIdx.x; j < (int)(kPoolHdrBytes / 4); j += (int)blockDim.x) reinterpret_cast<uint32_t*>(H)[j] = 0u;
  NC.fill(M, threadIdx.x, (int)blockDim.x);
  __syncthreads();
  const uint32_t n_tickets = fam_end - fam_begin;
  for (;;) {
    uint32_t k = 0, first = 0, np = 0;
    if (lane == 0)
      k = pool_take(H, npages, W.ticket, n_tickets, B, fam_begin, sizeof(Real), Scaled, pool_bytes, &first, &np);
    k = __builtin_amdgcn_readfirstlane(k);
    if (k >= n_tickets) break;
    first = __builtin_amdgcn_readfirstlane(first);
    np = __builtin_amdgcn_readfirstlane(np);
    const uint32_t f = __builtin_amdgcn_readfirstlane(B.order[fam_begin + k]);
    if (first == kNone) {  // the allocation gave up (not expected): the result says so
      if (lane == 0) W.fam_ll[f] = __builtin_nan("");
      continue;
    }
    const LdsCfg Lf{np * kPage, L.shared_bytes};
    unsigned char* const wb = pool + (size_t)first * kPage;
    fwd_family<Real, Scaled, StoreAll, true, false>(B, M, W, Lf, f, wb, sTab, NC, lane, W.ell + B.fam[f].ell0, W.coef,
                                                    0u);
    wave_sync_mem();  // the forward's rows, coefficients and flag are visible to the wave
    const bool flagged = Scaled && sizeof(Real) == sizeof(float) &&
                         __hip_atomic_load(&W.flag[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
    if (grad && !flagged) {  // else done by the FP64 pass
      bwd_family<Real, Scaled, true, false>(B, M, W, Lf, f, wb, sTab, accPS, accX, NC, lane, W.ell + B.fam[f].ell0,
                                            W.ibar + B.fam[f].ibar0, Scaled ? W.coef : nullptr, 0u);
    }
    wave_sync();  // the wave's last reads and writes of its pages are done
    if (lane == 0) pool_free(H, first, np);
  }
  if (grad && !W.acc_fam) {
    __syncthreads();
    for (int j = threadIdx.x; j < T2; j += (int)blockDim.x) {
      const Real v = accPS[j];
      if (v != Real(0)) unsafeAtomicAdd(W.acc + j, (double)v);
    }
    for (int j = AL.eps + threadIdx.x; j < AL.n; j += (int)blockDim.x) {
      const double v = accX[j];
      if (v != 0.0) unsafeAtomicAdd(W.acc + j, v);
    }
  }
}

// ---------------------------------------------------------------------------
// The conditioning term c = log P(condition) (src/condition.jl:11-29) of the
// current parameters and its derivatives in the root's ε, its children's ε
// and η: RootCondition P = 1 − g(ε_f) − g(ε_g) + g(ε_ρ), NonExtinct
// P = 1 − g(ε_ρ), g the root's geometric pgf; none: c = 0.
struct CondTerms {
  double c, dc_er, dc_ef, dc_eg, dc_eta;
};
__device__ inline CondTerms cond_terms(const ModelDev& M, int cond) {
  const NodeDev& rn = M.node[M.root];
  const double eta = M.par[3 * M.n + M.n_wgd];
  const double er = eps_top(M, M.root);
  const double ef = eps_top(M, rn.child0), eg = eps_top(M, rn.child1);
  auto gpgf = [&](double x) { return eta * x / (1.0 - (1.0 - eta) * x); };
  auto dg_ds = [&](double x) { double D = 1.0 - (1.0 - eta) * x; return eta / (D * D); };
  auto dg_deta = [&](double x) { double D = 1.0 - (1.0 - eta) * x; return x * (1.0 - x) / (D * D); };
  CondTerms T{0, 0, 0, 0, 0};
  if (cond == WHALE_COND_ROOT) {
    const double P = 1.0 - gpgf(ef) - gpgf(eg) + gpgf(er);
    T.c = P > 0 ? log(P) : -INFINITY;
    T.dc_ef = -dg_ds(ef) / P;
    T.dc_eg = -dg_ds(eg) / P;
    T.dc_er = dg_ds(er) / P;
    T.dc_eta = (-dg_deta(ef) - dg_deta(eg) + dg_deta(er)) / P;
  } else if (cond == WHALE_COND_NONEXTINCT) {
    const double P = 1.0 - gpgf(er);
    T.c = log(P);
    T.dc_er = -dg_ds(er) / P;
    T.dc_eta = -dg_deta(er) / P;
  }
  return T;
}

// k_mix: one workgroup.  The mixture's per-family log-sum-exp over its
// components (src/core.jl:66-76) and the responsibilities that weight each
// component's reverse sweep.
__global__ __launch_bounds__(1024) void k_mix(const MixComp* X, int K, const double* log_prior, uint32_t n_fam,
                                              int cond, double* out, double* resp) {
  __shared__ double red[1024];
  __shared__ double s_c[kMaxMix];
  __shared__ double s_w[kMaxMix][32];
  const int t = threadIdx.x;
  if (t < K) s_c[t] = cond_terms(X[t].M, cond).c;
  for (int i = t; i < kMaxMix * 32; i += blockDim.x) s_w[i / 32][i % 32] = 0.0;
  __syncthreads();
  double tot = 0.0, wsum[kMaxMix];
  for (int j = 0; j < kMaxMix; ++j) wsum[j] = 0.0;
  for (uint32_t i = t; i < n_fam; i += blockDim.x) {
    double l[kMaxMix], mx = -INFINITY;
    for (int j = 0; j < K; ++j) {
      l[j] = X[j].W.fam_ll[i] + log_prior[j] - s_c[j];
      mx = fmax(mx, l[j]);
    }
    double lse = -INFINITY;
    if (isfinite(mx)) {
      double acc = 0.0;
      for (int j = 0; j < K; ++j) acc += exp(l[j] - mx);
      lse = mx + log(acc);
    }
    for (int j = 0; j < K; ++j) {
      const double r = isfinite(lse) ? exp(l[j] - lse) : 0.0;
      X[j].W.seed[i] = r;
      wsum[j] += r;
      if (resp) resp[(size_t)i * K + j] = r;
    }
    tot += lse;
  }
  red[t] = tot;
  for (int j = 0; j < K; ++j) atomicAdd(&s_w[j][t % 32], wsum[j]);
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t < K) {
    double sw = 0.0;
    for (int q = 0; q < 32; ++q) sw += s_w[t][q];
    X[t].W.seed[n_fam] = sw;
  }
  if (t == 0) out[0] = isfinite(red[0]) ? red[0] : -INFINITY;
}

constexpr int kTailThreads = 1024;

// k_tail: one workgroup.  Σℓ over the families, the conditioning term, and
// the chain rule from the per-row table adjoints to ∂λ, ∂μ, ∂q, ∂η.  Down a
// branch the table adjoint is affine in the adjoint of its top ε:
//   ε̄(r−1) = ε̄(r)·∂ε/∂ε' + φ̄_r·∂φ/∂ε' + ψ̄_r·∂ψ/∂ε',
// and so are ᾱ and β̄.  Each node's three affine maps are built from its rows
// with one lane per node, all nodes at once; only applying them runs by depth
// (parents first, a node's top adjoint takes its parent's bottom one).
// Per node in cf (6 doubles): ε̄(bottom), ᾱ, β̄ as (slope, offset) pairs in ε̄(top).
__global__ __launch_bounds__(kTailThreads) void k_tail(ModelDev M, const double* fam_ll, uint32_t n_fam,
                                                       const double* acc_in, double* ebar, const double* drow,
                                                       const double* dnode, const int32_t* dlevels,
                                                       const int32_t* dnodes, int n_dlevels, int cond, int with_grad,
                                                       double* out, const double* n_eff) {
  __shared__ double red[kTailThreads];
  __shared__ double s_total, s_eta;
  const int t = threadIdx.x;
  const AccLayout AL(M.tab_rows, M.n, M.n_wgd);
  double* cf = ebar + M.n;
  // the affine maps first: they do not depend on the sum
  if (with_grad)
    for (int e = t; e < M.n; e += blockDim.x) {
      const NodeDev& d = M.node[e];
      double ga = 1.0, gb = 0.0;           // ε̄ = ga·ε̄(top) + gb
      double aa = 0.0, ab = 0.0, ba = 0.0, bb = 0.0;
      for (int r = d.R - 1; r >= 1; --r) {
        const double* G = drow + 9 * (size_t)(d.tab + r);
        const double fp = acc_in[AL.phi + d.tab + r], fs = acc_in[AL.psi + d.tab + r];
        aa += ga * G[0];
        ab += gb * G[0] + fp * G[3] + fs * G[6];
        ba += ga * G[1];
        bb += gb * G[1] + fp * G[4] + fs * G[7];
        gb = gb * G[2] + fp * G[5] + fs * G[8];
        ga = ga * G[2];
      }
      cf[6 * e + 0] = ga;
      cf[6 * e + 1] = gb;
      cf[6 * e + 2] = aa;
      cf[6 * e + 3] = ab;
      cf[6 * e + 4] = ba;
      cf[6 * e + 5] = bb;
    }
  // Σ ℓ_i
  double s = 0;
  for (uint32_t i = t; i < n_fam; i += blockDim.x) s += fam_ll[i];
  red[t] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const NodeDev& rn = M.node[M.root];
  const double eta = M.par[3 * M.n + M.n_wgd];
  const double er = eps_top(M, M.root);
  const CondTerms CT = cond_terms(M, cond);
  const double c = CT.c, dc_er = CT.dc_er, dc_ef = CT.dc_ef, dc_eg = CT.dc_eg, dc_eta = CT.dc_eta;
  const double N = (double)n_fam;
  if (t == 0) {
    const double tot = red[0] - N * c;
    s_total = isfinite(tot) ? tot : -INFINITY;
  }
  if (!with_grad) {
    if (t == 0) out[0] = s_total;
    return;
  }
  // ε̄ of every node's top row: DP contributions + conditioning + root A/B terms
  for (int e = t; e < M.n; e += blockDim.x) ebar[e] = acc_in[AL.eps + e];
  __syncthreads();
  if (t == 0) {
    const double A_ = acc_in[AL.rootA], B_ = acc_in[AL.rootB];
    const double xi = 1.0 - (1.0 - eta) * er;
    const double dA_deta = (-xi + (1.0 - eta) * er) / eta - (1.0 - eta) * xi / (eta * eta);
    const double dA_der = -(1.0 - eta) * (1.0 - eta) / eta;
    const double dB_deta = (1.0 - er) / (xi * xi) - 2.0 * eta * (1.0 - er) * er / (xi * xi * xi);
    const double dB_der = -eta / (xi * xi) + 2.0 * eta * (1.0 - er) * (1.0 - eta) / (xi * xi * xi);
    const double Ng = n_eff ? *n_eff : N;  // the families' weight in the gradient
    ebar[M.root] += A_ * dA_der + B_ * dB_der - Ng * dc_er;
    ebar[rn.child0] += -Ng * dc_ef;
    ebar[rn.child1] += -Ng * dc_eg;
    s_eta = A_ * dA_deta + B_ * dB_deta - Ng * dc_eta;
  }
  __syncthreads();
  double* dl = out + 1;
  double* dm = out + 1 + M.n;
  double* dq = out + 1 + 2 * M.n;
  for (int w = t; w < M.n_wgd; w += blockDim.x) dq[w] = acc_in[AL.q + w];
  __syncthreads();
  // parents before children
  for (int L = 0; L < n_dlevels; ++L) {
    for (int k = dlevels[L] + t; k < dlevels[L + 1]; k += blockDim.x) {
      const int e = dnodes[k];
      const NodeDev& d = M.node[e];
      const double gt = ebar[e];
      const double* c = cf + 6 * e;
      const double g = c[0] * gt + c[1];  // ε̄ of row 0
      const double ga = c[2] * gt + c[3], gb = c[4] * gt + c[5];
      dl[e] = d.R > 1 ? ga * dnode[4 * e + 0] + gb * dnode[4 * e + 2] : 0.0;
      dm[e] = d.R > 1 ? ga * dnode[4 * e + 1] + gb * dnode[4 * e + 3] : 0.0;
      // row 0 of ε: product of the children's tops, or the WGD retention map
      if (d.kind == WHALE_NODE_WGD) {
        const double ec = eps_top(M, d.child0);
        const double qq = M.par[3 * M.n + d.wgd];
        ebar[d.child0] += g * (2.0 * qq * ec + (1.0 - qq));
        dq[d.wgd] += g * (ec * ec - ec);
      } else if (d.kind == WHALE_NODE_SPECIATION || d.kind == WHALE_NODE_ROOT) {
        const double e0 = eps_top(M, d.child0), e1 = eps_top(M, d.child1);
        ebar[d.child0] += g * e1;
        ebar[d.child1] += g * e0;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = s_total;
    out[1 + 2 * M.n + M.n_wgd] = s_eta;
  }
}

// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static size_t real_size(int mode) { return mode == WHALE_FP32_SCALED ? sizeof(float) : sizeof(double); }

size_t fwd_shared_bytes(int mode, bool tabl, int tab_rows, int n, int nwgd) {
  return (tabl ? align16(2 * (size_t)tab_rows * real_size(mode)) : 0) + nodec_bytes(real_size(mode), n, nwgd);
}

size_t bwd_shared_bytes(int mode, bool tabl, int tab_rows, int acc_n, int n, int nwgd) {
  const size_t tab_b = align16(2 * (size_t)tab_rows * real_size(mode));
  return (tabl ? 2 * tab_b + align16(8 * (size_t)(acc_n - 2 * tab_rows)) : 0) + nodec_bytes(real_size(mode), n, nwgd);
}

template <typename K, typename... Args>
static hipError_t go(K kern, int grid, int fpw, size_t lds, hipStream_t st, Args... args) {
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kWG * fpw), lds, st, args...);
  return hipGetLastError();
}

template <typename R, bool S, bool A>
static hipError_t fwd_variant(const LaunchCfg& C, hipStream_t st, const BatchDev& B, const ModelDev& M,
                              const WorkRaw& W, int32_t fam_single, uint32_t fam_begin, uint32_t fam_end,
                              const uint32_t* fam_count, const MixComp* X, int K) {
  const size_t lds = C.lds();
  const WorkDev<R> w(W);
  if (C.fpw < 1 || C.fpw > kMaxFPW || K < 0 || K > kMaxMix || (K > 0 && (!X || fam_single >= 0)))
    return hipErrorInvalidValue;
  const int grid = C.grid * (K > 0 ? K : 1);
  if (C.gws) {  // stage buffers in the global workspace: one wave per workgroup, tables from global memory
    if (C.fpw != 1 || C.tabl || !W.gws || K > 0) return hipErrorInvalidValue;
    return go(k_forward<R, S, A, false, true>, grid, 1, lds, st, B, M, w, C.L, fam_single, fam_begin, fam_end,
              fam_count, X, K);
  }
  if (C.tabl)
    return go(k_forward<R, S, A, true, false>, grid, C.fpw, lds, st, B, M, w, C.L, fam_single, fam_begin, fam_end,
              fam_count, X, K);
  return go(k_forward<R, S, A, false, false>, grid, C.fpw, lds, st, B, M, w, C.L, fam_single, fam_begin, fam_end,
            fam_count, X, K);
}

template <typename R, bool S>
static hipError_t bwd_variant(const LaunchCfg& C, hipStream_t st, const BatchDev& B, const ModelDev& M,
                              const WorkRaw& W, int32_t fam_single, uint32_t fam_begin, uint32_t fam_end,
                              const uint32_t* fam_count, const MixComp* X, int K) {
  const size_t lds = C.lds();
  const WorkDev<R> w(W);
  if (C.fpw < 1 || C.fpw > kMaxFPW || K < 0 || K > kMaxMix || (K > 0 && (!X || fam_single >= 0)))
    return hipErrorInvalidValue;
  const int grid = C.grid * (K > 0 ? K : 1);
  if (C.gws) {
    if (C.fpw != 1 || !W.gws || K > 0) return hipErrorInvalidValue;
    if (C.tabl)
      return go(k_backward<R, S, true, true>, grid, 1, lds, st, B, M, w, C.L, fam_single, fam_begin, fam_end,
                fam_count, X, K);
    return go(k_backward<R, S, false, true>, grid, 1, lds, st, B, M, w, C.L, fam_single, fam_begin, fam_end,
              fam_count, X, K);
  }
  if (C.tabl)
    return go(k_backward<R, S, true, false>, grid, C.fpw, lds, st, B, M, w, C.L, fam_single, fam_begin, fam_end,
              fam_count, X, K);
  return go(k_backward<R, S, false, false>, grid, C.fpw, lds, st, B, M, w, C.L, fam_single, fam_begin, fam_end,
            fam_count, X, K);
}

hipError_t launch_forward(int mode, bool store_all, const LaunchCfg& C, hipStream_t st, const BatchDev& B,
                          const ModelDev& M, const WorkRaw& W, int32_t fam_single, uint32_t fam_begin,
                          uint32_t fam_end, const uint32_t* fam_count, const MixComp* X, int K) {
  if (mode == WHALE_FP64_PLAIN)
    return store_all ? fwd_variant<double, false, true>(C, st, B, M, W, fam_single, fam_begin, fam_end, fam_count, X, K)
                     : fwd_variant<double, false, false>(C, st, B, M, W, fam_single, fam_begin, fam_end, fam_count, X, K);
  if (mode == WHALE_FP32_SCALED)
    return store_all ? fwd_variant<float, true, true>(C, st, B, M, W, fam_single, fam_begin, fam_end, fam_count, X, K)
                     : fwd_variant<float, true, false>(C, st, B, M, W, fam_single, fam_begin, fam_end, fam_count, X, K);
  return store_all ? fwd_variant<double, true, true>(C, st, B, M, W, fam_single, fam_begin, fam_end, fam_count, X, K)
                   : fwd_variant<double, true, false>(C, st, B, M, W, fam_single, fam_begin, fam_end, fam_count, X, K);
}

hipError_t launch_backward(int mode, const LaunchCfg& C, hipStream_t st, const BatchDev& B, const ModelDev& M,
                           const WorkRaw& W, int32_t fam_single, uint32_t fam_begin, uint32_t fam_end,
                           const uint32_t* fam_count, const MixComp* X, int K) {
  if (mode == WHALE_FP64_PLAIN)
    return bwd_variant<double, false>(C, st, B, M, W, fam_single, fam_begin, fam_end, fam_count, X, K);
  if (mode == WHALE_FP32_SCALED)
    return bwd_variant<float, true>(C, st, B, M, W, fam_single, fam_begin, fam_end, fam_count, X, K);
  return bwd_variant<double, true>(C, st, B, M, W, fam_single, fam_begin, fam_end, fam_count, X, K);
}

template <typename R, bool S>
static hipError_t fused_variant(const LaunchCfg& C, hipStream_t st, const BatchDev& B, const ModelDev& M,
                                const WorkRaw& W, uint32_t fam_begin, uint32_t fam_end, const uint32_t* fam_count,
                                int grad, const FusedSlots& SL) {
  const size_t lds = C.lds();
  const WorkDev<R> w(W);
  if (C.fpw < 1 || C.fpw > kMaxFPW || !SL.base) return hipErrorInvalidValue;
  if (C.gws) {
    if (C.fpw != 1 || C.tabl || !W.gws) return hipErrorInvalidValue;
    return go(k_fused<R, S, false, true>, C.grid, 1, lds, st, B, M, w, C.L, fam_begin, fam_end, fam_count, grad, SL);
  }
  if (C.tabl)
    return go(k_fused<R, S, true, false>, C.grid, C.fpw, lds, st, B, M, w, C.L, fam_begin, fam_end, fam_count, grad,
              SL);
  return go(k_fused<R, S, false, false>, C.grid, C.fpw, lds, st, B, M, w, C.L, fam_begin, fam_end, fam_count, grad,
            SL);
}

hipError_t launch_fused(int mode, const LaunchCfg& C, hipStream_t st, const BatchDev& B, const ModelDev& M,
                        const WorkRaw& W, uint32_t fam_begin, uint32_t fam_end, const uint32_t* fam_count, int grad,
                        const FusedSlots& SL) {
  if (mode == WHALE_FP64_PLAIN)
    return fused_variant<double, false>(C, st, B, M, W, fam_begin, fam_end, fam_count, grad, SL);
  if (mode == WHALE_FP32_SCALED)
    return fused_variant<float, true>(C, st, B, M, W, fam_begin, fam_end, fam_count, grad, SL);
  return fused_variant<double, true>(C, st, B, M, W, fam_begin, fam_end, fam_count, grad, SL);
}

template <typename R, bool S>
static hipError_t pool_variant(bool store_all, const LaunchCfg& C, hipStream_t st, const BatchDev& B,
                               const ModelDev& M, const WorkRaw& W, uint32_t fam_begin, uint32_t fam_end, int grad) {
  const WorkDev<R> w(W);
  if (C.L.wave_bytes <= kPoolHdrBytes || (grad && !store_all)) return hipErrorInvalidValue;
  const size_t lds = (size_t)C.L.shared_bytes + C.L.wave_bytes;
  if (store_all)
    return go(k_pool<R, S, true>, C.grid, kPoolWaves, lds, st, B, M, w, C.L, fam_begin, fam_end, grad);
  return go(k_pool<R, S, false>, C.grid, kPoolWaves, lds, st, B, M, w, C.L, fam_begin, fam_end, grad);
}

hipError_t launch_pool(int mode, bool store_all, const LaunchCfg& C, hipStream_t st, const BatchDev& B,
                       const ModelDev& M, const WorkRaw& W, uint32_t fam_begin, uint32_t fam_end, int grad) {
  if (mode == WHALE_FP64_PLAIN) return pool_variant<double, false>(store_all, C, st, B, M, W, fam_begin, fam_end, grad);
  if (mode == WHALE_FP32_SCALED) return pool_variant<float, true>(store_all, C, st, B, M, W, fam_begin, fam_end, grad);
  return pool_variant<double, true>(store_all, C, st, B, M, W, fam_begin, fam_end, grad);
}

hipError_t launch_mix(hipStream_t st, const MixComp* X, int K, const double* log_prior, uint32_t n_fam, int cond,
                      double* out, double* resp) {
  if (K < 1 || K > kMaxMix) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mix, dim3(1), dim3(1024), 0, st, X, K, log_prior, n_fam, cond, out, resp);
  return hipGetLastError();
}

hipError_t launch_tables(hipStream_t st, const ModelDev& M, double* eps, double* phi, double* psi, double* drow,
                         double* dnode, const int32_t* hlevels, const int32_t* hnodes, int n_levels) {
  hipLaunchKernelGGL(k_tables, dim3(1), dim3(1024), sizeof(int32_t) * (M.n + 1), st, M, eps, phi, psi, drow, dnode,
                     hlevels, hnodes, n_levels);
  return hipGetLastError();
}

hipError_t launch_tail(hipStream_t st, const ModelDev& M, const double* fam_ll, uint32_t n_fam, const double* acc,
                       double* ebar, const double* drow, const double* dnode, const int32_t* dlevels,
                       const int32_t* dnodes, int n_dlevels, int cond, int with_grad, double* out,
                       const double* n_eff) {
  hipLaunchKernelGGL(k_tail, dim3(1), dim3(kTailThreads), 0, st, M, fam_ll, n_fam, acc, ebar, drow, dnode, dlevels, dnodes,
                     n_dlevels, cond, with_grad, out, n_eff);
  return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, int n) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_f64(hipStream_t st, double* p, double v, size_t n) {
  int grid = (int)std::min<size_t>(1024, (n + 255) / 256 + 1);
  hipLaunchKernelGGL(k_fill_f64, dim3(grid), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k_backtrack: stochastic backtracking of reconciled trees from the stored ℓ
// (backtrack!, src/track.jl:188-414).  One lane per (family, sample): a depth
// first walk over slice states (species node e, clade γ, slice row t) that
// picks each event with probability term / ℓ_e(γ, t), the terms taken in the
// reference's order: inside a branch "no event" (φ_t ℓ(γ, t-1)) then each
// duplication (p ψ_t ℓ(γ1, t-1) ℓ(γ2, t-1)); at a speciation node the two
// losses then both speciations of every split; at a WGD node non-retention
// ((1 - q + 2qε_f) ℓ_f(γ)) then each retained split; at the root each split's
// root duplication (ξ(1-η)/η) and both speciations (η(1-ε)/ξ²), then the two
// losses.  Uniform draws come from a counter-based generator of (seed,
// family, sample), one per inter- or intra-branch step in depth-first order.
// Scaled forms: every term is formed relative to 2^E(γ) of the state's clade.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bt_uniform(uint64_t& st) {
  st += 0x9E3779B97F4A7C15ull;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

struct BtCell {
  double v;  // stored value (0 when the clade is not compatible with the branch)
  int E;     // its clade's exponent (0 in the plain form)
};

struct BtCtx {
  const BtDev& D;
  const ModelDev& M;
  FamDev F;
  uint32_t f;
  bool f64;  // this family's ℓ is in the FP64 region
  __device__ BtCell cell(int e, int g, int t) const {
    BtCell c{0.0, 0};
    const int32_t s = D.node_seg[(size_t)f * M.n + e];
    if (s < 0) return c;
    const SegDev sg = D.seg[s];
    if ((uint32_t)t > sg.R || t < 1) return c;
    uint32_t col;
    if (e == M.root) {
      col = (uint32_t)g;  // the root's columns are all clades in order
    } else {
      // γ's column: the segment's clades are in increasing id
      uint32_t lo = D.seg_item[s], hi = lo + sg.nc;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (D.it_gid[mid] < (uint32_t)g) lo = mid + 1;
        else hi = mid;
      }
      if (lo == D.seg_item[s] + sg.nc || D.it_gid[lo] != (uint32_t)g) return c;
      col = lo - D.seg_item[s];
    }
    const uint64_t idx = F.ell0 + sg.ell + (uint64_t)(t - 1) * sg.nc + col;
    c.v = f64 ? D.ell64[idx] : (D.ell_float ? (double)((const float*)D.ell)[idx] : ((const double*)D.ell)[idx]);
    if (D.E && c.v != 0.0) c.E = D.E[F.clade0 + g];
    return c;
  }
  __device__ BtCell top(int e, int g) const { return cell(e, g, M.node[e].R); }
};

// term x·a·b relative to 2^Eg, a, b two cells
__device__ __forceinline__ double bt_rel2(double x, const BtCell& a, const BtCell& b, int Eg) {
  if (a.v == 0.0 || b.v == 0.0 || x == 0.0) return 0.0;
  return ldexp(x * a.v * b.v, a.E + b.E - Eg);
}
__device__ __forceinline__ double bt_rel1(double x, const BtCell& a, int Eg) {
  if (a.v == 0.0 || x == 0.0) return 0.0;
  return ldexp(x * a.v, a.E - Eg);
}

struct BtState {
  int e, g, t;
};

__global__ void k_backtrack(BtDev D, ModelDev M, uint32_t fam_begin, uint32_t n_fams, uint32_t n_samples, uint64_t seed,
                            uint32_t cap, int32_t* counts, int32_t* nodes, int32_t* stack) {
  const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (uint64_t)n_fams * n_samples) return;
  const uint32_t fi = (uint32_t)(id / n_samples), smp = (uint32_t)(id % n_samples);
  const uint32_t f = fam_begin + fi;
  BtCtx X{D, M, D.fam[f], f, D.fam64 ? D.fam64[f] != 0 : false};
  uint64_t st = seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(f + 1)) ^ (0xD1B54A32D192ED03ull * (uint64_t)(smp + 1));
  int32_t* out = nodes + id * (uint64_t)cap * 4;
  int32_t* stk = stack + id * (uint64_t)cap * 4;
  const double eta = M.par[3 * M.n + M.n_wgd];
  uint32_t n = 0, sp = 0;
  bool over = false;
  auto add = [&](const BtState& s, int par) -> int {
    if (n >= cap) { over = true; return -1; }
    int32_t* o = out + 4 * (uint64_t)n;
    o[0] = s.g; o[1] = s.e; o[2] = s.t; o[3] = par;
    return (int)n++;
  };
  const int gam = (int)X.F.n_clades - 1;
  BtState s{M.root, gam, 1};
  int cur = add(s, -1);
  for (;;) {
    // the state s at node `cur`; a branch of the walk ends at a leaf or a loss
    BtState nx[2];
    int nn = 0;
    bool leafend = false;
    if (s.g < 0) {
      leafend = true;
    } else if (s.t == 1) {
      const NodeDev& nd = M.node[s.e];
      if (nd.kind == WHALE_NODE_LEAF) {
        leafend = true;
      } else {
        const BtCell me = X.cell(s.e, s.g, 1);
        const int Eg = me.E;
        double r = bt_uniform(st) * me.v;
        BtState last[2];
        int nlast = 0;
        auto term = [&](double x, BtState a, BtState b, int cnt) {
          if (nn > 0 || !(x > 0.0)) return;
          last[0] = a; last[1] = b; nlast = cnt;
          r -= x;
          if (r < 0.0) { nx[0] = a; nx[1] = b; nn = cnt; }
        };
        const uint32_t ca = X.F.clade0 + (uint32_t)s.g;
        const uint32_t k0 = D.spl_off[ca], k1 = D.spl_off[ca + 1];
        if (nd.kind == WHALE_NODE_ROOT) {
          const double eps = eps_top(M, s.e), xi = 1.0 - (1.0 - eta) * eps;
          const double cA = xi * (1.0 - eta) / eta, cB = eta * (1.0 - eps) / (xi * xi);
          const int fc = nd.child0, gc = nd.child1;
          const int tf = M.node[fc].R, tg = M.node[gc].R;
          for (uint32_t k = k0; k < k1 && nn == 0; ++k) {
            const int g1 = (int)(D.spl_g[k] & 0xffff), g2 = (int)(D.spl_g[k] >> 16);
            const double p = D.split_p[k];
            term(bt_rel2(p * cA, X.cell(s.e, g1, 1), X.cell(s.e, g2, 1), Eg), {s.e, g1, 1}, {s.e, g2, 1}, 2);
            term(bt_rel2(p * cB, X.top(fc, g1), X.top(gc, g2), Eg), {fc, g1, tf}, {gc, g2, tg}, 2);
            term(bt_rel2(p * cB, X.top(gc, g1), X.top(fc, g2), Eg), {gc, g1, tg}, {fc, g2, tf}, 2);
          }
          term(bt_rel1(eps_top(M, gc) * cB, X.top(fc, s.g), Eg), {fc, s.g, tf}, {gc, -1, 0}, 2);
          term(bt_rel1(eps_top(M, fc) * cB, X.top(gc, s.g), Eg), {gc, s.g, tg}, {fc, -1, 0}, 2);
        } else if (nd.kind == WHALE_NODE_WGD) {
          const double q = M.par[3 * M.n + nd.wgd];
          const int fc = nd.child0, tf = M.node[fc].R;
          term(bt_rel1(1.0 - q + 2.0 * q * eps_top(M, fc), X.top(fc, s.g), Eg), {fc, s.g, tf}, {0, 0, 0}, 1);
          for (uint32_t k = k0; k < k1 && nn == 0; ++k) {
            const int g1 = (int)(D.spl_g[k] & 0xffff), g2 = (int)(D.spl_g[k] >> 16);
            term(bt_rel2(q * D.split_p[k], X.top(fc, g1), X.top(fc, g2), Eg), {fc, g1, tf}, {fc, g2, tf}, 2);
          }
        } else {
          const int fc = nd.child0, gc = nd.child1;
          const int tf = M.node[fc].R, tg = M.node[gc].R;
          term(bt_rel1(eps_top(M, gc), X.top(fc, s.g), Eg), {fc, s.g, tf}, {gc, -1, 0}, 2);
          term(bt_rel1(eps_top(M, fc), X.top(gc, s.g), Eg), {gc, s.g, tg}, {fc, -1, 0}, 2);
          for (uint32_t k = k0; k < k1 && nn == 0; ++k) {
            const int g1 = (int)(D.spl_g[k] & 0xffff), g2 = (int)(D.spl_g[k] >> 16);
            const double p = D.split_p[k];
            term(bt_rel2(p, X.top(fc, g1), X.top(gc, g2), Eg), {fc, g1, tf}, {gc, g2, tg}, 2);
            term(bt_rel2(p, X.top(gc, g1), X.top(fc, g2), Eg), {gc, g1, tg}, {fc, g2, tf}, 2);
          }
        }
        if (nn == 0) {  // rounding left r >= 0: the last non-zero term
          nn = nlast;
          nx[0] = last[0];
          nx[1] = last[1];
        }
        if (nn == 0) leafend = true;
      }
    } else if (D.clade_size[X.F.clade0 + (uint32_t)s.g] == 1) {
      s.t = 1;  // a one-gene clade does not duplicate along the branch
      continue;
    } else {
      const NodeDev& nd = M.node[s.e];
      const BtCell me = X.cell(s.e, s.g, s.t);
      const int Eg = me.E;
      double r = bt_uniform(st) * me.v;
      const double stay = bt_rel1(M.phi[nd.tab + s.t - 1], X.cell(s.e, s.g, s.t - 1), Eg);
      r -= stay;
      if (r < 0.0) {
        s.t -= 1;
        continue;
      }
      const double ps = M.psi[nd.tab + s.t - 1];
      const uint32_t ca = X.F.clade0 + (uint32_t)s.g;
      bool any = false;
      for (uint32_t k = D.spl_off[ca]; k < D.spl_off[ca + 1]; ++k) {
        const int g1 = (int)(D.spl_g[k] & 0xffff), g2 = (int)(D.spl_g[k] >> 16);
        const double x = bt_rel2(D.split_p[k] * ps, X.cell(s.e, g1, s.t - 1), X.cell(s.e, g2, s.t - 1), Eg);
        if (x > 0.0) {
          nx[0] = {s.e, g1, s.t - 1};
          nx[1] = {s.e, g2, s.t - 1};
          any = true;
          r -= x;
          if (r < 0.0) break;
        }
      }
      if (any) {
        nn = 2;  // the split where r fell below 0, or (rounding) the last non-zero one
      } else {
        s.t -= 1;  // only "no event" has weight
        continue;
      }
    }
    if (!leafend) {
      // next states in the reference's order: the first is walked first
      for (int j = nn - 1; j >= 0; --j) {
        if (sp >= cap) { over = true; break; }
        int32_t* q = stk + 4 * (uint64_t)sp++;
        q[0] = nx[j].e; q[1] = nx[j].g; q[2] = nx[j].t; q[3] = cur;
      }
    }
    if (over || sp == 0) break;
    // the next pending state: a new node when its (clade, branch) differs
    // from the node it came from ((b::BackTracker)(newstate), src/track.jl:172-180)
    const int32_t* q = stk + 4 * (uint64_t)(--sp);
    const BtState ns{q[0], q[1], q[2]};
    const int from = q[3];
    const int32_t* fn = out + 4 * (uint64_t)from;
    if (ns.g != fn[0] || ns.e != fn[1]) {
      cur = add(ns, from);
      if (cur < 0) break;
    } else {
      cur = from;
    }
    s = ns;
  }
  counts[id] = over ? -1 : (int32_t)n;
}

hipError_t launch_backtrack(hipStream_t st, const BtDev& D, const ModelDev& M, uint32_t fam_begin, uint32_t n_fams,
                            uint32_t n_samples, uint64_t seed, uint32_t cap, int32_t* counts, int32_t* nodes,
                            int32_t* stack) {
  const uint64_t lanes = (uint64_t)n_fams * n_samples;
  if (lanes == 0) return hipSuccess;
  const uint32_t grid = (uint32_t)((lanes + 63) / 64);
  hipLaunchKernelGGL(k_backtrack, dim3(grid), dim3(64), 0, st, D, M, fam_begin, n_fams, n_samples, seed, cap, counts,
                     nodes, stack);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k_det_sum: the batch's gradient accumulators as the sum of the per-family
// slots in family order (one lane per accumulator), so that the result does
// not depend on which workgroup evaluated which family, or when.
// ---------------------------------------------------------------------------
__global__ void k_det_sum(const double* slots, uint32_t n_fam, int ncol, double* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  double s = 0.0;
  for (uint32_t f = 0; f < n_fam; ++f) s += slots[(size_t)f * ncol + j];
  out[j] = s;
}

hipError_t launch_det_sum(hipStream_t st, const double* slots, uint32_t n_fam, int ncol, double* out) {
  hipLaunchKernelGGL(k_det_sum, dim3((ncol + 255) / 256), dim3(256), 0, st, slots, n_fam, ncol, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fixed workspace of the evaluation that allocates nothing
// ---------------------------------------------------------------------------
// Every GPU the library runs on holds one copy of this array until the
// process ends.  Host code never passes its name to a kernel (on the host it
// is a stand-in, not a device address): the device address of each GPU's copy
// comes from hipGetSymbolAddress, once per device, and only that is used.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

hipError_t scratch_address(int device, void** out) {
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *out = addr[device];
  return hipSuccess;
}

}  // namespace whale
