// gfx950 kernels of the amalgamated-likelihood DP.
//
//   k_tables       slice tables ε, φ, ψ of every branch (setmodel!/setslices!,
//                  src/model.jl:162-191; getα/_ϵ/_ϕ/_ψ, src/bdputil.jl:6-11)
//                  plus the per-row partial derivatives the reverse sweep needs
//   k_forward      the DP of every family (whale!/whalewgd!/within_branch!/
//                  whaleroot!, src/core.jl:83-158), one family per workgroup,
//                  stage by stage, rows in lockstep, row buffers in LDS
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
  explicit WorkDev(const WorkRaw& w)
      : ell((Real*)w.ell), coef((Real*)w.coef), E(w.E), ibar((Real*)w.ibar), fam_ll(w.fam_ll), acc(w.acc),
        ticket(w.ticket), flag(w.flag), fb(w.fb) {}
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
// k_tables: one workgroup; nodes by height (children first), one lane per node
// ---------------------------------------------------------------------------
// deriv layout per table row (9 doubles): dε/dα, dε/dβ, dε/dε', dφ/dα, dφ/dβ,
// dφ/dε', dψ/dα, dψ/dβ, dψ/dε'.  Per node (4 doubles): dα/dλ, dα/dμ, dβ/dλ, dβ/dμ.
__global__ void k_tables(ModelDev M, double* eps, double* phi, double* psi, double* drow, double* dnode,
                         const int32_t* hlevels, const int32_t* hnodes, int n_levels) {
  const double* lam = M.par;
  const double* mu = M.par + M.n;
  const double* p = M.par + 2 * M.n;
  const double* q = M.par + 3 * M.n;
  for (int L = 0; L < n_levels; ++L) {
    for (int k = hlevels[L] + threadIdx.x; k < hlevels[L + 1]; k += blockDim.x) {
      int e = hnodes[k];
      const NodeDev& d = M.node[e];
      int t0 = d.tab;
      double e0;
      if (d.kind == WHALE_NODE_WGD) {
        double ec = eps_top(M, d.child0);
        double qq = q[d.wgd];
        e0 = qq * (ec * ec) + (1.0 - qq) * ec;
      } else if (d.kind == WHALE_NODE_LEAF) {
        e0 = p[e];
      } else {
        e0 = eps_top(M, d.child0) * eps_top(M, d.child1);
      }
      eps[t0] = e0;
      phi[t0] = 1.0;
      psi[t0] = 1.0;
      double l = lam[e], u = mu[e], t = d.delta;
      double a = 0, b = 0, dal = 0, dam = 0, dbl = 0, dbm = 0;
      if (d.R > 1) {
        bool close = (l == u) || (isfinite(l) && isfinite(u) && fabs(l - u) <= 1e-6);
        if (close) {
          double x = 1.0 + l * t;
          a = l * t / x;
          dal = t / (x * x);
          dam = 0.0;
        } else {
          double E = exp(t * (l - u));
          double N = u * (E - 1.0), D = l * E - u;
          a = N / D;
          double dNl = u * t * E, dNm = (E - 1.0) - u * t * E;
          double dDl = E + l * t * E, dDm = -l * t * E - 1.0;
          dal = (dNl * D - N * dDl) / (D * D);
          dam = (dNm * D - N * dDm) / (D * D);
        }
        b = (l / u) * a;
        dbl = a / u + (l / u) * dal;
        dbm = -l * a / (u * u) + (l / u) * dam;
      }
      dnode[4 * e + 0] = dal;
      dnode[4 * e + 1] = dam;
      dnode[4 * e + 2] = dbl;
      dnode[4 * e + 3] = dbm;
      double ep = e0;
      for (int r = 1; r < d.R; ++r) {
        double den = 1.0 - b * ep;
        double num = a + (1.0 - a - b) * ep;
        double en = num / den;
        double ph = (1.0 - a) * (1.0 - b) / (den * den);
        double ps = (1.0 - a) * (1.0 - b) * b / (den * den * den);
        eps[t0 + r] = en;
        phi[t0 + r] = ph;
        psi[t0 + r] = ps;
        double* g = drow + 9 * (size_t)(t0 + r);
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
        ep = en;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// wave-level segmented sum: items of one segment sit on consecutive lanes
// ---------------------------------------------------------------------------
template <typename Real>
__device__ __forceinline__ void seg_add(Real v, uint32_t key, double* addr) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    Real vu = __shfl_up(v, off, 64);
    uint32_t ku = __shfl_up(key, off, 64);
    if (lane >= off && ku == key) v += vu;
  }
  uint32_t kn = __shfl_down(key, 1, 64);
  bool tail = (lane == 63) || (kn != key);
  if (tail && key != kNone && v != Real(0)) unsafeAtomicAdd(addr, (double)v);
}

template <typename Real>
__device__ __forceinline__ Real wave_sum(Real v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// coefficient of record r in the scaled forms: p·2^(E1+E2-E), 0 while any exponent is open
template <typename Real>
__device__ __forceinline__ Real scaled_coef(const BatchDev& B, const FamDev& F, const int32_t* E, uint32_t item0,
                                            uint32_t rec, int Eg) {
  uint32_t ops = B.rec_ops[rec];
  int e1 = E[F.clade0 + B.it_gid[item0 + (ops & 0xffff)]];
  int e2 = E[F.clade0 + B.it_gid[item0 + (ops >> 16)]];
  if (Eg == kUndef || e1 == kUndef || e2 == kUndef) return Real(0);
  return ldexp_r((Real)B.split_p[F.split0 + B.rec_sid[rec]], e1 + e2 - Eg);
}

// Every workgroup is a single wavefront (kWG == 64).  A wave issues its LDS
// instructions in order and the LDS executes them in order, so between the
// rows of a stage it is enough to stop the compiler from moving LDS accesses
// across the row boundary; no s_barrier and no wait on outstanding global
// stores is needed.  Stage boundaries (where global ℓ written earlier is read
// back) keep __syncthreads().
static_assert(kWG == 64, "row synchronisation assumes one wavefront per workgroup");
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// row-invariant data of one item, cached in registers for a stage's rows
// ---------------------------------------------------------------------------
constexpr int kCache = 4;  // items per lane held in registers (stages of up to 256 items)

struct ItemRow {
  uint32_t cell;    // ℓ cell of row 0, relative to the family
  uint32_t stride;  // |C_e| of the item's segment (row pitch)
  uint32_t tab;     // first table row of the segment's node
  uint32_t r0, r1;  // records
  uint32_t q0, q1;  // inverse records
  uint32_t R;       // rows of the segment
  uint32_t gid;     // family-local clade id
  uint32_t seg;     // absolute segment (reduction key)
};

__device__ __forceinline__ ItemRow load_item(const BatchDev& B, const ModelDev& M, uint32_t ia, uint32_t i) {
  ItemRow it;
  it.seg = B.it_seg[ia];
  const SegDev sg = B.seg[it.seg];
  it.cell = sg.ell + (i - sg.item0);
  it.stride = sg.nc;
  it.tab = M.node[sg.node].tab;
  it.r0 = B.it_rec[ia];
  it.r1 = B.it_rec[ia + 1];
  it.q0 = B.it_inv[ia];
  it.q1 = B.it_inv[ia + 1];
  it.R = sg.R;
  it.gid = B.it_gid[ia];
  return it;
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <typename Real, bool Scaled, bool StoreAll>
__global__ __launch_bounds__(kWG) void k_forward(BatchDev B, ModelDev M, WorkDev<Real> W, uint32_t cap,
                                                  uint32_t rcap, int32_t fam_single, uint32_t fam_begin,
                                                  uint32_t fam_end, const uint32_t* fam_count) {
  if (fam_count) fam_end = fam_begin + *fam_count;
  extern __shared__ __align__(16) unsigned char smem[];
  Real* bA = reinterpret_cast<Real*>(smem);
  Real* bB = bA + cap;
  Real* bF = bB + cap;
  Real* bG = bF + cap;
  Real* sCoef = bG + cap;                                    // rcap coefficients of the stage's records
  uint32_t* sOps = reinterpret_cast<uint32_t*>(sCoef + rcap);  // rcap operand pairs
  __shared__ uint32_t s_fam;
  __shared__ int s_dirty;
  const int lane = threadIdx.x;
  const Real* coefs = Scaled ? W.coef : reinterpret_cast<const Real*>(B.rec_p);
  bool first = true;
  for (;;) {
    uint32_t f;
    if (fam_single >= 0) {
      if (!first) break;
      f = (uint32_t)fam_single;
    } else {
      if (lane == 0) s_fam = fam_begin + atomicAdd(W.ticket, 1u);
      __syncthreads();
      uint32_t k = s_fam;
      __syncthreads();
      if (k >= fam_end) break;
      f = B.order[k];
    }
    first = false;
    const FamDev F = B.fam[f];
    Real* ell = W.ell + F.ell0;
    int32_t* E = W.E;
    // scaled forms: coefficients of the stage's records, into LDS and (for the reverse sweep) global
    auto refresh_coefs = [&](const StageDev& S, uint32_t rb) {
      for (uint32_t i = lane; i < S.n_items; i += kWG) {
        const uint32_t ia = S.item0 + i;
        const int Eg = E[F.clade0 + B.it_gid[ia]];
        for (uint32_t k = B.it_rec[ia]; k < B.it_rec[ia + 1]; ++k) {
          const Real c = scaled_coef<Real>(B, F, E, S.item0, k, Eg);
          sCoef[k - rb] = c;
          W.coef[k] = c;
        }
      }
    };
    if (Scaled) {
      for (uint32_t c = lane; c < F.n_clades; c += kWG) E[F.clade0 + c] = kUndef;
      __syncthreads();
    }
    // ---------------- non-root stages
    for (uint32_t s = 0; s < F.n_stages; ++s) {
      const StageDev S = B.stage[F.stage0 + s];
      const uint32_t n = S.n_items;
      const uint32_t rb = B.it_rec[S.item0];             // the stage's records are [rb, rb + nrec)
      const uint32_t nrec = B.it_rec[S.item0 + n] - rb;  // <= rcap (size class)
      for (uint32_t i = lane; i < n; i += kWG) {
        int32_t cf = B.it_cf[S.item0 + i], cg = B.it_cg[S.item0 + i];
        bF[i] = cf >= 0 ? ell[cf] : Real(0);
        bG[i] = cg >= 0 ? ell[cg] : Real(0);
      }
      for (uint32_t k = lane; k < nrec; k += kWG) {
        sOps[k] = B.rec_ops[rb + k];
        if (!Scaled) sCoef[k] = coefs[rb + k];
      }
      __syncthreads();
      // row 1
      for (uint32_t i = lane; i < n; i += kWG) {
        const uint32_t ia = S.item0 + i;
        const SegDev sg = B.seg[B.it_seg[ia]];
        const NodeDev& nd = M.node[sg.node];
        const uint32_t gid = B.it_gid[ia];
        const bool leafc = B.clade_size[F.clade0 + gid] == 1;
        const uint32_t r0 = B.it_rec[ia], r1 = B.it_rec[ia + 1];
        Real v = 0;
        if (nd.kind == WHALE_NODE_LEAF) {
          if (leafc) {
            if (Scaled) {
              int Eg = E[F.clade0 + gid];
              if (Eg == kUndef) {
                int ex;
                v = frexp_r((Real)nd.leafP, &ex);
                E[F.clade0 + gid] = ex;
              } else {
                v = ldexp_r((Real)nd.leafP, -Eg);
              }
            } else {
              v = (Real)nd.leafP;
            }
          }
        } else if (!Scaled) {
          if (nd.kind == WHALE_NODE_WGD) {
            const Real qq = (Real)M.par[3 * M.n + nd.wgd];
            const Real ef = (Real)eps_top(M, nd.child0);
            Real r = 0;
            for (uint32_t k = r0 - rb; k < r1 - rb; ++k) {
              uint32_t ops = sOps[k];
              r += sCoef[k] * bF[ops & 0xffff] * bF[ops >> 16];
            }
            v = r * qq + ((Real(1) - qq) * bF[i] + (Real(2) * qq) * ef * bF[i]);
          } else {
            const Real ef = (Real)eps_top(M, nd.child0), eg = (Real)eps_top(M, nd.child1);
            Real sp = 0;
            for (uint32_t k = r0 - rb; k < r1 - rb; ++k) {
              uint32_t ops = sOps[k];
              uint32_t a = ops & 0xffff, b = ops >> 16;
              sp += sCoef[k] * (bF[a] * bG[b] + bG[a] * bF[b]);
            }
            v = sp + (bF[i] * eg + bG[i] * ef);
          }
        } else {
          // scaled: explicit exponents (operands come from other segments)
          const bool wgd = nd.kind == WHALE_NODE_WGD;
          const Real qq = wgd ? (Real)M.par[3 * M.n + nd.wgd] : Real(1);
          const Real ef = (Real)eps_top(M, nd.child0);
          const Real eg = wgd ? Real(0) : (Real)eps_top(M, nd.child1);
          int Eg = E[F.clade0 + gid];
          XSum<Real> xs;
          Real base = 0;  // terms already at scale Eg
          for (uint32_t k = r0; k < r1; ++k) {
            uint32_t ops = B.rec_ops[k];
            uint32_t a = ops & 0xffff, b = ops >> 16;
            Real pk = (Real)B.split_p[F.split0 + B.rec_sid[k]];
            Real fa = bF[a], fb = bF[b];
            Real ga = wgd ? fa : bG[a], gb = wgd ? fb : bG[b];
            int ea = E[F.clade0 + B.it_gid[S.item0 + a]], eb = E[F.clade0 + B.it_gid[S.item0 + b]];
            Real m; int ex;
            if (fa != Real(0) && gb != Real(0)) { xprod(pk * qq, fa, ea, gb, eb, &m, &ex); xs.add(m, ex); }
            if (!wgd && ga != Real(0) && fb != Real(0)) { xprod(pk, ga, ea, fb, eb, &m, &ex); xs.add(m, ex); }
          }
          if (Eg == kUndef) {
            if (xs.k != kUndef) {
              int ex;
              v = ldexp_r(frexp_r(xs.s, &ex), -kBirth);
              E[F.clade0 + gid] = xs.k + ex + kBirth;
            }
          } else {
            if (wgd) base = ((Real(1) - qq) * bF[i] + (Real(2) * qq) * ef * bF[i]);
            else base = bF[i] * eg + bG[i] * ef;
            v = base + (xs.k == kUndef ? Real(0) : ldexp_r(xs.s, xs.k - Eg));
          }
        }
        bA[i] = v;
        const uint32_t j = i - sg.item0;
        if (StoreAll || sg.R == 1) ell[sg.ell + j] = v;
      }
      __syncthreads();
      if (Scaled) {
        refresh_coefs(S, rb);
        __syncthreads();
      }
      // rows 2..R
      Real* cur = bA;
      Real* nxt = bB;
      ItemRow cache[kCache];
#pragma unroll
      for (int c = 0; c < kCache; ++c) {
        const uint32_t i = lane + c * kWG;
        if (i < n) cache[c] = load_item(B, M, S.item0 + i, i);
      }
      for (uint32_t r = 1; r < S.r_max; ++r) {
        if (Scaled && lane == 0) s_dirty = 0;  // one wave per workgroup: ordered after last row's read
        auto update = [&](const ItemRow& it, uint32_t i) {
          Real v = cur[i];
          if (r < it.R) {
            const Real ph = (Real)M.phi[it.tab + r], ps = (Real)M.psi[it.tab + r];
            bool pending = false;
            if (Scaled && v == Real(0) && it.r1 > it.r0) pending = E[F.clade0 + it.gid] == kUndef;
            if (!pending) {
              Real acc = 0;
              for (uint32_t k = it.r0 - rb; k < it.r1 - rb; ++k) {
                uint32_t ops = sOps[k];
                acc += sCoef[k] * cur[ops & 0xffff] * cur[ops >> 16];
              }
              v = ph * v + ps * acc;
            } else {
              XSum<Real> xs;
              for (uint32_t k = it.r0; k < it.r1; ++k) {
                uint32_t ops = sOps[k - rb];
                uint32_t a = ops & 0xffff, b = ops >> 16;
                Real va = cur[a], vb = cur[b];
                if (va == Real(0) || vb == Real(0)) continue;
                int ea = E[F.clade0 + B.it_gid[S.item0 + a]], eb = E[F.clade0 + B.it_gid[S.item0 + b]];
                Real m; int ex;
                xprod((Real)B.split_p[F.split0 + B.rec_sid[k]], va, ea, vb, eb, &m, &ex);
                xs.add(m, ex);
              }
              if (xs.k != kUndef) {
                int ex;
                v = ldexp_r(frexp_r(ps * xs.s, &ex), -kBirth);
                E[F.clade0 + it.gid] = xs.k + ex + kBirth;
                s_dirty = 1;
              }
            }
            if (StoreAll || r == it.R - 1) ell[it.cell + (size_t)r * it.stride] = v;
          }
          nxt[i] = v;
        };
#pragma unroll
        for (int c = 0; c < kCache; ++c) {
          const uint32_t i = lane + c * kWG;
          if (i < n) update(cache[c], i);
        }
        for (uint32_t i = lane + kCache * kWG; i < n; i += kWG) update(load_item(B, M, S.item0 + i, i), i);
        wave_sync();
        if (Scaled && s_dirty) {
          // an exponent was fixed in this row (E is global: full barrier)
          __syncthreads();
          refresh_coefs(S, rb);
          __syncthreads();
        }
        Real* t = cur; cur = nxt; nxt = t;
      }
      __syncthreads();  // this stage's ℓ is read by later stages
    }
    // ---------------- root: all clades, by size level
    {
      const uint32_t n = F.n_clades;
      const uint32_t i0 = F.root_item0;
      const SegDev sg = B.seg[F.root_seg];
      const NodeDev& nd = M.node[M.root];
      const double eta = M.par[3 * M.n + M.n_wgd];
      const double er = eps_top(M, M.root);
      const double xi = 1.0 - (1.0 - eta) * er;
      const Real cA = (Real)((1.0 - eta) * xi / eta);
      const Real cB = (Real)(eta * (1.0 - er) / (xi * xi));
      const Real ef = (Real)eps_top(M, nd.child0), eg = (Real)eps_top(M, nd.child1);
      for (uint32_t i = lane; i < n; i += kWG) {
        int32_t cf = B.it_cf[i0 + i], cg = B.it_cg[i0 + i];
        bF[i] = cf >= 0 ? ell[cf] : Real(0);
        bG[i] = cg >= 0 ? ell[cg] : Real(0);
      }
      __syncthreads();
      for (uint32_t L = 0; L < F.n_lvl; ++L) {
        const uint32_t a0 = B.lvl[F.lvl0 + L], a1 = B.lvl[F.lvl0 + L + 1];
        for (uint32_t i = a0 + lane; i < a1; i += kWG) {
          const uint32_t ia = i0 + i;
          const uint32_t r0 = B.it_rec[ia], r1 = B.it_rec[ia + 1];
          Real v;
          if (!Scaled) {
            Real a = 0, b = 0;
            for (uint32_t k = r0; k < r1; ++k) {
              uint32_t ops = B.rec_ops[k];
              uint32_t x = ops & 0xffff, y = ops >> 16;
              a += coefs[k] * bA[x] * bA[y];
              b += coefs[k] * (bF[x] * bG[y] + bG[x] * bF[y]);
            }
            Real c = bF[i] * eg + bG[i] * ef;
            v = cA * a + cB * (b + c);
          } else {
            int Eg = E[F.clade0 + i];
            XSum<Real> xs;
            for (uint32_t k = r0; k < r1; ++k) {
              uint32_t ops = B.rec_ops[k];
              uint32_t x = ops & 0xffff, y = ops >> 16;
              Real pk = (Real)B.split_p[F.split0 + B.rec_sid[k]];
              int ex_ = E[F.clade0 + x], ey = E[F.clade0 + y];
              Real m; int ee;
              if (bA[x] != Real(0) && bA[y] != Real(0)) { xprod(pk * cA, bA[x], ex_, bA[y], ey, &m, &ee); xs.add(m, ee); }
              if (bF[x] != Real(0) && bG[y] != Real(0)) { xprod(pk * cB, bF[x], ex_, bG[y], ey, &m, &ee); xs.add(m, ee); }
              if (bG[x] != Real(0) && bF[y] != Real(0)) { xprod(pk * cB, bG[x], ex_, bF[y], ey, &m, &ee); xs.add(m, ee); }
            }
            if (Eg == kUndef) {
              v = 0;
              if (xs.k != kUndef) {
                int ex;
                v = ldexp_r(frexp_r(xs.s, &ex), -kBirth);
                E[F.clade0 + i] = xs.k + ex + kBirth;
              }
            } else {
              Real c = bF[i] * eg + bG[i] * ef;
              v = cB * c + (xs.k == kUndef ? Real(0) : ldexp_r(xs.s, xs.k - Eg));
            }
          }
          bA[i] = v;
          ell[sg.ell + i] = v;
        }
        __syncthreads();
      }
      if (lane == 0) {
        Real L = bA[n - 1];
        double ll;
        if (L > Real(0) && isfinite((double)L)) {
          ll = log((double)L);
          if (Scaled) ll += (double)E[F.clade0 + n - 1] * 0.69314718055994530942;
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
        for (uint32_t i = lane; i < n; i += kWG) {
          const uint32_t ia = i0 + i;
          int Eg = E[F.clade0 + i];
          for (uint32_t k = B.it_rec[ia]; k < B.it_rec[ia + 1]; ++k) W.coef[k] = scaled_coef<Real>(B, F, E, i0, k, Eg);
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// backward (reverse sweep); requires a forward with StoreAll
// ---------------------------------------------------------------------------
template <typename Real, bool Scaled>
__global__ __launch_bounds__(kWG) void k_backward(BatchDev B, ModelDev M, WorkDev<Real> W, uint32_t cap,
                                                   uint32_t rcap, int32_t fam_single, uint32_t fam_begin,
                                                   uint32_t fam_end, const uint32_t* fam_count) {
  if (fam_count) fam_end = fam_begin + *fam_count;
  extern __shared__ __align__(16) unsigned char smem[];
  Real* bA = reinterpret_cast<Real*>(smem);  // d logL / d ℓ, current row
  Real* bB = bA + cap;                        // next row
  Real* bV = bB + cap;                        // ℓ of the row below / child0 tops
  Real* bG = bV + cap;                        // child1 tops
  Real* sCoef = bG + cap;                                      // the stage's record coefficients
  uint32_t* sOps = reinterpret_cast<uint32_t*>(sCoef + rcap);  // and operand pairs
  __shared__ uint32_t s_fam;
  const int lane = threadIdx.x;
  const Real* coefs = Scaled ? W.coef : reinterpret_cast<const Real*>(B.rec_p);
  const AccLayout AL(M.tab_rows, M.n, M.n_wgd);
  double* acc = W.acc;
  bool first = true;
  for (;;) {
    uint32_t f;
    if (fam_single >= 0) {
      if (!first) break;
      f = (uint32_t)fam_single;
    } else {
      if (lane == 0) s_fam = fam_begin + atomicAdd(W.ticket, 1u);
      __syncthreads();
      uint32_t k = s_fam;
      __syncthreads();
      if (k >= fam_end) break;
      f = B.order[k];
    }
    first = false;
    if (Scaled && sizeof(Real) == sizeof(float) && W.flag[f]) continue;  // done by the FP64 pass
    const FamDev F = B.fam[f];
    const Real* ell = W.ell + F.ell0;
    Real* ibar = W.ibar + F.ibar0;
    // ---------------- root
    {
      const uint32_t n = F.n_clades;
      const uint32_t i0 = F.root_item0;
      const SegDev sg = B.seg[F.root_seg];
      const NodeDev& nd = M.node[M.root];
      const double eta = M.par[3 * M.n + M.n_wgd];
      const double er = eps_top(M, M.root);
      const double xi = 1.0 - (1.0 - eta) * er;
      const Real cA = (Real)((1.0 - eta) * xi / eta);
      const Real cB = (Real)(eta * (1.0 - er) / (xi * xi));
      const Real ef = (Real)eps_top(M, nd.child0), eg = (Real)eps_top(M, nd.child1);
      for (uint32_t i = lane; i < n; i += kWG) {
        int32_t cf = B.it_cf[i0 + i], cg = B.it_cg[i0 + i];
        bV[i] = ell[sg.ell + i];
        bB[i] = cf >= 0 ? ell[cf] : Real(0);  // child0 tops
        bG[i] = cg >= 0 ? ell[cg] : Real(0);
        bA[i] = 0;
      }
      __syncthreads();
      const Real Lr = bV[n - 1];
      const bool ok = Lr > Real(0) && isfinite((double)Lr);
      if (ok) {
        Real sA = 0, sBc = 0, sF = 0, sG = 0;
        for (int L = (int)F.n_lvl - 1; L >= 0; --L) {
          const uint32_t a0 = B.lvl[F.lvl0 + L], a1 = B.lvl[F.lvl0 + L + 1];
          for (uint32_t i = a0 + lane; i < a1; i += kWG) {
            const uint32_t ia = i0 + i;
            Real g = (i == n - 1) ? Real(1) / Lr : Real(0);
            for (uint32_t k = B.it_inv[ia]; k < B.it_inv[ia + 1]; ++k) {
              uint32_t ops = B.inv_ops[k];
              g += cA * coefs[B.inv_rec[k]] * bA[ops & 0xffff] * bV[ops >> 16];
            }
            bA[i] = g;
          }
          __syncthreads();
        }
        for (uint32_t i = lane; i < n; i += kWG) {
          const uint32_t ia = i0 + i;
          const Real g = bA[i];
          Real sf = 0, sgg = 0;
          for (uint32_t k = B.it_inv[ia]; k < B.it_inv[ia + 1]; ++k) {
            uint32_t ops = B.inv_ops[k];
            Real c = coefs[B.inv_rec[k]] * bA[ops & 0xffff];
            uint32_t sib = ops >> 16;
            sf += c * bG[sib];
            sgg += c * bB[sib];
          }
          ibar[sg.ibar + i] = cB * (sf + g * eg);
          ibar[sg.ibar + n + i] = cB * (sgg + g * ef);
          Real a = 0, b = 0;
          for (uint32_t k = B.it_rec[ia]; k < B.it_rec[ia + 1]; ++k) {
            uint32_t ops = B.rec_ops[k];
            uint32_t x = ops & 0xffff, y = ops >> 16;
            a += coefs[k] * bV[x] * bV[y];
            b += coefs[k] * (bB[x] * bG[y] + bG[x] * bB[y]);
          }
          Real c = bB[i] * eg + bG[i] * ef;
          sA += g * a;
          sBc += g * (b + c);
          sF += g * bB[i];
          sG += g * bG[i];
        }
        sA = wave_sum(sA);
        sBc = wave_sum(sBc);
        sF = wave_sum(sF);
        sG = wave_sum(sG);
        if (lane == 0) {
          unsafeAtomicAdd(acc + AL.rootA, (double)sA);
          unsafeAtomicAdd(acc + AL.rootB, (double)sBc);
          unsafeAtomicAdd(acc + AL.eps + nd.child1, (double)cB * (double)sF);
          unsafeAtomicAdd(acc + AL.eps + nd.child0, (double)cB * (double)sG);
        }
      } else {
        for (uint32_t i = lane; i < 2 * n; i += kWG) ibar[sg.ibar + i] = 0;
      }
      __syncthreads();
    }
    // ---------------- non-root stages, top down
    for (int s = (int)F.n_stages - 1; s >= 0; --s) {
      const StageDev S = B.stage[F.stage0 + s];
      const uint32_t n = S.n_items;
      const uint32_t rb = B.it_rec[S.item0];
      const uint32_t nrec = B.it_rec[S.item0 + n] - rb;  // <= rcap (size class)
      for (uint32_t i = lane; i < n; i += kWG) bA[i] = ibar[B.it_bar[S.item0 + i]];
      for (uint32_t k = lane; k < nrec; k += kWG) {
        sOps[k] = B.rec_ops[rb + k];
        sCoef[k] = coefs[rb + k];
      }
      __syncthreads();
      ItemRow cache[kCache];
#pragma unroll
      for (int c = 0; c < kCache; ++c) {
        const uint32_t i = lane + c * kWG;
        if (i < n) cache[c] = load_item(B, M, S.item0 + i, i);
      }
      for (int r = (int)S.r_max - 1; r >= 1; --r) {
        // ℓ of the row below, for the items whose segment reaches row r
#pragma unroll
        for (int c = 0; c < kCache; ++c) {
          const uint32_t i = lane + c * kWG;
          if (i < n && r < (int)cache[c].R) bV[i] = ell[cache[c].cell + (size_t)(r - 1) * cache[c].stride];
        }
        for (uint32_t i = lane + kCache * kWG; i < n; i += kWG) {
          const ItemRow it = load_item(B, M, S.item0 + i, i);
          if (r < (int)it.R) bV[i] = ell[it.cell + (size_t)(r - 1) * it.stride];
        }
        wave_sync();
        // all lanes call this together (segmented reductions)
        auto update = [&](const ItemRow& it, uint32_t i, bool valid) {
          Real cphi = 0, cpsi = 0;
          uint32_t key = kNone;
          double* addr = acc;
          if (valid) {
            Real g = bA[i];
            if (r < (int)it.R) {
              const Real ph = (Real)M.phi[it.tab + r], ps = (Real)M.psi[it.tab + r];
              Real sd = 0;
              for (uint32_t k = it.q0; k < it.q1; ++k) {
                uint32_t ops = B.inv_ops[k];
                sd += sCoef[B.inv_rec[k] - rb] * bA[ops & 0xffff] * bV[ops >> 16];
              }
              Real fo = 0;
              for (uint32_t k = it.r0 - rb; k < it.r1 - rb; ++k) {
                uint32_t ops = sOps[k];
                fo += sCoef[k] * bV[ops & 0xffff] * bV[ops >> 16];
              }
              bB[i] = ph * g + ps * sd;
              cphi = g * bV[i];
              cpsi = g * fo;
              key = it.seg;
              addr = acc + it.tab + r;
            } else {
              bB[i] = g;
            }
          }
          seg_add(cphi, key, addr);
          seg_add(cpsi, key, addr + AL.psi);
        };
#pragma unroll
        for (int c = 0; c < kCache; ++c) {
          if ((uint32_t)c * kWG < n) {
            const uint32_t i = lane + c * kWG;
            update(cache[c], i, i < n);
          }
        }
        for (uint32_t base = kCache * kWG; base < n; base += kWG) {
          const uint32_t i = base + lane;
          update(i < n ? load_item(B, M, S.item0 + i, i) : cache[0], i, i < n);
        }
        wave_sync();
        Real* t = bA; bA = bB; bB = t;
      }
      // row 1: adjoints of the child tops, ε and q
      for (uint32_t i = lane; i < n; i += kWG) {
        int32_t cf = B.it_cf[S.item0 + i], cg = B.it_cg[S.item0 + i];
        bV[i] = cf >= 0 ? ell[cf] : Real(0);
        bG[i] = cg >= 0 ? ell[cg] : Real(0);
      }
      __syncthreads();
      for (uint32_t base = 0; base < n; base += kWG) {
        const uint32_t i = base + lane;
        Real c0 = 0, c1 = 0, c2 = 0;
        uint32_t key = kNone;
        double* a0 = acc;
        double* a1 = acc;
        double* a2 = acc;
        if (i < n) {
          const uint32_t ia = S.item0 + i;
          const uint32_t si = B.it_seg[ia];
          const SegDev sg = B.seg[si];
          const NodeDev& nd = M.node[sg.node];
          const Real g = bA[i];
          const uint32_t j = i - sg.item0;
          if (nd.kind == WHALE_NODE_SPECIATION) {
            const Real ef = (Real)eps_top(M, nd.child0), eg = (Real)eps_top(M, nd.child1);
            Real sf = 0, sgg = 0;
            for (uint32_t k = B.it_inv[ia]; k < B.it_inv[ia + 1]; ++k) {
              uint32_t ops = B.inv_ops[k];
              Real c = sCoef[B.inv_rec[k] - rb] * bA[ops & 0xffff];
              uint32_t sib = ops >> 16;
              sf += c * bG[sib];
              sgg += c * bV[sib];
            }
            ibar[sg.ibar + j] = sf + g * eg;
            ibar[sg.ibar + sg.nc + j] = sgg + g * ef;
            c0 = g * bV[i];  // -> ε of child1's top
            c1 = g * bG[i];  // -> ε of child0's top
            key = si;
            a0 = acc + AL.eps + nd.child1;
            a1 = acc + AL.eps + nd.child0;
          } else if (nd.kind == WHALE_NODE_WGD) {
            const Real qq = (Real)M.par[3 * M.n + nd.wgd];
            const Real ef = (Real)eps_top(M, nd.child0);
            Real sf = 0;
            for (uint32_t k = B.it_inv[ia]; k < B.it_inv[ia + 1]; ++k) {
              uint32_t ops = B.inv_ops[k];
              sf += sCoef[B.inv_rec[k] - rb] * bA[ops & 0xffff] * bV[ops >> 16];
            }
            ibar[sg.ibar + j] = qq * sf + g * ((Real(1) - qq) + Real(2) * qq * ef);
            Real ret = 0;
            for (uint32_t k = B.it_rec[ia] - rb; k < B.it_rec[ia + 1] - rb; ++k) {
              uint32_t ops = sOps[k];
              ret += sCoef[k] * bV[ops & 0xffff] * bV[ops >> 16];
            }
            c0 = g * (ret - bV[i] + Real(2) * ef * bV[i]);  // -> q
            c1 = g * Real(2) * qq * bV[i];                   // -> ε of child0's top
            key = si;
            a0 = acc + AL.q + nd.wgd;
            a1 = acc + AL.eps + nd.child0;
          }
        }
        seg_add(c0, key, a0);
        seg_add(c1, key, a1);
        (void)c2; (void)a2;
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// tail: one workgroup.  Batch sum, conditioning, table adjoint, chain rule.
// out: [total, dλ(n), dμ(n), dq(n_wgd), dη]
// ---------------------------------------------------------------------------
__global__ void k_tail(ModelDev M, const double* fam_ll, uint32_t n_fam, const double* acc_in, double* ebar,
                       const double* drow, const double* dnode, const int32_t* dlevels, const int32_t* dnodes,
                       int n_dlevels, int cond, int with_grad, double* out) {
  __shared__ double red[256];
  __shared__ double s_total, s_eta;
  const int t = threadIdx.x;
  const AccLayout AL(M.tab_rows, M.n, M.n_wgd);
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
  const double ef = eps_top(M, rn.child0), eg = eps_top(M, rn.child1);
  auto gpgf = [&](double x) { return eta * x / (1.0 - (1.0 - eta) * x); };
  auto dg_ds = [&](double x) { double D = 1.0 - (1.0 - eta) * x; return eta / (D * D); };
  auto dg_deta = [&](double x) { double D = 1.0 - (1.0 - eta) * x; return x * (1.0 - x) / (D * D); };
  double c = 0, dc_er = 0, dc_ef = 0, dc_eg = 0, dc_eta = 0;
  if (cond == WHALE_COND_ROOT) {
    double P = 1.0 - gpgf(ef) - gpgf(eg) + gpgf(er);
    c = P > 0 ? log(P) : -INFINITY;
    dc_ef = -dg_ds(ef) / P;
    dc_eg = -dg_ds(eg) / P;
    dc_er = dg_ds(er) / P;
    dc_eta = (-dg_deta(ef) - dg_deta(eg) + dg_deta(er)) / P;
  } else if (cond == WHALE_COND_NONEXTINCT) {
    double P = 1.0 - gpgf(er);
    c = log(P);
    dc_er = -dg_ds(er) / P;
    dc_eta = -dg_deta(er) / P;
  }
  const double N = (double)n_fam;
  if (t == 0) {
    double tot = red[0] - N * c;
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
    ebar[M.root] += A_ * dA_der + B_ * dB_der - N * dc_er;
    ebar[rn.child0] += -N * dc_ef;
    ebar[rn.child1] += -N * dc_eg;
    s_eta = A_ * dA_deta + B_ * dB_deta - N * dc_eta;
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
      double g = ebar[e];
      double ga = 0, gb = 0;
      for (int r = d.R - 1; r >= 1; --r) {
        const double* G = drow + 9 * (size_t)(d.tab + r);
        const double fp = acc_in[AL.phi + d.tab + r], fs = acc_in[AL.psi + d.tab + r];
        ga += g * G[0] + fp * G[3] + fs * G[6];
        gb += g * G[1] + fp * G[4] + fs * G[7];
        g = g * G[2] + fp * G[5] + fs * G[8];
      }
      dl[e] = d.R > 1 ? ga * dnode[4 * e + 0] + gb * dnode[4 * e + 2] : 0.0;
      dm[e] = d.R > 1 ? ga * dnode[4 * e + 1] + gb * dnode[4 * e + 3] : 0.0;
      // row 1 of ε: product of the children's tops, or the WGD retention map
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
size_t lds_bytes(int mode, uint32_t cap, uint32_t rcap) {
  size_t r = mode == WHALE_FP32_SCALED ? sizeof(float) : sizeof(double);
  return (4 * (size_t)cap + rcap) * r + (size_t)rcap * sizeof(uint32_t);
}

template <typename K>
static hipError_t lds_attr(K kern, size_t lds) {
  static size_t done = 0;
  (void)done;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_forward(int mode, bool store_all, int grid, uint32_t cap, uint32_t rcap, hipStream_t st,
                          const BatchDev& B, const ModelDev& M, const WorkRaw& W, int32_t fam_single,
                          uint32_t fam_begin, uint32_t fam_end, const uint32_t* fam_count) {
  size_t lds = lds_bytes(mode, cap, rcap);
  dim3 g(grid), b(kWG);
#define WHALE_FWD(R, S, A)                                                                              \
  do {                                                                                                  \
    auto k = k_forward<R, S, A>;                                                                        \
    if (lds > 65536) { hipError_t e = lds_attr(k, lds); if (e != hipSuccess) return e; }               \
    hipLaunchKernelGGL(k, g, b, lds, st, B, M, WorkDev<R>(W), cap, rcap, fam_single, fam_begin, fam_end,    \
                       fam_count);                                                                 \
  } while (0)
  if (mode == WHALE_FP64_PLAIN) { if (store_all) WHALE_FWD(double, false, true); else WHALE_FWD(double, false, false); }
  else if (mode == WHALE_FP32_SCALED) { if (store_all) WHALE_FWD(float, true, true); else WHALE_FWD(float, true, false); }
  else { if (store_all) WHALE_FWD(double, true, true); else WHALE_FWD(double, true, false); }
#undef WHALE_FWD
  return hipGetLastError();
}

hipError_t launch_backward(int mode, int grid, uint32_t cap, uint32_t rcap, hipStream_t st, const BatchDev& B,
                           const ModelDev& M, const WorkRaw& W, int32_t fam_single, uint32_t fam_begin,
                           uint32_t fam_end, const uint32_t* fam_count) {
  size_t lds = lds_bytes(mode, cap, rcap);
  dim3 g(grid), b(kWG);
#define WHALE_BWD(R, S)                                                                                 \
  do {                                                                                                  \
    auto k = k_backward<R, S>;                                                                          \
    if (lds > 65536) { hipError_t e = lds_attr(k, lds); if (e != hipSuccess) return e; }               \
    hipLaunchKernelGGL(k, g, b, lds, st, B, M, WorkDev<R>(W), cap, rcap, fam_single, fam_begin, fam_end,    \
                       fam_count);                                                                 \
  } while (0)
  if (mode == WHALE_FP64_PLAIN) WHALE_BWD(double, false);
  else if (mode == WHALE_FP32_SCALED) WHALE_BWD(float, true);
  else WHALE_BWD(double, true);
#undef WHALE_BWD
  return hipGetLastError();
}

hipError_t launch_tables(hipStream_t st, const ModelDev& M, double* eps, double* phi, double* psi, double* drow,
                         double* dnode, const int32_t* hlevels, const int32_t* hnodes, int n_levels) {
  hipLaunchKernelGGL(k_tables, dim3(1), dim3(256), 0, st, M, eps, phi, psi, drow, dnode, hlevels, hnodes, n_levels);
  return hipGetLastError();
}

hipError_t launch_tail(hipStream_t st, const ModelDev& M, const double* fam_ll, uint32_t n_fam, const double* acc,
                       double* ebar, const double* drow, const double* dnode, const int32_t* dlevels,
                       const int32_t* dnodes, int n_dlevels, int cond, int with_grad, double* out) {
  hipLaunchKernelGGL(k_tail, dim3(1), dim3(256), 0, st, M, fam_ll, n_fam, acc, ebar, drow, dnode, dlevels, dnodes,
                     n_dlevels, cond, with_grad, out);
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

}  // namespace whale
