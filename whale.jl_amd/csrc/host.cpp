// ALE ingest and the per-family plan compiler (host C++).
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <unordered_map>
#include <numeric>
#include <thread>

#include "../../include/whale_amd.h"

namespace whale {

// ---------------------------------------------------------------------------
// ALE v0.4 text -> CCD.  Follows parse_aleobserve / parse_body /
// addleafclades! / addubiquitous! and CCD(ale, wm, spmap)
// (src/ccd.jl:102-121, 147-248).
// ---------------------------------------------------------------------------
namespace {

struct Line {
  std::vector<std::string> w;  // whitespace-separated fields
};

bool num(const std::string& s, double& v) {
  char* e = nullptr;
  v = std::strtod(s.c_str(), &e);
  return e && *e == 0 && !s.empty();
}
bool inum(const std::string& s, long& v) {
  char* e = nullptr;
  v = std::strtol(s.c_str(), &e, 10);
  return e && *e == 0 && !s.empty();
}

}  // namespace

std::string parse_ale(const char* text, size_t len, const std::unordered_map<std::string, int>& spmap,
                      CCDHost& out) {
  // split(alestring, "#")[2:end-1]
  std::vector<std::pair<size_t, size_t>> pieces;
  size_t st = 0;
  for (size_t i = 0; i <= len; ++i) {
    if (i == len || text[i] == '#') { pieces.push_back({st, i}); st = i + 1; }
  }
  if (pieces.size() != 10) return "Not a valid .ale file (expected 8 sections)";
  std::map<std::string, std::vector<Line>> sec;
  for (size_t k = 1; k + 1 < pieces.size(); ++k) {
    std::string s(text + pieces[k].first, text + pieces[k].second);
    std::string t;
    t.reserve(s.size());
    for (size_t i = 0; i < s.size(); ++i) {
      if (s[i] == ':' && i + 1 < s.size() && s[i + 1] == '\t') { ++i; continue; }  // ":\t" => ""
      t += s[i];
    }
    size_t nl = t.find('\n');
    std::string header = t.substr(0, nl);
    std::vector<Line> lines;
    size_t pos = nl == std::string::npos ? t.size() : nl + 1;
    while (pos < t.size()) {
      size_t e = t.find('\n', pos);
      if (e == std::string::npos) e = t.size();
      if (e > pos) {
        Line L;
        size_t a = pos;
        while (a < e) {
          while (a < e && isspace((unsigned char)t[a])) ++a;
          size_t b = a;
          while (b < e && !isspace((unsigned char)t[b])) ++b;
          if (b > a) L.w.emplace_back(t.substr(a, b - a));
          a = b;
        }
        lines.push_back(std::move(L));
      }
      pos = e + 1;
    }
    sec[header] = std::move(lines);
  }
  for (const char* h : {"observations", "Bip_counts", "Dip_counts", "leaf-id", "set-id"})
    if (!sec.count(h)) return std::string("ALE section missing: ") + h;
  for (const char* h : {"Bip_counts", "Dip_counts", "leaf-id", "set-id"})
    if (sec[h].size() == 1) return std::string("ALE section with a single line is not a table: ") + h;
  double obs;
  if (sec["observations"].size() != 1 || sec["observations"][0].w.empty() || !num(sec["observations"][0].w[0], obs))
    return "bad #observations";
  std::map<long, double> bip;
  for (auto& L : sec["Bip_counts"]) {
    long k; double v;
    if (L.w.size() < 2 || !inum(L.w[0], k) || !num(L.w[1], v)) return "bad #Bip_counts line";
    bip[k] = v;
  }
  struct D { long a, b; double c; };
  std::map<long, std::vector<D>> dip;
  for (auto& L : sec["Dip_counts"]) {
    long k, a, b; double c;
    if (L.w.size() < 4 || !inum(L.w[0], k) || !inum(L.w[1], a) || !inum(L.w[2], b) || !num(L.w[3], c))
      return "bad #Dip_counts line";
    dip[k].push_back({a, b, c});
  }
  std::map<long, std::string> leafname;  // ALE leaf id -> name
  for (auto& L : sec["leaf-id"]) {
    long id;
    if (L.w.size() < 2 || !inum(L.w[1], id)) return "bad #leaf-id line";
    leafname[id] = L.w[0];
  }
  std::map<long, std::vector<long>> sets;  // sorted by set id, as sort!(OrderedDict(...))
  for (auto& L : sec["set-id"]) {
    long k;
    if (L.w.empty() || !inum(L.w[0], k)) return "bad #set-id line";
    std::vector<long> v;
    for (size_t i = 1; i < L.w.size(); ++i) { long x; if (!inum(L.w[i], x)) return "bad #set-id member"; v.push_back(x); }
    sets[k] = std::move(v);
  }
  // addleafclades!: leaf sets get Bip = observations and no splits; leaf ids
  // inside other sets are replaced by the leaf's set id.
  std::map<long, std::string> leafclade;  // leaf set id -> name
  std::map<long, long> leaf2set;
  for (auto& kv : sets) {
    if (kv.second.size() == 1) {
      auto it = leafname.find(kv.second[0]);
      if (it == leafname.end()) return "set-id names an unknown leaf";
      bip[kv.first] = obs;
      dip[kv.first].clear();
      leafclade[kv.first] = it->second;
      leaf2set[kv.second[0]] = kv.first;
    } else {
      for (auto& x : kv.second) {
        auto it = leaf2set.find(x);
        if (it == leaf2set.end()) return "set-id refers to a leaf before its leaf set";
        x = it->second;
      }
    }
  }
  for (auto& kv : leaf2set) sets[kv.second] = {kv.second};
  // addubiquitous!: the root clade's splits are all complementary pairs (i<j)
  // of the set ids 1..nset, each with count Bip_j.
  long n = (long)leafclade.size();
  long nset = (long)sets.size();
  for (long i = 1; i <= nset; ++i) if (!sets.count(i)) return "set ids are not contiguous";
  long G = nset + 1;
  // Complements are found through an additive hash of the members (a random
  // 64-bit key per leaf): the complement of set i hashes to (sum over all
  // leaves) - h(i).  Every candidate is then checked exactly (sizes add up to
  // the leaves, no common member), and pairs come out in the (i, j) order of
  // the reference's double loop.
  auto key = [](long x) {
    uint64_t z = (uint64_t)x * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  std::vector<uint64_t> h(nset + 1, 0);
  for (long i = 1; i <= nset; ++i)
    for (long x : sets[i]) {
      if (x < 1 || x > nset) return "bad set member";
      h[i] += key(x);
    }
  uint64_t all = 0;
  for (auto& kv : leafclade) all += key(kv.first);
  std::unordered_map<uint64_t, std::vector<long>> by_hash;
  by_hash.reserve((size_t)nset * 2);
  for (long i = 1; i <= nset; ++i) by_hash[h[i]].push_back(i);
  std::vector<char> mark(nset + 1, 0);
  std::vector<D> root;
  double N = 0;
  for (long i = 1; i <= nset; ++i) {
    auto it = by_hash.find(all - h[i]);
    if (it == by_hash.end()) continue;
    const long si = (long)sets[i].size();
    bool marked = false;
    for (long j : it->second) {
      if (j <= i || si + (long)sets[j].size() != n) continue;
      if (!marked) {
        for (long x : sets[i]) mark[x] = 1;
        marked = true;
      }
      bool disjoint = true;
      for (long x : sets[j]) if (mark[x]) { disjoint = false; break; }
      if (!disjoint) continue;
      if (!bip.count(i) || !bip.count(j)) return "missing Bip count";
      if (bip[i] != bip[j]) return "complementary clades have different Bip counts";
      N += bip[i];
      root.push_back({i, j, bip[j]});
    }
    if (marked)
      for (long x : sets[i]) mark[x] = 0;
  }
  if (root.empty()) return "no split of the ubiquitous clade";
  bip[G] = N;
  dip[G] = root;
  {
    std::vector<long> u = sets[root.back().a];
    for (long x : sets[root.back().b]) if (std::find(u.begin(), u.end(), x) == u.end()) u.push_back(x);
    sets[G] = u;
  }
  // CCD: order by (size, set id); p = Dip / Bip
  std::vector<std::pair<long, long>> ord;
  for (auto& kv : sets) ord.push_back({(long)kv.second.size(), kv.first});
  std::sort(ord.begin(), ord.end());
  std::unordered_map<long, int> newid;
  for (size_t i = 0; i < ord.size(); ++i) newid[ord[i].second] = (int)i;
  size_t nc = ord.size();
  out.size.assign(nc, 0);
  out.leaf_species.assign(nc, -1);
  out.split_off.assign(nc + 1, 0);
  out.g1.clear(); out.g2.clear(); out.p.clear();
  for (size_t i = 0; i < nc; ++i) {
    long k = ord[i].second;
    out.size[i] = (int32_t)ord[i].first;
    out.split_off[i] = (int32_t)out.g1.size();
    if (ord[i].first == 1) {
      const std::string& nm = leafclade[k];
      std::string pre = nm.substr(0, nm.find('_'));
      auto it = spmap.find(pre);
      if (it == spmap.end()) return "gene " + nm + ": species " + pre + " is not in the species tree";
      out.leaf_species[i] = it->second;
    }
    auto bi = bip.find(k);
    if (bi == bip.end()) return "clade without a Bip count";
    for (auto& d : dip[k]) {
      auto a = newid.find(d.a), b = newid.find(d.b);
      if (a == newid.end() || b == newid.end()) return "split names an unknown clade";
      out.g1.push_back(a->second);
      out.g2.push_back(b->second);
      out.p.push_back(d.c / bi->second);
    }
    if (ord[i].first > 1 && out.g1.size() == (size_t)out.split_off[i]) return "non-leaf clade without splits";
  }
  out.split_off[nc] = (int32_t)out.g1.size();
  return "";
}

// ---------------------------------------------------------------------------
// plan compiler
// ---------------------------------------------------------------------------
namespace {

struct FamPlan {
  std::string err;
  uint32_t n_clades = 0, max_items = 0, max_recs = 0, root_recs = 0;
  std::vector<StageDev> stage;      // item0 relative to family, seg0 relative
  std::vector<SegDev> seg;          // ell/ibar relative to family
  uint32_t root_seg = 0, root_item0 = 0;
  std::vector<uint32_t> it_seg, it_gid, it_rec;  // it_seg, it_rec family-relative
  std::vector<int32_t> it_cf, it_cg, it_bar;
  std::vector<uint32_t> rec_ops, rec_sid;
  std::vector<uint16_t> rec_own;                              // stage-local item owning each record
  std::vector<uint32_t> lvl;
  uint64_t ell = 0, ibar = 0;
  int64_t cells = 0, split_terms = 0, sum_compat = 0, row_steps = 0;
  double cost = 0;
};

void plan_family(const ModelHost& m, int32_t nc, const int32_t* csize, const int32_t* leaf_sp,
                 const int64_t* soff, const int32_t* g1, const int32_t* g2, FamPlan& F) {
  const int W = m.words;
  F.n_clades = nc;
  if (nc < 3) { F.err = "family with fewer than two genes"; return; }
  if (nc > 65535) { F.err = "family with more than 65535 clades"; return; }
  // species bitsets of clades
  std::vector<uint64_t> cb((size_t)nc * W, 0);
  for (int c = 0; c < nc; ++c) {
    if (c > 0 && csize[c] < csize[c - 1]) { F.err = "clades are not sorted by size"; return; }
    int64_t a = soff[c], b = soff[c + 1];
    if (csize[c] == 1) {
      int s = leaf_sp[c];
      if (s < 0 || s >= m.n_species) { F.err = "leaf clade with a bad species"; return; }
      if (b != a) { F.err = "leaf clade with splits"; return; }
      cb[(size_t)c * W + (s >> 6)] |= 1ull << (s & 63);
      continue;
    }
    if (b <= a) { F.err = "non-leaf clade without splits"; return; }
    for (int64_t k = a; k < b; ++k) {
      int x = g1[k], y = g2[k];
      if (x < 0 || y < 0 || x >= c || y >= c) { F.err = "split part is not an earlier clade"; return; }
      if (csize[x] + csize[y] != csize[c]) { F.err = "split part sizes do not add up"; return; }
    }
    int x = g1[a], y = g2[a];
    for (int w = 0; w < W; ++w) cb[(size_t)c * W + w] = cb[(size_t)x * W + w] | cb[(size_t)y * W + w];
  }
  if (csize[nc - 1] <= csize[nc - 2]) {  // the last clade must be the unique largest (ubiquitous) one
    F.err = "the last clade is not the ubiquitous clade";
    return;
  }
  // compatible clades per node: species(γ) ⊆ clade(e)  (src/model.jl:39 iscompatible)
  std::vector<std::vector<int32_t>> comp(m.n);
  std::vector<std::vector<int32_t>> idx(m.n);
  for (int e = 0; e < m.n; ++e) {
    const uint64_t* eb = &m.bits[(size_t)e * W];
    auto& C = comp[e];
    for (int c = 0; c < nc; ++c) {
      const uint64_t* x = &cb[(size_t)c * W];
      bool sub = true;
      for (int w = 0; w < W; ++w) if (x[w] & ~eb[w]) { sub = false; break; }
      if (sub) C.push_back(c);
    }
    if (e == m.root && (int)C.size() != nc) { F.err = "a clade is not compatible with the root"; return; }
  }
  // segments: every node with compatible clades, root last
  std::vector<int> segof(m.n, -1);
  std::vector<int> segnodes;
  for (int e : m.order) if (!comp[e].empty() && e != m.root) segnodes.push_back(e);
  // stages by height; segments of one stage must not share clades (MUL trees)
  std::vector<std::vector<int>> stages;
  {
    std::vector<int> stamp(nc, -1);
    for (int h = 0; h <= m.max_height; ++h) {
      std::vector<int> pending;
      for (int e : segnodes) if (m.height[e] == h) pending.push_back(e);
      while (!pending.empty()) {
        int sid = (int)stages.size();
        stages.push_back({});
        std::vector<int> rest;
        for (int e : pending) {
          bool clash = false;
          for (int c : comp[e]) if (stamp[c] == sid) { clash = true; break; }
          if (clash) { rest.push_back(e); continue; }
          for (int c : comp[e]) stamp[c] = sid;
          stages.back().push_back(e);
        }
        pending.swap(rest);
      }
    }
  }
  // A stage runs max R rows for all its segments, so the rows of a family are
  // the sum of its stages' longest branches.  Stages by height put long
  // branches of different lineages in different stages; a segment may run in
  // any stage after its children's and before its parent's.  A local search
  // moves segments (longest first) to the stage that lowers the modelled cost
  // Σ (max R − 1 + 3) · (2 + ⌈items/64⌉ + 2⌈records/64⌉) — rows times the
  // 64-lane chunks of one row, record chunks weighing double (they gather two
  // operands, and scatter twice in the reverse sweep), plus a stage's set-up —
  // without raising the family's largest stage (its LDS need) or putting
  // clades twice in a stage.
  if (stages.size() > 1) {
    // the height stages take every kGap-th slot; the slots between are empty
    // at first, so a segment can also run between two height stages
    constexpr int kGap = 2;
    const int K = kGap * ((int)stages.size() - 1) + 1;
    std::vector<int> stg(m.n, -1);
    std::vector<int64_t> nrec_of(m.n, 0);
    std::vector<std::vector<int>> mem(K);
    std::vector<int64_t> ni(K, 0), nr(K, 0);
    for (int h = 0; h < (int)stages.size(); ++h)
      for (int e : stages[h]) {
        const int k = kGap * h;
        stg[e] = k;
        mem[k].push_back(e);
        for (int32_t c : comp[e]) nrec_of[e] += soff[c + 1] - soff[c];
        ni[k] += (int64_t)comp[e].size();
        nr[k] += nrec_of[e];
      }
    int64_t cap_i = 0, cap_r = 0;
    for (int k = 0; k < K; ++k) cap_i = std::max(cap_i, ni[k]), cap_r = std::max(cap_r, nr[k]);
    auto rmax_of = [&](int k, int skip) {
      int r = 0;
      for (int e : mem[k]) if (e != skip) r = std::max(r, m.nodes[e].R);
      return r;
    };
    auto cost = [](int rmax, int64_t items, int64_t recs) -> int64_t {
      if (items == 0) return 0;
      return (int64_t)(rmax - 1 + 3) * (2 + (items + 63) / 64 + 2 * ((recs + 63) / 64));
    };
    std::vector<int> rmx(K);
    for (int k = 0; k < K; ++k) rmx[k] = rmax_of(k, -1);
    std::vector<int> by_r(segnodes);
    std::stable_sort(by_r.begin(), by_r.end(), [&](int a, int b) { return m.nodes[a].R > m.nodes[b].R; });
    std::vector<int> mark(nc, -1);
    for (int pass = 0; pass < 8; ++pass) {
      bool moved = false;
      for (int e : by_r) {
        const NodeDev& nd = m.nodes[e];
        const int a = stg[e];
        int lo = 0, hi = K - 1;
        for (int ch : {nd.child0, nd.child1})
          if (ch >= 0 && stg[ch] >= 0) lo = std::max(lo, stg[ch] + 1);
        if (nd.parent >= 0 && nd.parent != m.root) hi = stg[nd.parent] - 1;
        const int64_t ce = (int64_t)comp[e].size(), re = nrec_of[e];
        const int ra = rmax_of(a, e);
        const int64_t drop = cost(ra, ni[a] - ce, nr[a] - re) - cost(rmx[a], ni[a], nr[a]);
        int best = a;
        int64_t gain = 0;
        for (int b = lo; b <= hi; ++b) {
          if (b == a || ni[b] + ce > cap_i || nr[b] + re > cap_r) continue;
          const int64_t d = drop + cost(std::max(rmx[b], nd.R), ni[b] + ce, nr[b] + re) - cost(rmx[b], ni[b], nr[b]);
          if (d >= gain) continue;
          // no clade twice in one stage (MUL trees)
          for (int32_t c : comp[e]) mark[c] = e;
          bool clash = false;
          for (int x : mem[b]) {
            for (int32_t c : comp[x]) if (mark[c] == e) { clash = true; break; }
            if (clash) break;
          }
          for (int32_t c : comp[e]) mark[c] = -1;
          if (clash) continue;
          gain = d;
          best = b;
        }
        if (best == a) continue;
        mem[a].erase(std::find(mem[a].begin(), mem[a].end(), e));
        mem[best].push_back(e);
        ni[a] -= ce, nr[a] -= re, ni[best] += ce, nr[best] += re;
        rmx[a] = ra;
        rmx[best] = std::max(rmx[best], nd.R);
        stg[e] = best;
        moved = true;
      }
      if (!moved) break;
    }
    // non-empty stages in order, segments in the species tree's order
    std::vector<std::vector<int>> out(K);
    for (int e : segnodes) out[stg[e]].push_back(e);
    stages.clear();
    for (auto& st : out) if (!st.empty()) stages.push_back(std::move(st));
  }
  // segment records, ℓ and ī offsets
  auto nkids = [&](int e) { const NodeDev& d = m.nodes[e]; return (d.child0 >= 0) + (d.child1 >= 0); };
  uint64_t ell = 0, ibar = 0;
  for (auto& st : stages) {
    for (int e : st) {
      SegDev s{};
      s.node = (uint32_t)e;
      s.nc = (uint32_t)comp[e].size();
      s.R = (uint32_t)m.nodes[e].R;
      s.ell = (uint32_t)ell;
      ell += (uint64_t)s.R * s.nc;
      s.ibar = nkids(e) ? (uint32_t)ibar : kNone;
      ibar += (uint64_t)nkids(e) * s.nc;
      segof[e] = (int)F.seg.size();
      F.seg.push_back(s);
    }
  }
  {
    SegDev s{};
    s.node = (uint32_t)m.root;
    s.nc = (uint32_t)nc;
    s.R = 1;
    s.ell = (uint32_t)ell;
    ell += nc;
    s.ibar = (uint32_t)ibar;
    ibar += 2ull * nc;
    segof[m.root] = (int)F.seg.size();
    F.root_seg = (uint32_t)F.seg.size();
    F.seg.push_back(s);
  }
  if (ell > 0xffffffffull || ibar > 0xffffffffull) { F.err = "family too large"; return; }
  F.ell = ell;
  F.ibar = ibar;
  for (int e = 0; e < m.n; ++e) {
    if (segof[e] < 0) continue;
    idx[e].assign(nc, -1);
    for (size_t j = 0; j < comp[e].size(); ++j) idx[e][comp[e][j]] = (int32_t)j;
  }
  // items, records, inverse records
  auto emit_stage = [&](const std::vector<int>& segs, bool is_root) {
    StageDev sd{};
    sd.item0 = (uint32_t)F.it_seg.size();
    sd.seg0 = (uint32_t)segof[segs[0]];
    sd.n_segs = (uint32_t)segs.size();
    uint32_t local = 0;
    for (int e : segs) {
      SegDev& s = F.seg[segof[e]];
      s.item0 = local;
      local += s.nc;
      sd.r_max = std::max(sd.r_max, s.R);
    }
    sd.n_items = local;
    if (local > 0xffff) { F.err = "stage with more than 65535 items"; return sd; }
    if (!is_root) F.max_items = std::max(F.max_items, local);
    for (int e : segs) {
      const SegDev& s = F.seg[segof[e]];
      const NodeDev& nd = m.nodes[e];
      int par = nd.parent;
      int slot = -1;
      if (par >= 0) slot = (m.nodes[par].child0 == e) ? 0 : 1;
      for (uint32_t j = 0; j < s.nc; ++j) {
        int gid = comp[e][j];
        uint32_t me = s.item0 + j;
        F.it_seg.push_back((uint32_t)segof[e]);
        F.it_gid.push_back((uint32_t)gid);
        F.it_rec.push_back((uint32_t)F.rec_ops.size());
        int32_t cfv = -1, cgv = -1;
        int kids[2] = {nd.child0, nd.child1};
        for (int k = 0; k < 2; ++k) {
          int ch = kids[k];
          if (ch < 0 || segof[ch] < 0) continue;
          int x = idx[ch][gid];
          if (x < 0) continue;
          const SegDev& cs = F.seg[segof[ch]];
          int32_t cell = (int32_t)(cs.ell + (uint64_t)(cs.R - 1) * cs.nc + x);
          (k == 0 ? cfv : cgv) = cell;
        }
        F.it_cf.push_back(cfv);
        F.it_cg.push_back(cgv);
        if (is_root) F.it_bar.push_back(-1);
        else {
          const SegDev& ps = F.seg[segof[par]];
          int x = idx[par][gid];
          if (x < 0) { F.err = "internal error: clade not compatible with the parent"; return sd; }
          F.it_bar.push_back((int32_t)(ps.ibar + (uint64_t)slot * ps.nc + x));
        }
        for (int64_t k = soff[gid]; k < soff[gid + 1]; ++k) {
          uint32_t l1 = s.item0 + (uint32_t)idx[e][g1[k]], l2 = s.item0 + (uint32_t)idx[e][g2[k]];
          uint32_t r = (uint32_t)F.rec_ops.size();
          F.rec_ops.push_back(l1 | (l2 << 16));
          F.rec_sid.push_back((uint32_t)(k - soff[0]));
          F.rec_own.push_back((uint16_t)me);
        }
      }
    }
    {
      uint32_t nrec = (uint32_t)F.rec_ops.size() - F.it_rec[sd.item0];
      sd.rec0 = F.it_rec[sd.item0];
      sd.n_recs = nrec;
      if (is_root) F.root_recs = nrec;
      else F.max_recs = std::max(F.max_recs, nrec);
    }
    return sd;
  };
  for (auto& st : stages) {
    StageDev sd = emit_stage(st, false);
    if (!F.err.empty()) return;
    F.stage.push_back(sd);
  }
  F.root_item0 = (uint32_t)F.it_seg.size();
  StageDev rsd = emit_stage({m.root}, true);
  if (!F.err.empty()) return;
  (void)rsd;
  // root levels: runs of clades in size order of which none is a part of
  // another, so a level reads only finished parts.  Clades of one size never
  // are; a run goes on past a size until a clade has a part inside it.
  F.lvl.push_back(0);
  for (int c = 1, start = 0; c < nc; ++c) {
    int32_t dep = -1;
    for (int64_t k = soff[c]; k < soff[c + 1]; ++k) dep = std::max(dep, std::max(g1[k], g2[k]));
    if (dep >= start) {
      F.lvl.push_back((uint32_t)c);
      start = c;
    }
  }
  F.lvl.push_back((uint32_t)nc);
  // workload counts (the reference's cell and split-term counts)
  for (auto& s : F.seg) {
    F.cells += (int64_t)s.R * s.nc;
    F.sum_compat += s.nc;
  }
  for (auto& st : stages) {
    for (int e : st) {
      const SegDev& s = F.seg[segof[e]];
      F.row_steps += s.R;
      int64_t sp = 0;
      for (int32_t c : comp[e]) sp += soff[c + 1] - soff[c];
      F.split_terms += (int64_t)(s.R - 1) * sp + (m.nodes[e].kind == WHALE_NODE_LEAF ? 0 : sp * (m.nodes[e].kind == WHALE_NODE_WGD ? 1 : 2));
    }
  }
  F.split_terms += 2 * (soff[nc] - soff[0]);
  uint64_t steps = 1 + F.lvl.size();
  for (auto& sd : F.stage) steps += sd.r_max;
  F.cost = (double)steps * 64.0 + (double)F.cells + (double)F.split_terms;
}

}  // namespace

std::string build_plan(const ModelHost& m, int n_fam, const int64_t* fam_clade_off, const int32_t* clade_size,
                       const int32_t* leaf_species, const int64_t* split_off, const int32_t* g1,
                       const int32_t* g2, const double* p, PlanHost& P, int nthreads) {
  std::vector<FamPlan> fp(n_fam);
  int nt = std::max(1, std::min(nthreads, n_fam));
  std::vector<std::thread> th;
  for (int w = 0; w < nt; ++w) {
    th.emplace_back([&, w]() {
      for (int f = w; f < n_fam; f += nt) {
        int64_t c0 = fam_clade_off[f], c1 = fam_clade_off[f + 1];
        plan_family(m, (int32_t)(c1 - c0), clade_size + c0, leaf_species + c0, split_off + c0, g1, g2, fp[f]);
        // g1/g2 are absolute arrays indexed by split; plan_family receives soff into them
      }
    });
  }
  for (auto& t : th) t.join();
  for (int f = 0; f < n_fam; ++f)
    if (!fp[f].err.empty()) return "family " + std::to_string(f) + ": " + fp[f].err;
  // concatenate
  P = PlanHost();
  P.cost_order.resize(n_fam);
  std::iota(P.cost_order.begin(), P.cost_order.end(), 0u);
  std::stable_sort(P.cost_order.begin(), P.cost_order.end(),
                   [&](uint32_t a, uint32_t b) { return fp[a].cost > fp[b].cost; });
  P.fam.resize(n_fam);
  P.fam_seg0.resize(n_fam);
  P.fam_nseg.resize(n_fam);
  uint64_t ell = 0, ibar = 0;
  for (int f = 0; f < n_fam; ++f) {
    FamPlan& F = fp[f];
    FamDev& d = P.fam[f];
    uint32_t item_base = (uint32_t)P.it_seg.size();
    uint32_t seg_base = (uint32_t)P.seg.size();
    uint32_t rec_base = (uint32_t)P.rec_ops.size();
    d.n_clades = F.n_clades;
    d.n_stages = (uint32_t)F.stage.size();
    d.stage0 = (uint32_t)P.stage.size();
    d.root_item0 = item_base + F.root_item0;
    d.clade0 = (uint32_t)fam_clade_off[f];
    d.split0 = (uint32_t)split_off[fam_clade_off[f]];
    d.lvl0 = (uint32_t)P.lvl.size();
    d.n_lvl = (uint32_t)F.lvl.size() - 1;
    d.max_items = F.max_items;
    d.max_recs = F.max_recs;
    d.root_recs = F.root_recs;
    d.root_seg = seg_base + F.root_seg;
    d.rec0 = rec_base;
    d.n_recs = (uint32_t)F.rec_ops.size();
    d.ell0 = ell;
    d.ibar0 = ibar;
    P.fam_seg0[f] = seg_base;
    P.fam_nseg[f] = (uint32_t)F.seg.size();
    for (auto s : F.stage) {
      s.item0 += item_base;
      s.seg0 += seg_base;
      s.rec0 += rec_base;
      P.stage.push_back(s);
    }
    for (auto& s : F.seg) P.seg.push_back(s);
    for (size_t i = 0; i < F.it_seg.size(); ++i) {
      P.it_seg.push_back(F.it_seg[i] + seg_base);
      P.it_gid.push_back(F.it_gid[i]);
      P.it_rec.push_back(F.it_rec[i] + rec_base);
      P.it_cf.push_back(F.it_cf[i]);
      P.it_cg.push_back(F.it_cg[i]);
      P.it_bar.push_back(F.it_bar[i]);
    }
    for (size_t r = 0; r < F.rec_ops.size(); ++r) {
      P.rec_ops.push_back(F.rec_ops[r]);
      P.rec_sid.push_back(F.rec_sid[r]);
      P.rec_own.push_back(F.rec_own[r]);
      P.rec_p.push_back(p[d.split0 + F.rec_sid[r]]);
    }
    for (auto v : F.lvl) P.lvl.push_back(v);
    ell += F.ell;
    ibar += F.ibar;
    P.max_items = std::max(P.max_items, std::max(F.max_items, F.n_clades));
    P.max_recs = std::max(P.max_recs, F.max_recs);
    P.max_root_recs = std::max(P.max_root_recs, F.root_recs);
    P.cells += F.cells;
    P.split_terms += F.split_terms;
    P.sum_compat += F.sum_compat;
    P.row_steps += F.row_steps;
    P.n_clades += F.n_clades;
    {  // release the family's temporaries, keep its cost for nothing else
      FamPlan tmp;
      tmp.cost = F.cost;
      std::swap(tmp, F);
    }
  }
  // sentinel: it_rec of one past the last item
  P.it_rec.push_back((uint32_t)P.rec_ops.size());
  P.ell_cells = ell;
  P.ibar_cells = ibar;
  int64_t nclades = fam_clade_off[n_fam];
  P.clade_size.resize(nclades);
  for (int64_t c = 0; c < nclades; ++c) P.clade_size[c] = (uint16_t)std::min(clade_size[c], 65535);
  int64_t nsplit = split_off[nclades];
  P.n_splits = nsplit;
  P.split_p.assign(p, p + nsplit);
  return "";
}

}  // namespace whale
