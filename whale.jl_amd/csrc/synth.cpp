// Seeded synthetic workloads: species trees, DL gene families and their CCDs
// written as ALE v0.4 text.
//
// Semantic template: the reference's forward simulator (src/simulation.jl:
// 14-25 `simulate`, 64-87 `randtree`, 94-133 `dlsim!`) and ALEobserve's
// conditional clade counting, whose output format the reference parses in
// src/ccd.jl:147-248.  Choices that the reference does not fix are marked
// "builder choice".
//
//  * species tree: pure-birth (Yule, rate 1) until n tips, the present one
//    further Exp(n) waiting time after the last birth, rescaled to height 1.
//    Leaves are named S1..Sn.  WGD nodes (unary, named wgd_1..) sit at a
//    uniform position on distinct random non-root branches.
//  * gene families: continuous-time duplication/loss along every branch from
//    one lineage at the root, WGD retention with probability q, pruning of
//    lost lineages.  A family's leaf count is drawn from the workload's size
//    distribution; the family's own rates are shifted towards that size and
//    the simulation is repeated until the count lands in [0.8 m, 1.25 m]
//    (builder choice: it keeps rejection cheap for 4..200 leaves on 50 taxa).
//  * CCD: `ntrees` unrooted trees, each the true gene tree after U{0..3}
//    random nearest-neighbour interchanges; every directed edge contributes a
//    clade and its split, as ALEobserve counts them.  Families whose CCD
//    exceeds the clade/split caps are redrawn.
//
// Every family is generated from its own stream (seed, family index), so the
// output does not depend on the number of threads.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

// SplitMix64 seeding + xoshiro256** stream
struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) { for (auto& v : s) v = splitmix(seed); }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double unif() { return (next() >> 11) * 0x1.0p-53; }
  double expo(double rate) { return -std::log1p(-unif()) / rate; }
  double normal() {
    double u1 = unif(), u2 = unif();
    return std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(2 * M_PI * u2);
  }
  int below(int n) { return (int)(unif() * n); }
};

uint64_t fam_seed(uint64_t seed, uint64_t i) {
  uint64_t x = seed ^ (0xD1B54A32D192ED03ULL * (i + 1));
  return Rng::splitmix(x);
}

// ------------------------------------------------------------------ species tree
struct SNode {
  int parent = -1;
  std::vector<int> ch;
  double len = 0;   // branch length above
  std::string name;
};

std::string to_newick(const std::vector<SNode>& t, int x) {
  char buf[64];
  std::string s;
  if (!t[x].ch.empty()) {
    s += "(";
    for (size_t i = 0; i < t[x].ch.size(); ++i) { if (i) s += ","; s += to_newick(t, t[x].ch[i]); }
    s += ")";
  }
  s += t[x].name;
  if (t[x].parent >= 0) { snprintf(buf, sizeof buf, ":%.17g", t[x].len); s += buf; }
  return s;
}

std::vector<SNode> yule(Rng& rng, int n, int nwgd) {
  std::vector<SNode> t(1);
  std::vector<double> birth(1, 0.0);  // time at which each node starts
  std::vector<int> alive;
  // root splits at time 0
  for (int c = 0; c < 2; ++c) { t.push_back(SNode()); t.back().parent = 0; t[0].ch.push_back((int)t.size() - 1); birth.push_back(0); alive.push_back((int)t.size() - 1); }
  double now = 0;
  while ((int)alive.size() < n) {
    now += rng.expo((double)alive.size());
    int k = rng.below((int)alive.size());
    int x = alive[k];
    t[x].len = now - birth[x];
    alive.erase(alive.begin() + k);
    for (int c = 0; c < 2; ++c) { t.push_back(SNode()); t.back().parent = x; t[x].ch.push_back((int)t.size() - 1); birth.push_back(now); alive.push_back((int)t.size() - 1); }
  }
  now += rng.expo((double)n);
  for (int x : alive) t[x].len = now - birth[x];
  for (auto& nd : t) nd.len /= now;
  int li = 0;
  // name leaves S1..Sn in left-to-right order
  std::vector<int> leaves;
  struct W { std::vector<SNode>& t; std::vector<int>& out; void go(int x) { if (t[x].ch.empty()) out.push_back(x); for (int c : t[x].ch) go(c); } } w{t, leaves};
  w.go(0);
  for (int x : leaves) t[x].name = "S" + std::to_string(++li);
  // WGDs on distinct random non-root branches, uniform position
  std::vector<int> cand;
  for (int x = 1; x < (int)t.size(); ++x) cand.push_back(x);
  for (int w2 = 0; w2 < nwgd && !cand.empty(); ++w2) {
    int k = rng.below((int)cand.size());
    int x = cand[k];
    cand.erase(cand.begin() + k);
    double u = rng.unif();
    int p = t[x].parent;
    SNode m;
    m.parent = p;
    m.name = "wgd_" + std::to_string(w2 + 1);
    m.len = t[x].len * (1 - u);
    int mi = (int)t.size();
    t.push_back(m);
    for (auto& c : t[p].ch) if (c == x) c = mi;
    t[mi].ch.push_back(x);
    t[x].parent = mi;
    t[x].len = t[x].len * u;
  }
  return t;
}

// ------------------------------------------------------------------ gene tree simulation
struct GNode {
  int parent = -1;
  int ch[2] = {-1, -1};
  int species = -1;  // leaf species index (1-based S index)
  bool dead = false;
};

struct Sim {
  const std::vector<SNode>& sp;
  std::vector<int> sp_index;  // tree node -> S index for leaves
  Rng& rng;
  double lam, mu, q;
  std::vector<GNode> g;
  size_t cap;
  bool overflow = false;
  int newnode(int parent) { g.push_back(GNode()); g.back().parent = parent; return (int)g.size() - 1; }
  void attach(int par, int c) { if (par < 0) return; if (g[par].ch[0] < 0) g[par].ch[0] = c; else g[par].ch[1] = c; }
  // lineage `x` (a gene node) enters the top of species branch s
  void branch(int s, int x) {
    if (overflow) return;
    double t = sp[s].len;
    // event-driven per lineage
    std::vector<int> out;
    std::vector<std::pair<int, double>> stack{{x, 0.0}};
    while (!stack.empty()) {
      auto [y, t0] = stack.back();
      stack.pop_back();
      double tt = t0 + rng.expo(lam + mu);
      if (tt >= t) { out.push_back(y); continue; }
      if (rng.unif() < lam / (lam + mu)) {
        int a = newnode(y), b = newnode(y);
        g[y].ch[0] = a; g[y].ch[1] = b;
        stack.push_back({a, tt}); stack.push_back({b, tt});
        if (g.size() > cap) { overflow = true; return; }
      } else {
        g[y].dead = true;
      }
    }
    for (int y : out) node(s, y);
  }
  // lineage y arrives at the bottom of species branch s (at node s)
  void node(int s, int y) {
    if (overflow) return;
    const SNode& n = sp[s];
    if (n.ch.empty()) { g[y].species = sp_index[s]; return; }
    if (n.ch.size() == 1) {  // WGD
      if (rng.unif() < q) {
        int a = newnode(y), b = newnode(y);
        g[y].ch[0] = a; g[y].ch[1] = b;
        branch(n.ch[0], a);
        branch(n.ch[0], b);
      } else {
        branch(n.ch[0], y);
      }
      return;
    }
    int a = newnode(y), b = newnode(y);
    g[y].ch[0] = a; g[y].ch[1] = b;
    branch(n.ch[0], a);
    branch(n.ch[1], b);
  }
};

// rooted binary gene tree after pruning: adjacency with leaf species
struct GTree {
  std::vector<std::array<int, 3>> adj;  // up to 3 neighbours (unrooted), -1 unused
  std::vector<int> species;             // leaf species (1-based S index), -1 internal
  int nleaves = 0;
};

// prune dead lineages and suppress unary nodes; returns an unrooted tree
bool prune(const std::vector<GNode>& g, GTree& out) {
  // postorder compute: a node's pruned representative
  std::vector<int> rep(g.size(), -2);
  std::vector<std::array<int, 2>> kids;
  std::vector<int> spc;
  std::vector<int> order;
  {
    std::vector<std::pair<int, int>> st{{0, 0}};
    while (!st.empty()) {
      int x = st.back().first, k = st.back().second;
      int nch = (g[x].ch[0] >= 0) + (g[x].ch[1] >= 0);
      if (k < nch) { st.back().second = k + 1; st.push_back({g[x].ch[k], 0}); continue; }
      order.push_back(x);
      st.pop_back();
    }
  }
  for (int x : order) {
    if (g[x].ch[0] < 0) {
      if (g[x].dead || g[x].species < 0) { rep[x] = -1; continue; }
      rep[x] = (int)kids.size(); kids.push_back({-1, -1}); spc.push_back(g[x].species);
      continue;
    }
    int a = rep[g[x].ch[0]], b = g[x].ch[1] >= 0 ? rep[g[x].ch[1]] : -1;
    if (a < 0 && b < 0) rep[x] = -1;
    else if (a < 0) rep[x] = b;
    else if (b < 0) rep[x] = a;
    else { rep[x] = (int)kids.size(); kids.push_back({a, b}); spc.push_back(-1); }
  }
  int root = rep[0];
  if (root < 0) return false;
  int nl = 0;
  for (int s : spc) nl += s >= 0;
  out.nleaves = nl;
  if (nl < 2) return true;
  // unrooted adjacency: drop the root (degree 2), joining its two children
  size_t N = kids.size();
  out.adj.assign(N, {-1, -1, -1});
  out.species = spc;
  auto link = [&](int a, int b) {
    for (int k = 0; k < 3; ++k) if (out.adj[a][k] < 0) { out.adj[a][k] = b; break; }
    for (int k = 0; k < 3; ++k) if (out.adj[b][k] < 0) { out.adj[b][k] = a; break; }
  };
  for (size_t x = 0; x < N; ++x) {
    if ((int)x == root) continue;
    for (int k = 0; k < 2; ++k) if (kids[x][k] >= 0) link((int)x, kids[x][k]);
  }
  if (kids[root][0] >= 0) link(kids[root][0], kids[root][1]);
  out.adj[root] = {-1, -1, -1};
  out.species[root] = -2;  // removed
  return true;
}

// ------------------------------------------------------------------ clade counting
using Bits = std::array<uint64_t, 4>;  // up to 256 leaves
struct BitsHash {
  size_t operator()(const Bits& b) const {
    uint64_t h = 0x9E3779B97F4A7C15ULL;
    for (auto w : b) { h ^= w + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2); }
    return (size_t)h;
  }
};

struct CladeRec {
  double bip = 0;
  std::map<std::pair<int, int>, double> dips;
  int size = 0;
  int first_leaf = -1;
};

struct Counter {
  std::unordered_map<Bits, int, BitsHash> idx;
  std::vector<Bits> bits;
  std::vector<CladeRec> recs;
  int get(const Bits& b) {
    auto it = idx.find(b);
    if (it != idx.end()) return it->second;
    int k = (int)recs.size();
    idx.emplace(b, k);
    bits.push_back(b);
    CladeRec r;
    for (auto w : b) r.size += __builtin_popcountll(w);
    recs.push_back(r);
    return k;
  }
};

// count the clades and splits of one unrooted tree (leaf slot per node given)
void count_tree(const GTree& t, const std::vector<int>& leafslot, Counter& cnt, const Bits& full) {
  size_t N = t.adj.size();
  int start = -1;
  for (size_t x = 0; x < N; ++x) if (t.species[x] >= 0) { start = (int)x; break; }
  // root at a leaf: compute subtree bitsets below each directed edge parent->child
  std::vector<int> par(N, -1), order;
  std::vector<int> st{start};
  par[start] = start;
  while (!st.empty()) {
    int x = st.back(); st.pop_back();
    order.push_back(x);
    for (int k = 0; k < 3; ++k) { int y = t.adj[x][k]; if (y >= 0 && par[y] < 0) { par[y] = x; st.push_back(y); } }
  }
  std::vector<Bits> down(N, Bits{0, 0, 0, 0});
  for (auto it = order.rbegin(); it != order.rend(); ++it) {
    int x = *it;
    if (t.species[x] >= 0) { int s = leafslot[x]; down[x][s >> 6] |= 1ULL << (s & 63); }
    if (x != start) for (int w = 0; w < 4; ++w) down[par[x]][w] |= down[x][w];
  }
  // clade of directed edge: child side (x below par[x]) = down[x]; parent side = full ^ down[x]
  auto comp = [&](const Bits& b) { Bits r; for (int w = 0; w < 4; ++w) r[w] = full[w] ^ b[w]; return r; };
  for (int x : order) {
    if (x == start) continue;
    // clade down[x], split into x's children (if internal)
    int cd = cnt.get(down[x]);
    cnt.recs[cd].bip += 1;
    if (t.species[x] < 0) {
      int kids[2], nk = 0;
      for (int k = 0; k < 3; ++k) { int y = t.adj[x][k]; if (y >= 0 && y != par[x]) kids[nk++] = y; }
      int a = cnt.get(down[kids[0]]), b = cnt.get(down[kids[1]]);
      cnt.recs[cd].dips[{std::min(a, b), std::max(a, b)}] += 1;
    } else {
      cnt.recs[cd].first_leaf = x;
    }
    // complement clade (side containing par[x]); split into par[x]'s other neighbours
    Bits up = comp(down[x]);
    int cu = cnt.get(up);
    cnt.recs[cu].bip += 1;
    int p = par[x];
    if (t.species[p] < 0) {
      int others[2], no = 0;
      for (int k = 0; k < 3; ++k) { int y = t.adj[p][k]; if (y >= 0 && y != x) others[no++] = y; }
      Bits cl[2];
      for (int k = 0; k < 2; ++k) cl[k] = (others[k] == par[p] && p != start) ? comp(down[p]) : down[others[k]];
      int a = cnt.get(cl[0]), b = cnt.get(cl[1]);
      cnt.recs[cu].dips[{std::min(a, b), std::max(a, b)}] += 1;
    } else {
      cnt.recs[cu].first_leaf = p;
    }
  }
}

bool nni(GTree& t, Rng& rng) {
  // internal edges (u,v) with both endpoints internal
  std::vector<std::pair<int, int>> edges;
  for (size_t x = 0; x < t.adj.size(); ++x) {
    if (t.species[x] != -1) continue;
    for (int k = 0; k < 3; ++k) { int y = t.adj[x][k]; if (y > (int)x && t.species[y] == -1) edges.push_back({(int)x, y}); }
  }
  if (edges.empty()) return false;
  auto [u, v] = edges[rng.below((int)edges.size())];
  int ou[2], nu = 0, ov[2], nv = 0;
  for (int k = 0; k < 3; ++k) { if (t.adj[u][k] != v) ou[nu++] = t.adj[u][k]; if (t.adj[v][k] != u) ov[nv++] = t.adj[v][k]; }
  int b = ou[rng.below(2)], c = ov[rng.below(2)];
  auto repl = [&](int x, int from, int to) { for (int k = 0; k < 3; ++k) if (t.adj[x][k] == from) { t.adj[x][k] = to; return; } };
  repl(u, b, c); repl(v, c, b); repl(b, u, v); repl(c, v, u);
  return true;
}

struct Family {
  std::string ale;
  int nleaves = 0, nclades = 0, nsplits = 0;
};

std::string gene_newick(const GTree& t, const std::vector<std::string>& names) {
  // unrooted: write as a trifurcation (or bifurcation) around an internal node
  int start = -1;
  for (size_t x = 0; x < t.adj.size(); ++x) if (t.species[x] == -1) { start = (int)x; break; }
  if (start < 0) {
    std::string s = "(";
    bool first = true;
    for (size_t x = 0; x < t.adj.size(); ++x) if (t.species[x] >= 0) { if (!first) s += ","; s += names[x]; first = false; }
    return s + ");";
  }
  struct W {
    const GTree& t; const std::vector<std::string>& nm;
    std::string go(int x, int from) {
      if (t.species[x] >= 0) return nm[x];
      std::string s = "(";
      bool first = true;
      for (int k = 0; k < 3; ++k) { int y = t.adj[x][k]; if (y < 0 || y == from) continue; if (!first) s += ","; s += go(y, x); first = false; }
      return s + ")";
    }
  } w{t, names};
  return w.go(start, -1) + ";";
}

bool make_family(const std::vector<SNode>& sp, const std::vector<int>& sp_index, int nspecies, uint64_t seed,
                 int target, int size_lo, int size_hi, double lam, double mu, double q, int ntrees,
                 int max_clades, int max_splits, Family& out) {
  Rng rng(seed);
  double r = std::log((double)target / nspecies);
  double lf = lam + std::max(0.0, r), mf = mu + std::max(0.0, -r);
  int lo = std::max({3, size_lo, (int)std::floor(0.8 * target)});
  int hi = std::max(lo, std::min(size_hi, (int)std::ceil(1.25 * target)));
  GTree gt;
  bool ok = false;
  for (int attempt = 0; attempt < 20000 && !ok; ++attempt) {
    if (attempt > 0 && attempt % 2000 == 0) { lo = std::max({3, size_lo, lo - 1}); hi = std::min(size_hi, hi + 1); }
    Sim s{sp, sp_index, rng, lf, mf, q, {}, (size_t)(40 * hi + 1000)};
    s.newnode(-1);
    s.node(0, 0);
    if (s.overflow) continue;
    GTree t;
    if (!prune(s.g, t)) continue;
    if (t.nleaves < lo || t.nleaves > hi || t.nleaves < 3 || t.nleaves > 256) continue;
    gt = t;
    ok = true;
  }
  if (!ok) return false;
  // leaf names and slots; ALE leaf ids are 1..m in slot order
  int m = gt.nleaves;
  std::vector<int> slot(gt.adj.size(), -1);
  std::vector<std::string> names(gt.adj.size());
  std::vector<int> per_species(nspecies + 1, 0);
  int k = 0;
  for (size_t x = 0; x < gt.adj.size(); ++x) if (gt.species[x] >= 0) {
    slot[x] = k++;
    names[x] = "S" + std::to_string(gt.species[x]) + "_g" + std::to_string(++per_species[gt.species[x]]);
  }
  Bits full{0, 0, 0, 0};
  for (int s = 0; s < m; ++s) full[s >> 6] |= 1ULL << (s & 63);
  for (int redraw = 0; redraw < 50; ++redraw) {
    Counter cnt;
    for (int i = 0; i < ntrees; ++i) {
      GTree t = gt;
      int nn = rng.below(4);
      for (int j = 0; j < nn; ++j) nni(t, rng);
      count_tree(t, slot, cnt, full);
    }
    int nclades = (int)cnt.recs.size() + 1;
    int nsplits = 0;
    for (auto& r2 : cnt.recs) nsplits += (int)r2.dips.size();
    nsplits += (int)cnt.recs.size() / 2;  // root splits: one per bipartition
    if (nclades > max_clades || nsplits > max_splits) { ntrees = std::max(1, ntrees * 3 / 4); continue; }
    // set ids: leaf clades get a permutation of 1..m, other clades m+1.. in discovery order
    std::vector<int> perm(m);
    std::iota(perm.begin(), perm.end(), 1);
    for (int i = m - 1; i > 0; --i) std::swap(perm[i], perm[rng.below(i + 1)]);
    std::vector<int> setid(cnt.recs.size(), 0);
    int nxt = m + 1;
    for (size_t c = 0; c < cnt.recs.size(); ++c) {
      if (cnt.recs[c].size == 1) setid[c] = perm[slot[cnt.recs[c].first_leaf]];
      else setid[c] = nxt++;
    }
    std::string s;
    char buf[128];
    s += "#constructor_string\n" + gene_newick(gt, names) + "\n";
    s += "#observations\n" + std::to_string(ntrees) + "\n";
    s += "#Bip_counts\n";
    std::vector<int> bysid(cnt.recs.size());
    std::iota(bysid.begin(), bysid.end(), 0);
    std::sort(bysid.begin(), bysid.end(), [&](int a, int b) { return setid[a] < setid[b]; });
    for (int c : bysid) if (cnt.recs[c].size > 1) { snprintf(buf, sizeof buf, "%d\t%.0f\n", setid[c], cnt.recs[c].bip); s += buf; }
    s += "#Bip_bls\n";
    for (int c : bysid) { snprintf(buf, sizeof buf, "%d\t1\n", setid[c]); s += buf; }
    s += "#Dip_counts\n";
    for (int c : bysid) for (auto& kv : cnt.recs[c].dips) {
      snprintf(buf, sizeof buf, "%d\t%d\t%d\t%.0f\n", setid[c], setid[kv.first.first], setid[kv.first.second], kv.second);
      s += buf;
    }
    s += "#last_leafset_id\n" + std::to_string(cnt.recs.size()) + "\n";
    s += "#leaf-id\n";
    for (size_t x = 0; x < gt.adj.size(); ++x) if (slot[x] >= 0) { s += names[x] + "\t" + std::to_string(slot[x] + 1) + "\n"; }
    s += "#set-id\n";
    for (int c : bysid) {
      s += std::to_string(setid[c]) + "\t:";
      const Bits& b = cnt.bits[c];
      for (int l = 0; l < m; ++l) if ((b[l >> 6] >> (l & 63)) & 1) s += "\t" + std::to_string(l + 1);
      s += "\n";
    }
    s += "#END\n";
    out.ale = std::move(s);
    out.nleaves = m;
    out.nclades = nclades;
    out.nsplits = nsplits;
    return true;
  }
  return false;
}

struct FamSet {
  std::vector<Family> fams;
  std::string err;
};

}  // namespace

extern "C" {

// Species tree for a workload, as Newick (root without a length).  Returns the
// string length, or -1 if `cap` is too small.
int ws_species_tree(uint64_t seed, int ntaxa, int nwgd, char* out, int cap) {
  Rng rng(Rng::splitmix(seed));
  auto t = yule(rng, ntaxa, nwgd);
  std::string s = to_newick(t, 0) + ";";
  if ((int)s.size() + 1 > cap) return -1;
  memcpy(out, s.c_str(), s.size() + 1);
  return (int)s.size();
}

// size_kind: 0 uniform integer on [lo, hi]; 1 log-normal (median, sigma) truncated to [lo, hi]
void* ws_families(const char* newick, uint64_t seed, int nfam, int size_kind, double median, double sigma,
                  int lo, int hi, int ntrees, int max_clades, int max_splits, double lam, double mu, double q,
                  int nthreads) {
  // parse the species tree (the Newick written by ws_species_tree)
  std::vector<SNode> sp;
  {
    std::string s(newick);
    size_t i = 0;
    struct P {
      std::string& s; size_t& i; std::vector<SNode>& t;
      int go(int parent) {
        int id = (int)t.size(); t.push_back(SNode()); t[id].parent = parent;
        if (s[i] == '(') {
          ++i;
          while (true) { int c = go(id); t[id].ch.push_back(c); if (s[i] == ',') { ++i; continue; } ++i; break; }
        }
        std::string nm;
        while (i < s.size() && !strchr(",():;", s[i])) nm += s[i++];
        t[id].name = nm;
        if (s[i] == ':') { ++i; size_t j = i; while (!strchr(",();", s[j])) ++j; t[id].len = strtod(s.substr(i, j - i).c_str(), nullptr); i = j; }
        return id;
      }
    } p{s, i, sp};
    p.go(-1);
  }
  std::vector<int> sp_index(sp.size(), -1);
  int nspecies = 0;
  for (size_t x = 0; x < sp.size(); ++x) if (sp[x].ch.empty()) { sp_index[x] = std::stoi(sp[x].name.substr(1)); ++nspecies; }
  auto* fs = new FamSet;
  fs->fams.resize(nfam);
  std::vector<int> target(nfam);
  for (int i = 0; i < nfam; ++i) {
    Rng rng(fam_seed(seed ^ 0x5bd1e995ULL, i));
    int m;
    if (size_kind == 0) m = lo + rng.below(hi - lo + 1);
    else {
      double v;
      do { v = median * std::exp(sigma * rng.normal()); } while (v < lo - 0.5 || v > hi + 0.5);
      m = (int)std::lround(v);
    }
    target[i] = std::min(hi, std::max(lo, m));
  }
  std::vector<int> failed(nfam, 0);
  int nt = std::max(1, nthreads);
  std::vector<std::thread> th;
  for (int w = 0; w < nt; ++w) {
    th.emplace_back([&, w]() {
      for (int i = w; i < nfam; i += nt) {
        if (!make_family(sp, sp_index, nspecies, fam_seed(seed, i), target[i], lo, hi, lam, mu, q, ntrees,
                         max_clades, max_splits, fs->fams[i]))
          failed[i] = 1;
      }
    });
  }
  for (auto& t : th) t.join();
  for (int i = 0; i < nfam; ++i) if (failed[i]) { fs->err = "family " + std::to_string(i) + " could not be generated"; break; }
  return fs;
}

const char* ws_error(void* h) { return ((FamSet*)h)->err.c_str(); }
int ws_count(void* h) { return (int)((FamSet*)h)->fams.size(); }
long ws_text_len(void* h, int i) { return (long)((FamSet*)h)->fams[i].ale.size(); }
const char* ws_text(void* h, int i) { return ((FamSet*)h)->fams[i].ale.c_str(); }
int ws_stats(void* h, int i, int* nleaves, int* nclades, int* nsplits) {
  auto& f = ((FamSet*)h)->fams[i];
  *nleaves = f.nleaves; *nclades = f.nclades; *nsplits = f.nsplits;
  return 0;
}
void ws_free(void* h) { delete (FamSet*)h; }

}  // extern "C"
