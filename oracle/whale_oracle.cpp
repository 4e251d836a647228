// CPU oracle for the amalgamated-likelihood reconciliation DP of Whale.jl.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain CPU restatement of the
// reference algorithm.  It is used by tests/, by __graft_entry__.smoke() and by
// bench.py's cpu_baseline leg, for reference values and the timed CPU
// baseline -- never by the product library, which must not link, load or call
// anything under oracle/.
//
// Everything the path needs is restated here independently of the product:
//   * a Newick reader and NewickTree-style insertnode! (halfway placement)
//   * WhaleModel construction: node order, ids, WGD ids, MUL groups, slices
//       reference: src/model.jl:16-22, 32-42, 96-145
//   * rate parameterisations DLWGD / ConstantDLWGD (getθ, getp)
//       reference: src/rmodels.jl:14, 31-33, 55-66
//   * slice tables setmodel!/setnode!/setwgdnode!/setslices!, getα, _ϵ, _ϕ, _ψ
//       reference: src/model.jl:162-191, src/bdputil.jl:3-11
//   * the ALE v0.4 parser and CCD construction
//       reference: src/ccd.jl:13-139, 147-248
//   * the DP whale!/whalewgd!/within_branch!/whaleroot! and the Π terms
//       reference: src/core.jl:29-64, 83-199
//   * conditioning (RootCondition, NonExtinctCondition, NoCondition, geompgf)
//       reference: src/condition.jl:1-29, src/bdputil.jl:67
//
// The scalar type is a template parameter: double (bit-faithful to the
// reference's Float64 path: same operation order, no FMA contraction),
// long double (x87 extended: 64-bit mantissa, 15-bit exponent; the
// underflow-free ground truth for large families) and Dual<N> (forward-mode
// dual numbers with ForwardDiff's semantics, used for the reference-style
// chunked gradient).
//
// Known deviations from the reference, all below 1e-15 relative:
//   * the batch sum over families is a sequential loop; Julia's `sum` uses a
//     @simd (reassociating) loop, so the last bits of the batch value may
//     differ.
//
// Each function is pinned by tests/test_oracle.py: the reference's golden
// log-likelihoods, closed forms of the birth-death pgf and its derivatives,
// exact small cases, and central differences for the dual gradient.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <functional>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------
template <int N>
struct Dual {
  double v;
  double d[N];
  Dual() : v(0) { for (int i = 0; i < N; ++i) d[i] = 0; }
  Dual(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0; }  // NOLINT
};
template <int N> inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> inline Dual<N> operator-(const Dual<N>& a) { Dual<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r; r.v = a.v / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
  return r;
}
template <int N> inline Dual<N> exp(const Dual<N>& a) { Dual<N> r; r.v = std::exp(a.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * r.v; return r; }
template <int N> inline Dual<N> log(const Dual<N>& a) { Dual<N> r; r.v = std::log(a.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / a.v; return r; }

inline double val(double x) { return x; }
inline long double val(long double x) { return x; }
template <int N> inline double val(const Dual<N>& x) { return x.v; }
using std::exp;
using std::log;

// ---------------------------------------------------------------------------
// Newick (subset used by the reference's fixtures) and NewickTree-like edits
// ---------------------------------------------------------------------------
struct TNode {
  std::string name;
  double dist = NAN;
  int parent = -1;
  std::vector<int> ch;
};
struct Tree {
  std::vector<TNode> n;
  int root = -1;
};

struct NwParser {
  const std::string& s;
  size_t i = 0;
  Tree& t;
  NwParser(const std::string& s_, Tree& t_) : s(s_), t(t_) {}
  void ws() { while (i < s.size() && isspace((unsigned char)s[i])) ++i; }
  int node(int parent) {
    ws();
    int id = (int)t.n.size();
    t.n.push_back(TNode());
    t.n[id].parent = parent;
    if (i < s.size() && s[i] == '(') {
      ++i;
      while (true) {
        int c = node(id);
        t.n[id].ch.push_back(c);
        ws();
        if (i < s.size() && s[i] == ',') { ++i; continue; }
        if (i < s.size() && s[i] == ')') { ++i; break; }
        throw std::runtime_error("newick: expected , or )");
      }
    }
    ws();
    std::string nm;
    while (i < s.size() && !strchr(",():;", s[i]) && !isspace((unsigned char)s[i])) nm += s[i++];
    t.n[id].name = nm;
    ws();
    if (i < s.size() && s[i] == ':') {
      ++i; ws();
      size_t j = i;
      while (j < s.size() && !strchr(",();", s[j]) && !isspace((unsigned char)s[j])) ++j;
      t.n[id].dist = std::strtod(s.substr(i, j - i).c_str(), nullptr);
      i = j;
    }
    return id;
  }
};

Tree parse_newick(const std::string& s) {
  Tree t;
  NwParser p(s, t);
  t.root = p.node(-1);
  p.ws();
  if (p.i >= s.size() || s[p.i] != ';') throw std::runtime_error("newick: missing ;");
  return t;
}

void leaves_of(const Tree& t, int x, std::vector<int>& out) {
  if (t.n[x].ch.empty()) { out.push_back(x); return; }
  for (int c : t.n[x].ch) leaves_of(t, c, out);
}

int find_leaf(const Tree& t, const std::string& name) {
  std::vector<int> L;
  leaves_of(t, t.root, L);
  for (int x : L) if (t.n[x].name == name) return x;
  throw std::runtime_error("no leaf named " + name);
}

int lca(const Tree& t, int a, int b) {
  std::set<int> anc;
  for (int x = a; x >= 0; x = t.n[x].parent) anc.insert(x);
  for (int x = b; x >= 0; x = t.n[x].parent) if (anc.count(x)) return x;
  throw std::runtime_error("no lca");
}

// insertnode!(n, name=...) with the default dist = distance(n)/2: a new node is
// placed halfway along the branch above n and takes n's slot among its
// parent's children.
void insert_above(Tree& t, int x, const std::string& name) {
  int p = t.n[x].parent;
  if (p < 0) throw std::runtime_error("cannot insert above root");
  double d = t.n[x].dist;
  int m = (int)t.n.size();
  t.n.push_back(TNode());
  t.n[m].name = name;
  t.n[m].dist = d / 2;
  t.n[m].parent = p;
  t.n[m].ch.push_back(x);
  for (auto& c : t.n[p].ch) if (c == x) c = m;
  t.n[x].parent = m;
  t.n[x].dist = d - d / 2;
}

// ---------------------------------------------------------------------------
// WhaleModel structure (reference: src/model.jl:16-22, 96-145)
// ---------------------------------------------------------------------------
struct MNode {
  int id = 0;                 // 1-based reference id
  int parent = 0;             // parent id (0 for root)
  std::vector<int> children;  // child ids in tree order
  std::string name;
  bool wgd = false, leaf = false, root = false;
  int wgdid = 0;              // 1-based, preorder numbering of WGD nodes
  double leafP = 0;           // leafℙ
  int k = 0;                  // number of slices; rows = k + 1
  double delta = 0;           // slice length t/k (0 if k == 0)
  std::vector<int> clade;     // sorted species ids (setclade!)
};

struct Model {
  std::vector<MNode> nodes;   // indexed by id-1
  std::vector<int> order;     // ids in wm.order
  int n_nonwgd = 0, n_wgd = 0;
  int cond = 0;               // 0 RootCondition, 1 NonExtinct, 2 NoCondition
  std::map<std::string, int> spmap;  // leaf name -> species (leaf) id
  int root_id() const { return order.back(); }
  const MNode& nd(int id) const { return nodes[id - 1]; }
};

static bool starts_with_wgd(const std::string& s) { return s.compare(0, 3, "wgd") == 0; }

Model build_model(const Tree& t, double dt, int minn, int maxn, int cond) {
  Model m;
  m.cond = cond;
  std::vector<int> tl;
  leaves_of(t, t.root, tl);
  std::map<std::string, int> counts;
  for (int x : tl) counts[t.n[x].name]++;
  std::map<std::string, int> mulgroups;
  for (auto& kv : counts) if (kv.second > 1) mulgroups[kv.first] = kv.second;

  // walk: preorder creation (WGD ids), postorder push
  std::vector<int> post;          // tree indices in postorder
  std::vector<int> wgdid(t.n.size(), 0);
  int wc = 0;
  {
    struct W {
      const Tree& t; int& wc; std::vector<int>& wgdid; std::vector<int>& post;
      void go(int x) {
        if (starts_with_wgd(t.n[x].name)) { ++wc; wgdid[x] = wc; }
        for (int c : t.n[x].ch) go(c);
        post.push_back(x);
      }
    } w{t, wc, wgdid, post};
    w.go(t.root);
  }
  // order = union(getleaves(root), postorder)
  std::vector<int> ord = tl;
  std::set<int> seen(tl.begin(), tl.end());
  for (int x : post) if (!seen.count(x)) { ord.push_back(x); seen.insert(x); }
  int nonwgd = 0;
  for (int x : post) if (!starts_with_wgd(t.n[x].name)) ++nonwgd;
  std::vector<int> idof(t.n.size(), 0);
  int i = nonwgd + 1, j = 1;
  for (int x : ord) {
    if (starts_with_wgd(t.n[x].name)) idof[x] = i++;
    else idof[x] = j++;
  }
  m.n_nonwgd = nonwgd;
  m.n_wgd = wc;
  m.nodes.resize(ord.size());
  std::map<std::string, int> mulid;
  for (int x : ord) {
    MNode& n = m.nodes[idof[x] - 1];
    n.id = idof[x];
    n.parent = t.n[x].parent >= 0 ? idof[t.n[x].parent] : 0;
    for (int c : t.n[x].ch) n.children.push_back(idof[c]);
    n.name = t.n[x].name;
    n.wgd = starts_with_wgd(n.name);
    n.leaf = t.n[x].ch.empty();
    n.root = (x == t.root);
    n.wgdid = wgdid[x];
    auto mg = mulgroups.find(n.name);
    n.leafP = mg != mulgroups.end() ? 1.0 / mg->second : (n.leaf ? 1.0 : 0.0);
    double tt = t.n[x].dist;
    if (std::isnan(tt)) n.k = 0;
    else n.k = std::min(maxn, std::max(minn, (int)std::ceil(tt / dt)));
    n.delta = n.k == 0 ? 0.0 : tt / n.k;
    if (!n.wgd && mg != mulgroups.end()) mulid[n.name] = n.id;
    if (!n.leaf && !n.wgd && n.children.size() != 2)
      throw std::runtime_error("speciation node " + n.name + " must have two children");
    if (n.wgd && n.children.size() != 1) throw std::runtime_error("WGD node " + n.name + " must have one child");
    m.order.push_back(n.id);
  }
  // setclade!
  struct C {
    Model& m; std::map<std::string, int>& mulid;
    std::vector<int> go(int id) {
      MNode& n = m.nodes[id - 1];
      std::set<int> s;
      if (n.leaf) {
        auto it = mulid.find(n.name);
        s.insert(it != mulid.end() ? it->second : n.id);
      } else {
        for (int c : n.children) { auto v = go(c); s.insert(v.begin(), v.end()); }
      }
      n.clade.assign(s.begin(), s.end());
      return n.clade;
    }
  } c{m, mulid};
  c.go(m.root_id());
  // spmap: name(l) => id(l) over getleaves(root(wm)); later leaves win
  for (int x : tl) m.spmap[t.n[x].name] = idof[x];
  return m;
}

// ---------------------------------------------------------------------------
// Parameters (reference: src/rmodels.jl)
// ---------------------------------------------------------------------------
template <typename T>
struct Theta {
  T lam, mu, q, eta, p;
};

// rates_kind 0: DLWGD (log-scale λ[id], μ[id]); 1: ConstantDLWGD (natural).
template <typename T>
struct Rates {
  int kind = 0;
  std::vector<T> lam, mu, q;  // DLWGD: per id (log scale); Constant: size 1
  T eta;
  std::vector<double> p;      // sampling probabilities per leaf id (not differentiated)
  Theta<T> get(const Model& m, const MNode& n) const {
    Theta<T> th;
    T nan = T(NAN);
    th.q = nan; th.p = T(0.0); th.eta = eta;
    double pv = (!p.empty() && n.leaf) ? p[n.id - 1] : 0.0;
    if (kind == 1) {
      th.lam = lam[0]; th.mu = mu[0];
      th.p = T(pv);
      if (n.wgd) th.q = q[n.wgdid - 1];
      return th;
    }
    if (n.wgd) {
      const MNode* c = &n;
      while (c->wgd) c = &m.nd(c->children[0]);  // nonwgdchild
      th.lam = exp(lam[c->id - 1]);
      th.mu = exp(mu[c->id - 1]);
      th.q = q[n.wgdid - 1];
      return th;
    }
    if (n.id > (int)lam.size()) { th.lam = nan; th.mu = nan; }
    else { th.lam = exp(lam[n.id - 1]); th.mu = exp(mu[n.id - 1]); }
    th.p = T(pv);
    return th;
  }
};

// ---------------------------------------------------------------------------
// slice tables (reference: src/model.jl:162-191, src/bdputil.jl:6-11)
// ---------------------------------------------------------------------------
template <typename T>
struct Tables {
  // per node id-1: rows (k+1) of eps, phi, psi
  std::vector<std::vector<T>> eps, phi, psi;
};

template <typename T>
T get_alpha(const T& l, const T& u, double t) {
  // isapprox(λ, μ, atol=1e-6): rtol is 0 when atol > 0
  double lv = (double)val(l), uv = (double)val(u);
  bool close = (lv == uv) || (std::isfinite(lv) && std::isfinite(uv) && std::fabs(lv - uv) <= 1e-6);
  T tt = T(t);
  T one = T(1.0);
  if (close) return l * tt / (one + l * tt);
  T e = exp(tt * (l - u));
  return u * (e - one) / (l * e - u);
}

template <typename T>
void set_slices(std::vector<T>& eps, std::vector<T>& phi, std::vector<T>& psi,
                const T& l, const T& u, double delta) {
  T one = T(1.0);
  for (size_t i = 1; i < eps.size(); ++i) {
    T a = get_alpha(l, u, delta);
    T b = (l / u) * a;
    T e = eps[i - 1];
    T den = one - b * e;
    eps[i] = (a + (one - a - b) * e) / den;
    phi[i] = (one - a) * (one - b) / (den * den);
    psi[i] = (one - a) * (one - b) * b / (den * den * den);
  }
}

template <typename T>
Tables<T> build_tables(const Model& m, const Rates<T>& r) {
  Tables<T> tb;
  size_t nn = m.nodes.size();
  tb.eps.resize(nn); tb.phi.resize(nn); tb.psi.resize(nn);
  for (int id : m.order) {
    const MNode& n = m.nd(id);
    int R = n.k + 1;
    auto& E = tb.eps[id - 1]; auto& F = tb.phi[id - 1]; auto& P = tb.psi[id - 1];
    E.assign(R, T(1.0)); F.assign(R, T(1.0)); P.assign(R, T(1.0));
    Theta<T> th = r.get(m, n);
    if (n.wgd) {
      T e = tb.eps[n.children[0] - 1].back();
      E[0] = th.q * (e * e) + (T(1.0) - th.q) * e;
    } else {
      if (n.leaf) E[0] = th.p;
      else {
        T pr = tb.eps[n.children[0] - 1].back();
        for (size_t c = 1; c < n.children.size(); ++c) pr = pr * tb.eps[n.children[c] - 1].back();
        E[0] = pr;
      }
      F[0] = T(1.0);
    }
    set_slices(E, F, P, th.lam, th.mu, n.delta);
  }
  return tb;
}

// ---------------------------------------------------------------------------
// ALE parser and CCD (reference: src/ccd.jl)
// ---------------------------------------------------------------------------
struct Split { int g1, g2; double p; };
struct Clade {
  int id;                  // 1-based new id
  int size;
  std::vector<Split> splits;
  std::vector<int> species;  // sorted species ids
};
struct CCD {
  std::vector<Clade> clades;             // by id
  std::vector<std::vector<int>> compat;  // per node id-1: clade ids (increasing)
  std::vector<std::vector<int>> index;   // per clade id-1: per node id-1 -> column (1-based) or 0
};

static std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string w;
  while (is >> w) out.push_back(w);
  return out;
}

static std::vector<std::string> split_char(const std::string& s, char c) {
  std::vector<std::string> out;
  size_t st = 0;
  for (size_t i = 0; i <= s.size(); ++i) {
    if (i == s.size() || s[i] == c) { out.push_back(s.substr(st, i - st)); st = i + 1; }
  }
  return out;
}

static long parse_int(const std::string& s) {
  char* e;
  long v = std::strtol(s.c_str(), &e, 10);
  if (*e) throw std::runtime_error("ale: expected integer, got '" + s + "'");
  return v;
}

static double parse_num(const std::string& s) {
  char* e;
  double v = std::strtod(s.c_str(), &e);
  if (*e) throw std::runtime_error("ale: expected number, got '" + s + "'");
  return v;
}

CCD parse_ale(const std::string& text, const Model& m) {
  // sections = split(alestring, "#")[2:end-1]
  std::vector<std::string> secs = split_char(text, '#');
  if (secs.size() < 2) throw std::runtime_error("Not a valid .ale file");
  std::vector<std::string> sections(secs.begin() + 1, secs.end() - 1);
  if (sections.size() != 8) throw std::runtime_error("Not a valid .ale file");
  double observations = 0;
  std::map<long, double> bip;                           // Bip_counts
  struct DipRow { long g1, g2; double c; };
  std::map<long, std::vector<DipRow>> dip;              // Dip_counts, in file order
  std::map<long, std::string> leaf_id;                  // ALE leaf id -> name
  std::map<long, std::vector<long>> set_id;
  for (auto sec : sections) {
    // replace(section, ":\t" => "")
    std::string s2;
    for (size_t i = 0; i < sec.size(); ++i) {
      if (sec[i] == ':' && i + 1 < sec.size() && sec[i + 1] == '\t') { ++i; continue; }
      s2 += sec[i];
    }
    std::vector<std::string> lines = split_char(s2, '\n');
    std::string header = lines[0];
    std::vector<std::string> xs;
    for (size_t i = 1; i < lines.size(); ++i) if (!lines[i].empty()) xs.push_back(lines[i]);
    if (header == "observations") {
      if (xs.size() != 1) throw std::runtime_error("ale: observations");
      observations = parse_num(split_ws(xs[0])[0]);
    } else if (header == "Bip_counts") {
      if (xs.size() == 1) throw std::runtime_error("ale: single-line Bip_counts is not a table");
      for (auto& x : xs) { auto w = split_ws(x); bip[parse_int(w[0])] = parse_num(w[1]); }
    } else if (header == "Dip_counts") {
      if (xs.size() == 1) throw std::runtime_error("ale: single-line Dip_counts is not a table");
      for (auto& x : xs) {
        auto w = split_ws(x);
        dip[parse_int(w[0])].push_back({parse_int(w[1]), parse_int(w[2]), parse_num(w[3])});
      }
    } else if (header == "leaf-id") {
      if (xs.size() == 1) throw std::runtime_error("ale: single-line leaf-id is not a table");
      for (auto& x : xs) { auto w = split_ws(x); leaf_id[parse_int(w[1])] = w[0]; }
    } else if (header == "set-id") {
      if (xs.size() == 1) throw std::runtime_error("ale: single-line set-id is not a table");
      for (auto& x : xs) {
        auto w = split_ws(x);
        std::vector<long> v;
        for (size_t i = 1; i < w.size(); ++i) v.push_back(parse_int(w[i]));
        set_id[parse_int(w[0])] = v;
      }
    }
  }
  // addleafclades! (sorted by set id)
  std::map<long, std::string> leafclades;
  std::map<long, long> themap;
  for (auto& kv : set_id) {
    long k = kv.first;
    auto& v = kv.second;
    if (v.size() == 1) {
      bip[k] = observations;
      dip[k].clear();
      leafclades[k] = leaf_id.at(v[0]);
      themap[v[0]] = k;
    } else {
      std::vector<long> nv;
      for (long x : v) nv.push_back(themap.at(x));
      v = nv;
    }
  }
  for (auto& kv : themap) set_id[kv.second] = {kv.second};
  // addubiquitous!
  long n = (long)leafclades.size();
  long nset = (long)set_id.size();
  long G = nset + 1;
  dip[G].clear();
  double N = 0;
  for (long i = 1; i <= nset; ++i) {
    for (long j = i + 1; j <= nset; ++j) {
      auto& li = set_id.at(i);
      auto& lj = set_id.at(j);
      if ((long)(li.size() + lj.size()) != n) continue;
      std::set<long> si(li.begin(), li.end());
      bool disj = true;
      for (long x : lj) if (si.count(x)) { disj = false; break; }
      if (disj) {
        if (bip.at(i) != bip.at(j)) throw std::runtime_error("ale: complementary Bip counts differ");
        N += bip.at(i);
        dip[G].push_back({i, j, bip.at(j)});
      }
    }
  }
  if (dip[G].empty()) throw std::runtime_error("ale: no root splits");
  bip[G] = N;
  {
    auto tr = dip[G].back();
    std::vector<long> u = set_id.at(tr.g1);
    for (long x : set_id.at(tr.g2)) if (std::find(u.begin(), u.end(), x) == u.end()) u.push_back(x);
    set_id[G] = u;
  }
  // CCD(ale, wm, spmap): sort by (size, set id)
  std::vector<std::pair<long, long>> ord;
  for (auto& kv : set_id) ord.push_back({(long)kv.second.size(), kv.first});
  std::sort(ord.begin(), ord.end());
  std::map<long, int> idmap;
  for (size_t i = 0; i < ord.size(); ++i) idmap[ord[i].second] = (int)i + 1;
  CCD c;
  for (size_t i = 0; i < ord.size(); ++i) {
    long k = ord[i].second;
    Clade cl;
    cl.id = (int)i + 1;
    cl.size = (int)set_id[k].size();
    double v = bip.at(k);
    for (auto& t3 : dip[k]) cl.splits.push_back({idmap.at(t3.g1), idmap.at(t3.g2), t3.c / v});
    std::set<int> sp;
    for (long lid : set_id[k]) {
      const std::string& nm = leafclades.at(lid);
      std::string pre = nm.substr(0, nm.find('_'));
      auto it = m.spmap.find(pre);
      if (it == m.spmap.end()) throw std::runtime_error("ale: unknown species " + pre);
      sp.insert(it->second);
    }
    cl.species.assign(sp.begin(), sp.end());
    c.clades.push_back(cl);
  }
  // index_and_getℓ
  size_t nn = m.nodes.size();
  c.compat.assign(nn, {});
  c.index.assign(c.clades.size(), std::vector<int>(nn, 0));
  for (int id : m.order) {
    const MNode& nd = m.nd(id);
    int i2 = 1;
    for (auto& g : c.clades) {
      if (!std::includes(nd.clade.begin(), nd.clade.end(), g.species.begin(), g.species.end())) continue;
      c.compat[id - 1].push_back(g.id);
      c.index[g.id - 1][id - 1] = i2++;
    }
  }
  return c;
}

// ---------------------------------------------------------------------------
// DP (reference: src/core.jl)
// ---------------------------------------------------------------------------
template <typename T>
struct Ell {
  // per node id-1: column-major (R x C)
  std::vector<std::vector<T>> a;
  std::vector<int> R;
};

template <typename T>
struct DP {
  const Model& m;
  const CCD& x;
  const Tables<T>& tb;
  const Rates<T>& rates;
  Ell<T> l;
  DP(const Model& m_, const CCD& x_, const Tables<T>& tb_, const Rates<T>& r_) : m(m_), x(x_), tb(tb_), rates(r_) {
    size_t nn = m.nodes.size();
    l.a.resize(nn); l.R.resize(nn);
    for (size_t e = 0; e < nn; ++e) {
      l.R[e] = m.nodes[e].k + 1;
      l.a[e].assign((size_t)l.R[e] * x.compat[e].size(), T(0.0));
    }
  }
  T& at(int e, int row, int col) { return l.a[e - 1][(size_t)(col - 1) * l.R[e - 1] + (row - 1)]; }
  T getl(int e, int g, int row) {
    int i = x.index[g - 1][e - 1];
    return i == 0 ? T(0.0) : at(e, row, i);
  }
  T getl(int e, int g) {
    int i = x.index[g - 1][e - 1];
    return i == 0 ? T(0.0) : at(e, l.R[e - 1], i);
  }
  const T& eps_top(int e) const { return tb.eps[e - 1].back(); }
  static bool isleafc(const Clade& g) { return g.size == 1; }

  T Pspec(const Clade& g, int f, int gg) {
    T p = T(0.0);
    for (auto& t : g.splits) p = p + T(t.p) * (getl(f, t.g1) * getl(gg, t.g2) + getl(gg, t.g1) * getl(f, t.g2));
    return p;
  }
  T Ploss(const Clade& g, int f, int gg) { return getl(f, g.id) * eps_top(gg) + getl(gg, g.id) * eps_top(f); }
  T Pdup(const Clade& g, int e, int i) {
    T p = T(0.0);
    for (auto& t : g.splits) p = p + T(t.p) * getl(e, t.g1, i - 1) * getl(e, t.g2, i - 1);
    return tb.psi[e - 1][i - 1] * p;
  }
  void within_branch(const MNode& n, const Clade& g, int e, int j, bool leaf) {
    int R = n.k + 1;
    for (int i = 2; i <= R; ++i) {
      at(e, i, j) = at(e, i, j) + tb.phi[e - 1][i - 1] * at(e, i - 1, j);
      if (!leaf) at(e, i, j) = at(e, i, j) + Pdup(g, e, i);
    }
  }
  void whale(const MNode& n) {
    if (n.wgd) return whalewgd(n);
    if (n.root) return whaleroot(n);
    int e = n.id;
    for (int c : x.compat[e - 1]) {
      const Clade& g = x.clades[c - 1];
      int j = x.index[c - 1][e - 1];
      bool leaf = isleafc(g);
      if (leaf && n.leaf) {
        at(e, 1, j) = T(n.leafP);
      } else if (!n.leaf) {
        T p = Pspec(g, n.children[0], n.children[1]) + Ploss(g, n.children[0], n.children[1]);
        at(e, 1, j) = at(e, 1, j) + p;
      }
      within_branch(n, g, e, j, leaf);
    }
  }
  void whalewgd(const MNode& n) {
    int e = n.id;
    T q = rates.get(m, n).q;
    int f = n.children[0];
    for (int c : x.compat[e - 1]) {
      const Clade& g = x.clades[c - 1];
      int j = x.index[c - 1][e - 1];
      T p = T(0.0);
      bool leaf = isleafc(g);
      if (!leaf) {
        T r = T(0.0);
        for (auto& t : g.splits) r = r + T(t.p) * getl(f, t.g1) * getl(f, t.g2);
        p = p + r * q;
      }
      p = p + ((T(1.0) - q) * getl(f, g.id) + (T(2.0) * q) * eps_top(f) * getl(f, g.id));
      at(e, 1, j) = p;
      within_branch(n, g, e, j, leaf);
    }
  }
  void whaleroot(const MNode& n) {
    T eta = rates.get(m, n).eta;
    int e = n.id;
    T eps = eps_top(e);
    T one = T(1.0);
    T xi = one - (one - eta) * eps;
    int f = n.children[0], gg = n.children[1];
    for (auto& g : x.clades) {
      bool leaf = isleafc(g);
      T a = T(0.0), b = T(0.0);
      T c = Ploss(g, f, gg);
      if (!leaf) {
        T pr = T(0.0);
        for (auto& t : g.splits) pr = pr + T(t.p) * getl(e, t.g1, 1) * getl(e, t.g2, 1);
        a = a + pr;
        b = b + Pspec(g, f, gg);
      }
      at(e, 1, g.id) = (one - eta) * xi * a / eta + (eta * (one - eps) / (xi * xi)) * (b + c);
    }
  }
  T run() {
    for (int id : m.order) whale(m.nd(id));
    T L = at(m.root_id(), 1, (int)x.clades.size());
    if (val(L) > 0) return log(L);
    return T(-INFINITY);
  }
};

template <typename T>
T geompgf(const T& p, const T& s) { return p * s / (T(1.0) - (T(1.0) - p) * s); }

template <typename T>
T condition(const Model& m, const Tables<T>& tb, const Rates<T>& r) {
  const MNode& root = m.nd(m.root_id());
  T eta = r.get(m, root).eta;
  if (m.cond == 2) return T(0.0);
  T er = geompgf(eta, tb.eps[root.id - 1].back());
  if (m.cond == 1) return log(T(1.0) - er);
  T ef = geompgf(eta, tb.eps[root.children[0] - 1].back());
  T eg = geompgf(eta, tb.eps[root.children[1] - 1].back());
  T p = T(1.0) - ef - eg + er;
  return val(p) > 0 ? log(p) : T(-INFINITY);
}

template <typename T>
Rates<T> make_rates(int kind, const T* theta, int nl, int nq, const double* p, int np) {
  Rates<T> r;
  r.kind = kind;
  if (kind == 1) {
    r.lam = {theta[0]}; r.mu = {theta[1]};
    for (int i = 0; i < nq; ++i) r.q.push_back(theta[2 + i]);
    r.eta = theta[2 + nq];
  } else {
    for (int i = 0; i < nl; ++i) r.lam.push_back(theta[i]);
    for (int i = 0; i < nl; ++i) r.mu.push_back(theta[nl + i]);
    for (int i = 0; i < nq; ++i) r.q.push_back(theta[2 * nl + i]);
    r.eta = theta[2 * nl + nq];
  }
  for (int i = 0; i < np; ++i) r.p.push_back(p[i]);
  return r;
}

// logpdf over a batch: per-family unconditioned values and the conditioned sum
template <typename T>
T batch_logpdf(const Model& m, const std::vector<CCD>& ccds, const Rates<T>& r, std::vector<T>* fam, int nthreads) {
  Tables<T> tb = build_tables(m, r);
  std::vector<T> ls(ccds.size());
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
  for (long i = 0; i < (long)ccds.size(); ++i) {
    DP<T> dp(m, ccds[i], tb, r);
    ls[i] = dp.run();
  }
  T s = T(0.0);
  for (auto& v : ls) s = s + v;
  T c = condition(m, tb, r);
  T tot = s - T((double)ccds.size()) * c;
  if (fam) *fam = ls;
  if (!std::isfinite((double)val(tot))) return T(-INFINITY);
  return tot;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C API (ctypes) -- used by tests, smoke() and the CPU baseline only
// ---------------------------------------------------------------------------
using namespace orc;

static thread_local std::string g_err;

struct OrcModel { Model m; };
struct OrcCCDs { std::vector<CCD> v; };

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// inserts: ";"-separated entries "name=leafA,leafB" applied in order
// (insertnode!(getlca(t, leafA, leafB), name=name)).  cond: 0 Root, 1 NonExtinct, 2 None.
void* orc_model_new(const char* newick, const char* inserts, double dt, int minn, int maxn, int cond) {
  try {
    Tree t = parse_newick(newick);
    if (inserts && *inserts) {
      for (auto& ent : split_char(inserts, ';')) {
        if (ent.empty()) continue;
        auto eq = ent.find('=');
        std::string name = ent.substr(0, eq);
        auto ab = split_char(ent.substr(eq + 1), ',');
        int a = find_leaf(t, ab[0]), b = find_leaf(t, ab[1]);
        insert_above(t, lca(t, a, b), name);
      }
    }
    auto* om = new OrcModel{build_model(t, dt, minn, maxn, cond)};
    return om;
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

void orc_model_free(void* p) { delete (OrcModel*)p; }

int orc_model_counts(void* p, int* n_nodes, int* n_nonwgd, int* n_wgd, int* n_leaves) {
  auto* om = (OrcModel*)p;
  *n_nodes = (int)om->m.nodes.size();
  *n_nonwgd = om->m.n_nonwgd;
  *n_wgd = om->m.n_wgd;
  int nl = 0;
  for (auto& n : om->m.nodes) nl += n.leaf;
  *n_leaves = nl;
  return 0;
}

// per node (by id-1): parent id, wgdid, k, leafP, is leaf, is wgd, is root, name
int orc_model_node(void* p, int id, int* parent, int* wgdid, int* k, double* delta, double* leafP, int* flags, char* name, int name_cap) {
  auto* om = (OrcModel*)p;
  const MNode& n = om->m.nd(id);
  *parent = n.parent; *wgdid = n.wgdid; *k = n.k; *delta = n.delta; *leafP = n.leafP;
  *flags = (n.leaf ? 1 : 0) | (n.wgd ? 2 : 0) | (n.root ? 4 : 0);
  snprintf(name, name_cap, "%s", n.name.c_str());
  return 0;
}

int orc_model_order(void* p, int* order) {
  auto* om = (OrcModel*)p;
  for (size_t i = 0; i < om->m.order.size(); ++i) order[i] = om->m.order[i];
  return 0;
}

void* orc_ccds_new(void* mp, const char** texts, int n) {
  try {
    auto* om = (OrcModel*)mp;
    auto* c = new OrcCCDs;
    c->v.resize(n);
    for (int i = 0; i < n; ++i) c->v[i] = parse_ale(texts[i], om->m);
    return c;
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

void orc_ccds_free(void* p) { delete (OrcCCDs*)p; }

int orc_ccd_nclades(void* p, int i) { return (int)((OrcCCDs*)p)->v[i].clades.size(); }

// flattened clades of family i: size, first species id (leaf clades), splits CSR
int orc_ccd_nsplits(void* p, int i) {
  int s = 0;
  for (auto& c : ((OrcCCDs*)p)->v[i].clades) s += (int)c.splits.size();
  return s;
}

int orc_ccd_export(void* p, int i, int* size, int* leaf_species, int* split_off, int* g1, int* g2, double* pr) {
  auto& x = ((OrcCCDs*)p)->v[i];
  int s = 0;
  for (size_t c = 0; c < x.clades.size(); ++c) {
    auto& cl = x.clades[c];
    size[c] = cl.size;
    leaf_species[c] = cl.size == 1 ? cl.species[0] : 0;
    split_off[c] = s;
    for (auto& t : cl.splits) { g1[s] = t.g1; g2[s] = t.g2; pr[s] = t.p; ++s; }
  }
  split_off[x.clades.size()] = s;
  return 0;
}

// prec: 0 double, 1 long double.  Returns the conditioned batch value; writes the
// per-family (unconditioned) values into fam if non-null, and the conditioning
// constant into cond_out if non-null.
double orc_logpdf(void* mp, void* cp, int kind, const double* theta, int nl, int nq,
                  const double* p, int np, int prec, int nthreads, double* fam, double* cond_out) {
  try {
    auto& m = ((OrcModel*)mp)->m;
    auto& cc = ((OrcCCDs*)cp)->v;
    int nt = kind == 1 ? 3 + nq : 2 * nl + nq + 1;
    if (prec == 1) {
      std::vector<long double> th(theta, theta + nt);
      auto r = make_rates<long double>(kind, th.data(), nl, nq, p, np);
      std::vector<long double> f;
      long double tot = batch_logpdf(m, cc, r, &f, nthreads);
      if (fam) for (size_t i = 0; i < f.size(); ++i) fam[i] = (double)f[i];
      if (cond_out) { auto tb = build_tables(m, r); *cond_out = (double)condition(m, tb, r); }
      return (double)tot;
    }
    auto r = make_rates<double>(kind, theta, nl, nq, p, np);
    std::vector<double> f;
    double tot = batch_logpdf(m, cc, r, &f, nthreads);
    if (fam) for (size_t i = 0; i < f.size(); ++i) fam[i] = f[i];
    if (cond_out) { auto tb = build_tables(m, r); *cond_out = condition(m, tb, r); }
    return tot;
  } catch (std::exception& e) { g_err = e.what(); return NAN; }
}

// ForwardDiff-style gradient: chunks of 12 partials (ForwardDiff's default
// chunk size is min(12, length(x))), one dual pass of the whole batch per chunk.
int orc_grad(void* mp, void* cp, int kind, const double* theta, int nl, int nq,
             const double* p, int np, int nthreads, double* value, double* grad) {
  try {
    constexpr int CH = 12;
    using D = Dual<CH>;
    auto& m = ((OrcModel*)mp)->m;
    auto& cc = ((OrcCCDs*)cp)->v;
    int nt = kind == 1 ? 3 + nq : 2 * nl + nq + 1;
    for (int c0 = 0; c0 < nt; c0 += CH) {
      std::vector<D> th(nt);
      for (int i = 0; i < nt; ++i) {
        th[i] = D(theta[i]);
        if (i >= c0 && i < c0 + CH) th[i].d[i - c0] = 1.0;
      }
      auto r = make_rates<D>(kind, th.data(), nl, nq, p, np);
      D tot = batch_logpdf<D>(m, cc, r, nullptr, nthreads);
      if (value) *value = tot.v;
      for (int i = c0; i < std::min(nt, c0 + CH); ++i) grad[i] = tot.d[i - c0];
    }
    return 0;
  } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// slice tables for inspection (double): rows of node id written to eps/phi/psi
int orc_tables(void* mp, int kind, const double* theta, int nl, int nq, const double* p, int np,
               int id, double* eps, double* phi, double* psi) {
  auto& m = ((OrcModel*)mp)->m;
  auto r = make_rates<double>(kind, theta, nl, nq, p, np);
  auto tb = build_tables(m, r);
  for (size_t i = 0; i < tb.eps[id - 1].size(); ++i) {
    eps[i] = tb.eps[id - 1][i]; phi[i] = tb.phi[id - 1][i]; psi[i] = tb.psi[id - 1][i];
  }
  return (int)tb.eps[id - 1].size();
}

// one-slice quantities getslice(λ, μ, t, ϵ) (reference: src/model.jl:194-198)
void orc_getslice(double l, double u, double t, double e, double* out3) {
  double a = get_alpha(l, u, t);
  double b = (l / u) * a;
  double den = 1.0 - b * e;
  out3[0] = (a + (1.0 - a - b) * e) / den;
  out3[1] = (1.0 - a) * (1.0 - b) / (den * den);
  out3[2] = (1.0 - a) * (1.0 - b) * b / (den * den * den);
}

// full ℓ matrices of one family in the reference layout (per node id: R x C,
// column-major), double precision.  Returns the number of values written.
long orc_ell(void* mp, void* cp, int fam, int kind, const double* theta, int nl, int nq,
             const double* p, int np, double* out, long cap) {
  auto& m = ((OrcModel*)mp)->m;
  auto& x = ((OrcCCDs*)cp)->v[fam];
  auto r = make_rates<double>(kind, theta, nl, nq, p, np);
  auto tb = build_tables(m, r);
  DP<double> dp(m, x, tb, r);
  dp.run();
  long k = 0;
  for (size_t e = 0; e < m.nodes.size(); ++e)
    for (double v : dp.l.a[e]) { if (k < cap) out[k] = v; ++k; }
  return k;
}

// Stochastic backtracking of one family (reference: src/track.jl:188-414)
// from the double ℓ of the oracle's own DP: n_samples reconciled trees, each
// as its nodes in creation order, 4 ints per node (0-based clade or -1 for a
// loss, 0-based species node, slice row (1-based, as SliceState.t), parent
// node or -1), at most `cap` nodes per sample (counts[s] = -1 past that).
// The uniform draws come from a counter-based generator of (seed, family,
// sample) and are drawn in the reference's order (depth first, one draw per
// inter- or intra-branch step), so a device sampler that follows the same
// order draws the same numbers.
static double bt_uniform(uint64_t& st) {
  st += 0x9E3779B97F4A7C15ull;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

int orc_backtrack(void* mp, void* cp, int fam, int kind, const double* theta, int nl, int nq, const double* p,
                  int np, uint64_t seed, int n_samples, int cap, int* counts, int* nodes) {
  try {
    auto& m = ((OrcModel*)mp)->m;
    auto& x = ((OrcCCDs*)cp)->v[fam];
    auto r = make_rates<double>(kind, theta, nl, nq, p, np);
    auto tb = build_tables(m, r);
    DP<double> dp(m, x, tb, r);
    dp.run();
    const int root = m.root_id();
    const int gam = (int)x.clades.size();
    for (int smp = 0; smp < n_samples; ++smp) {
      uint64_t st = seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(fam + 1)) ^ (0xD1B54A32D192ED03ull * (uint64_t)(smp + 1));
      int* out = nodes + (size_t)smp * cap * 4;
      int n = 0;
      bool over = false;
      auto add = [&](int g, int e, int t, int par) -> int {
        if (n >= cap) { over = true; return -1; }
        int* o = out + 4 * n;
        o[0] = g - 1; o[1] = e - 1; o[2] = t; o[3] = par;
        return n++;
      };
      struct St { int e, g, t; };
      // bt(node, state): the recursion of backtrack!(b) with b.node = node
      std::function<void(int, St)> bt;
      auto go = [&](int node, St s) {  // (b::BackTracker)(newstate), then backtrack!
        if (over) return;
        const int* nd = out + 4 * node;
        if (s.g - 1 != nd[0] || s.e - 1 != nd[1]) {
          int c = add(s.g, s.e, s.t, node);
          if (c < 0) return;
          bt(c, s);
        } else {
          bt(node, s);
        }
      };
      bt = [&](int node, St s) {
        if (over) return;
        if (s.g == 0) return;  // loss
        const MNode& n_ = m.nd(s.e);
        if (s.t == 1) {
          if (n_.leaf) return;
          double rr = bt_uniform(st) * dp.getl(s.e, s.g, 1);
          const Clade& g = x.clades[s.g - 1];
          std::vector<St> next;
          auto pick = [&](double term, std::vector<St> nx) {
            if (!next.empty()) return;
            if (term > 0) { rr -= term; if (rr < 0) next = nx; }
          };
          // a term that ends the scan with rr still >= 0 (rounding): the last
          // non-zero one is taken
          std::vector<St> last;
          auto note = [&](double term, std::vector<St> nx) { if (term > 0) last = nx; pick(term, nx); };
          if (n_.root) {
            const double eta = r.get(m, n_).eta, eps = dp.eps_top(s.e), xi = 1.0 - (1.0 - eta) * eps;
            const double cA = xi * (1.0 - eta) / eta, cB = eta * (1.0 - eps) / (xi * xi);
            const int f = n_.children[0], gg = n_.children[1];
            const int tf = m.nd(f).k + 1, tg = m.nd(gg).k + 1;
            for (auto& t3 : g.splits) {
              note(t3.p * dp.getl(s.e, t3.g1, 1) * dp.getl(s.e, t3.g2, 1) * cA, {{s.e, t3.g1, 1}, {s.e, t3.g2, 1}});
              note(t3.p * dp.getl(f, t3.g1) * dp.getl(gg, t3.g2) * cB, {{f, t3.g1, tf}, {gg, t3.g2, tg}});
              note(t3.p * dp.getl(gg, t3.g1) * dp.getl(f, t3.g2) * cB, {{gg, t3.g1, tg}, {f, t3.g2, tf}});
            }
            note(dp.getl(f, s.g) * dp.eps_top(gg) * cB, {{f, s.g, tf}, {gg, 0, 0}});
            note(dp.getl(gg, s.g) * dp.eps_top(f) * cB, {{gg, s.g, tg}, {f, 0, 0}});
          } else if (n_.wgd) {
            const double q = r.get(m, n_).q;
            const int f = n_.children[0], tf = m.nd(f).k + 1;
            note((1.0 - q + 2.0 * q * dp.eps_top(f)) * dp.getl(f, s.g), {{f, s.g, tf}});
            for (auto& t3 : g.splits)
              note(q * t3.p * dp.getl(f, t3.g1) * dp.getl(f, t3.g2), {{f, t3.g1, tf}, {f, t3.g2, tf}});
          } else {
            const int f = n_.children[0], gg = n_.children[1];
            const int tf = m.nd(f).k + 1, tg = m.nd(gg).k + 1;
            note(dp.getl(f, s.g) * dp.eps_top(gg), {{f, s.g, tf}, {gg, 0, 0}});
            note(dp.getl(gg, s.g) * dp.eps_top(f), {{gg, s.g, tg}, {f, 0, 0}});
            for (auto& t3 : g.splits) {
              note(t3.p * dp.getl(f, t3.g1) * dp.getl(gg, t3.g2), {{f, t3.g1, tf}, {gg, t3.g2, tg}});
              note(t3.p * dp.getl(gg, t3.g1) * dp.getl(f, t3.g2), {{gg, t3.g1, tg}, {f, t3.g2, tf}});
            }
          }
          if (next.empty()) next = last;
          for (auto& nx : next) go(node, nx);
          return;
        }
        // within the branch
        const Clade& g = x.clades[s.g - 1];
        if (g.size == 1) return bt(node, {s.e, s.g, 1});
        double rr = bt_uniform(st) * dp.getl(s.e, s.g, s.t);
        const double stay = tb.phi[s.e - 1][s.t - 1] * dp.getl(s.e, s.g, s.t - 1);
        rr -= stay;
        if (rr < 0) return bt(node, {s.e, s.g, s.t - 1});
        const double ps = tb.psi[s.e - 1][s.t - 1];
        // the duplication where rr falls below 0, or (rounding) the last
        // non-zero one; with no duplication weight at all, "no event"
        std::vector<St> next;
        for (auto& t3 : g.splits) {
          const double term = t3.p * dp.getl(s.e, t3.g1, s.t - 1) * dp.getl(s.e, t3.g2, s.t - 1) * ps;
          if (term > 0) {
            next = {{s.e, t3.g1, s.t - 1}, {s.e, t3.g2, s.t - 1}};
            rr -= term;
            if (rr < 0) break;
          }
        }
        if (next.empty()) return bt(node, {s.e, s.g, s.t - 1});
        for (auto& nx : next) go(node, nx);
      };
      add(gam, root, 1, -1);
      bt(0, {root, gam, 1});
      counts[smp] = over ? -1 : n;
    }
    return 0;
  } catch (std::exception& e) { g_err = e.what(); return -1; }
}

}  // extern "C"
