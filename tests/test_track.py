"""The oracle's backtracking sampler (src/track.jl:188-414), which the device
sampler is compared with: its trees hold every gene once, the per-pair event
posteriors of each family sum to one (the reference's own check,
test/runtests.jl:95-109), and the first event at the root follows the
decomposition of the root's value into its terms (src/core.jl:147), checked
by a chi-square test at a fixed seed.  No GPU."""
import math

import numpy as np
import pytest

from conftest import EXTREE, KAT_INSERTS, example1_texts
from oracle.oracle import OracleModel, DLWGD as ODLWGD


@pytest.fixture(scope="module")
def kat(w):
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    oc = om.ccds(example1_texts())
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    wm = w.WhaleModel(w.DLWGD(λ=[1.0] * 17, μ=[1.0] * 17, q=[0.2, 0.1], η=0.9), t, 0.05)  # structure only
    return om, oc, wm, ODLWGD([1.0] * 17, [1.0] * 17, [0.2, 0.1], 0.9)


def test_trees_hold_every_gene_once(w, kat):
    om, oc, wm, r = kat
    for fam in range(len(oc)):
        e = oc.export(fam)
        genes = sorted(np.where(e["size"] == 1)[0])
        for x in om.backtrack(oc, fam, r, seed=3, n_samples=20):
            t = w.RecTree(x)
            assert t.genes() == genes
            assert t.nodes[0, 3] == -1 and np.all(t.nodes[1:, 3] < np.arange(1, len(t)))


def test_pair_posteriors_sum_to_one(w, kat):
    """test/runtests.jl:95-109: 100 backtracks of the first family, every
    gene pair's event posteriors add up to 1."""
    om, oc, wm, r = kat
    trees = [w.RecTree(x) for x in om.backtrack(oc, 0, r, seed=11, n_samples=100)]
    pp = w.pair_posteriors(trees, wm)
    nleaf = int(np.sum(oc.export(0)["size"] == 1))
    assert len(pp) == nleaf * (nleaf - 1) // 2
    labels = set(w.pair_labels(wm))
    for pair, d in pp.items():
        assert set(d) <= labels
        assert abs(sum(d.values()) - 1.0) <= 1e-9, (pair, d)


def root_event_weights(om, oc, fam, r, eta):
    """The root value of the ubiquitous clade Γ split into its three terms
    (src/core.jl:147) over the value, from the oracle's ℓ: a duplication at
    the root ((1-η)ξ/η Σ p ℓρ(γ1)ℓρ(γ2)), a speciation (η(1-ε)/ξ² Σ p (ℓf ℓg
    + ℓg ℓf)) and a speciation with a loss (η(1-ε)/ξ² (ℓf(Γ)εg + ℓg(Γ)εf))."""
    e = oc.export(fam)
    nodes = {d["id"]: d for d in om.nodes()}
    root = next(i for i, d in nodes.items() if d["root"])
    f, g = sorted(i for i, d in nodes.items() if d["parent"] == root)
    ell = om.ell(oc, fam, r)
    nc = len(e["size"])
    species = []
    for c in range(nc):
        if e["size"][c] == 1:
            species.append({int(e["leaf_species"][c])})
        else:
            a = e["split_off"][c]
            species.append(species[e["g1"][a] - 1] | species[e["g2"][a] - 1])
    under = {}
    for i in om.order():  # children before parents
        ch = [j for j, x in nodes.items() if x["parent"] == i]
        under[i] = {i} if nodes[i]["leaf"] else set().union(*[under[j] for j in ch])
    # ℓ per node id: (k+1) x |compat| column-major, compat in increasing clade id
    ofs, tops, root_vals = 0, {}, None
    for i in sorted(nodes):
        d = nodes[i]
        R = d["k"] + 1
        comp = list(range(nc)) if d["root"] else [c for c in range(nc) if species[c] <= under[i]]
        block = ell[ofs:ofs + R * len(comp)].reshape(len(comp), R)
        ofs += R * len(comp)
        tops[i] = {c: block[j, -1] for j, c in enumerate(comp)}
        if d["root"]:
            root_vals = block[:, 0]
    assert ofs == len(ell)
    eps_f = om.tables(r, f, nodes[f]["k"] + 1)[0][-1]
    eps_g = om.tables(r, g, nodes[g]["k"] + 1)[0][-1]
    eps = eps_f * eps_g
    xi = 1 - (1 - eta) * eps
    cA, cB = (1 - eta) * xi / eta, eta * (1 - eps) / xi ** 2
    G = nc - 1
    lf = lambda c: tops[f].get(c, 0.0)
    lg = lambda c: tops[g].get(c, 0.0)
    dup = spec = 0.0
    for k in range(e["split_off"][G], e["split_off"][G + 1]):
        c1, c2, p = e["g1"][k] - 1, e["g2"][k] - 1, e["p"][k]
        dup += cA * p * root_vals[c1] * root_vals[c2]
        spec += cB * p * (lf(c1) * lg(c2) + lg(c1) * lf(c2))
    loss = cB * (lf(G) * eps_g + lg(G) * eps_f)
    assert abs(dup + spec + loss - root_vals[G]) <= 1e-12 * root_vals[G]
    return np.array([dup, spec, loss]) / root_vals[G], root


def classify_root_event(t, root):
    ch = t.children()[0]
    if all(t.nodes[c, 1] == root - 1 for c in ch):
        return 0
    if any(t.nodes[c, 0] < 0 for c in ch):
        return 2
    return 1


@pytest.mark.parametrize("fam,lam,mu,eta", [(3, -1.5, -1.0, 0.7), (8, -1.0, -0.5, 0.6), (11, -2.0, -1.5, 0.8)])
def test_root_event_frequencies_follow_the_root_terms(w, kat, fam, lam, mu, eta):
    """The first event of every tree, at the root clade Γ, follows the root
    value's terms, chi-square at a fixed seed."""
    om, oc, wm, _ = kat
    r = ODLWGD([lam] * 17, [mu] * 17, [0.2, 0.1], eta)
    want, root = root_event_weights(om, oc, fam, r, eta)
    print("family", fam, "root event weights", want)
    assert np.sum(want > 0.02) >= 2  # a case where the draw decides something
    n = 20000
    got = np.zeros(3)
    for x in om.backtrack(oc, fam, r, seed=2024 + fam, n_samples=n):
        got[classify_root_event(w.RecTree(x), root)] += 1
    print("sampled", got / n)
    m = want > 0
    assert np.all(got[~m] == 0)
    chi2 = float(np.sum((got[m] - n * want[m]) ** 2 / (n * want[m])))
    assert chi2 < 13.8  # chi-square, at most 2 degrees of freedom, p = 0.001
