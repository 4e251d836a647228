"""NowhereExtinctCondition (src/condition.jl:31-36) on the host: pinned by the
reference's own monotonicity check (test/runtests.jl:55-64) and by a direct
sum over lineage counts with the linear BDP transition probabilities
(src/bdputil.jl:30-36), which shares no code with the pgf route.  The
gradient (complex-step) is pinned by central differences of the value.
No GPU."""
import math

import numpy as np

import pytest

from conftest import EXTREE, KAT_INSERTS, unpack_rates


def kat_tree(w):
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    return t


def test_nowhere_extinct_grows_with_retention(w):
    t = kat_tree(w)
    prev = -math.inf
    for q in [i / 10 for i in range(11)]:
        m = w.WhaleModel(w.ConstantDLWGD(λ=0.3, μ=0.4, q=[q, q], η=0.66), t, 0.05,
                         condition=w.NowhereExtinctCondition())
        c = w.condition(m)
        assert c > prev
        prev = c
    assert prev < 0.0


def _alpha(lam, mu, t):
    if abs(lam - mu) <= 1e-6:
        return lam * t / (1.0 + lam * t)
    return mu * (math.exp(t * (lam - mu)) - 1.0) / (lam * math.exp(t * (lam - mu)) - mu)


def _tp(a, b, t, lam, mu):
    if a == 0:
        return 1.0 if b == 0 else 0.0
    al = _alpha(lam, mu, t)
    be = lam / mu * al
    return sum(math.comb(a, k) * math.comb(a + b - k - 1, a - 1) * al ** (a - k) * be ** (b - k)
               * (1.0 - al - be) ** k for k in range(0, min(a, b) + 1))


def _brute_nowhere(m, nmax=18):
    """P(every leaf has a gene) by summing over lineage counts node by node
    (a leaf needs only P(not extinct) = 1 - tp(a, 0); other branches sum the
    counts at their bottom up to 2a + nmax)."""
    from functools import lru_cache
    lam, mu, q, eta, _, _ = m.node_params()
    nodes = m.s.nodes

    @lru_cache(maxsize=None)
    def g(i, a):  # P(all leaves below node i non-empty | a lineages at the top of its branch)
        d = nodes[i - 1]
        lm, mm, t = float(lam[i - 1]), float(mu[i - 1]), float(d.distance)
        if d.isleaf():
            return 1.0 - _tp(a, 0, t, lm, mm)
        tot = 0.0
        for b in range(1, 2 * a + nmax + 1):
            pb = _tp(a, b, t, lm, mm)
            if d.wgd:
                qq = float(q[d.wgdid - 1])
                c = d.children[0]
                tot += pb * sum(math.comb(b, j) * qq ** j * (1 - qq) ** (b - j) * g(c, b + j)
                                for j in range(0, b + 1))
            else:
                tot += pb * math.prod(g(c, b) for c in d.children)
        return tot

    root = next(d for d in nodes if d.isroot())
    return sum(eta * (1 - eta) ** (n - 1) * math.prod(g(c, n) for c in root.children)
               for n in range(1, nmax + 1))


@pytest.mark.parametrize("rates", [(0.3, 0.4, 0.35), (0.5, 0.5, 0.6), (0.2, 0.6, 0.8)])
def test_nowhere_extinct_matches_count_sum(w, rates):
    lam, mu, q = rates
    t = w.readnw("((A:0.6,B:0.4):0.5,C:0.9);")
    w.insertnode(w.getlca(t, "A", "A"), name="wgd_1")
    m = w.WhaleModel(w.ConstantDLWGD(λ=lam, μ=mu, q=[q], η=0.85), t, 0.05,
                     condition=w.NowhereExtinctCondition())
    got = math.exp(w.condition(m))
    ref = _brute_nowhere(m)
    print(rates, got, ref)
    assert abs(got - ref) <= 1e-9


def _check_grad_vs_central_differences(w, m, h=1e-6, tol=1e-6):
    from whale_jl_amd.mixture import nowhere_extinct, nowhere_extinct_grad
    g = m.chain_gradient(*nowhere_extinct_grad(m))
    th = m.rates.pack()
    fd = np.zeros_like(th)
    for i in range(len(th)):
        up, dn = th.copy(), th.copy()
        up[i] += h
        dn[i] -= h
        fd[i] = (nowhere_extinct(m(unpack_rates(m.rates, up))) - nowhere_extinct(m(unpack_rates(m.rates, dn)))) / (2 * h)
    assert np.max(np.abs(g - fd)) <= tol * max(1.0, np.max(np.abs(fd))), (g, fd)
    # zero only where c does not depend on the entry (the root's λ, μ; μ of a
    # λ = μ branch, whose pgf form has λ alone)
    assert np.count_nonzero(g) >= len(g) - 3


def test_nowhere_extinct_gradient_constant_rates_with_wgds(w):
    """Complex-step ∂ log c against central differences of the value,
    ConstantDLWGD on a tree with two WGD nodes: [λ, μ, q1, q2, η]."""
    m = w.WhaleModel(w.ConstantDLWGD(λ=0.3, μ=0.45, q=[0.2, 0.35], η=0.7), kat_tree(w), 0.05,
                     condition=w.NowhereExtinctCondition())
    _check_grad_vs_central_differences(w, m)


def test_nowhere_extinct_gradient_branch_rates(w):
    """The same with branch-wise log-scale rates, λ ≠ μ and λ = μ branches."""
    t = w.readnw("((A:0.6,B:0.4):0.5,C:0.9);")
    n = len(w.postwalk(t))
    lam = [math.log(0.2 + 0.05 * i) for i in range(n)]
    mu = [math.log(0.3 + 0.04 * i) for i in range(n)]
    mu[1] = lam[1]  # the λ = μ form of the pgf (within 1e-6 under the steps)
    m = w.WhaleModel(w.DLWGD(λ=lam, μ=mu, q=[], η=0.8), t, 0.05, condition=w.NowhereExtinctCondition())
    _check_grad_vs_central_differences(w, m)
