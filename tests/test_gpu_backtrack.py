"""Backtracking on the device (whale_backtrack, src/track.jl:188-414) against
the oracle's sampler: with the same seed both draw the same uniform numbers
in the same order, so from the same ℓ they sample the same reconciled trees.
The FP64 form's ℓ equals the oracle's to rounding, so every tree must match;
the scaled forms' ℓ differ in the last bits of FP32 or FP64, which flips a
draw only when it falls within that distance of an event boundary.  Also
the reference's own check on the trees (test/runtests.jl:95-109) and its MUL
tree case (test/runtests.jl:111-126)."""
import os

import numpy as np
import pytest

from conftest import EXTREE, GOLDEN, KAT_INSERTS, MULTREE, example1_texts
from oracle.oracle import OracleModel, DLWGD as ODLWGD, ConstantDLWGD as OConst

pytestmark = pytest.mark.gpu


def kat(w, lam=1.0, mu=1.0):
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    m = w.WhaleModel(w.DLWGD(λ=[lam] * 17, μ=[mu] * 17, q=[0.2, 0.1], η=0.9), t, 0.05)
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    return m, c, om, om.ccds(example1_texts()), ODLWGD([lam] * 17, [mu] * 17, [0.2, 0.1], 0.9)


def same_fraction(w, trees, om, oc, r, seed, n):
    same = total = 0
    for fam, ts in enumerate(trees):
        ref = om.backtrack(oc, fam, r, seed=seed, n_samples=n)
        for t, x in zip(ts, ref):
            total += 1
            same += int(np.array_equal(t.nodes, x))
    return same / total


def test_device_trees_equal_oracle_trees_fp64(w):
    m, c, om, oc, r = kat(w, -0.5, -0.8)
    w.logpdf_(m, c)  # logpdf!: stores ℓ
    trees = w.backtrack(m, c, n=40, seed=99)
    assert len(trees) == 12 and all(len(ts) == 40 for ts in trees)
    frac = same_fraction(w, trees, om, oc, r, 99, 40)
    print("fp64: identical trees", frac)
    assert frac == 1.0


@pytest.mark.parametrize("mode,least", [("fp64s", 1.0), ("fp32", 0.99)])
def test_device_trees_equal_oracle_trees_scaled(w, mode, least):
    m, c, om, oc, r = kat(w, -0.5, -0.8)
    w.logpdf_(m, c, mode=mode)
    trees = w.backtrack(m, c, n=40, seed=5)
    frac = same_fraction(w, trees, om, oc, r, 5, 40)
    print(mode, "identical trees", frac)
    assert frac >= least


def test_pair_posteriors_sum_to_one(w):
    """test/runtests.jl:95-109: logpdf!, 100 backtracks of the first family,
    every gene pair's event posteriors sum to 1."""
    m, c, om, oc, r = kat(w)
    w.logpdf_(m, c)
    trees = w.backtrack(m, c, n=100, seed=1, families=[0])[0]
    pp = w.pair_posteriors(trees, m)
    n = int(np.sum(c[0].size == 1))
    assert len(pp) == n * (n - 1) // 2
    for pair, d in pp.items():
        assert abs(sum(d.values()) - 1.0) <= 1e-9, (pair, d)
    # the same trees as the oracle's
    ref = om.backtrack(oc, 0, r, seed=1, n_samples=100)
    assert all(np.array_equal(t.nodes, x) for t, x in zip(trees, ref))


def test_after_gradient_and_on_mul_tree(w):
    """The gradient stores ℓ too; the MUL tree of test/runtests.jl:111-126."""
    t = w.readnw(MULTREE)
    n = len(w.postwalk(t))
    m = w.WhaleModel(w.DLWGD(λ=[-1.0] * n, μ=[-1.0] * n, η=0.9), t, 0.05)
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    w.gradient(m, c)
    trees = w.backtrack(m, c, n=100, seed=3)
    om = OracleModel(MULTREE, [], 0.05)
    oc = om.ccds(example1_texts())
    r = ODLWGD([-1.0] * n, [-1.0] * n, [], 0.9)
    for fam, ts in enumerate(trees):
        genes = sorted(np.where(c[fam].size == 1)[0])
        assert all(x.genes() == genes for x in ts)
    assert same_fraction(w, trees, om, oc, r, 3, 100) == 1.0


def test_global_memory_family(w):
    """The 4,917-clade family, whose stages run in global memory."""
    tree = open(os.path.join(GOLDEN, "example-4", "tree.nw")).read().strip()
    ale = open(os.path.join(GOLDEN, "example-4", "cytp450.ale")).read()
    m = w.WhaleModel(w.ConstantDLWGD(λ=0.1, μ=0.2, η=1 / 1.5), w.readnw(tree), 0.1)
    c = w.ccds_from_texts([ale], m)
    w.logpdf_(m, c, mode="fp64s")
    trees = w.backtrack(m, c, n=20, seed=8)[0]
    genes = sorted(np.where(c[0].size == 1)[0])
    assert all(t.genes() == genes for t in trees)
    om = OracleModel(tree, [], 0.1)
    oc = om.ccds([ale])
    ref = om.backtrack(oc, 0, OConst(0.1, 0.2, [], 1 / 1.5), seed=8, n_samples=20)
    assert sum(np.array_equal(t.nodes, x) for t, x in zip(trees, ref)) >= 19


def test_needs_stored_ell(w):
    m, c, om, oc, r = kat(w)
    w.logpdf(m, c)  # no stored rows
    with pytest.raises(Exception, match="stored"):
        w.backtrack(m, c, n=1)
