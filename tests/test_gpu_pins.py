"""HIP-path parity on the cases whose oracle functions are pinned by
independent routes in test_oracle.py: leaf sampling probabilities p ≠ 0
(src/rmodels.jl:14, src/model.jl:170), the MUL-tree leafℙ = 1/m
(src/model.jl:110-111) expanded by hand, NonExtinctCondition and
RootCondition against the extinction ODE, and the gradient at the KAT-3
parameters whose root rates are NaN (test/runtests.jl:44-49)."""
import math
import os

import numpy as np
import pytest

from conftest import EXTREE, GOLDEN, KAT_INSERTS, example1_texts
from oracle.oracle import OracleModel, DLWGD as ODLWGD
from test_oracle import TWO_GENES, TWO_PARALOGS, _extinction_by_ode, mul_two_copies_by_hand

pytestmark = pytest.mark.gpu


def kat_tree(w):
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    return t


@pytest.fixture(scope="module")
def kat_p(w):
    rng = np.random.default_rng(31)
    lam = list(rng.normal(-0.4, 0.3, 17))
    mu = list(rng.normal(-0.7, 0.3, 17))
    p = list(rng.uniform(0.05, 0.45, 9))  # one distinct value per leaf id
    r = w.DLWGD(λ=lam, μ=mu, q=[0.35, 0.15], p=p, η=0.85)
    m = w.WhaleModel(r, kat_tree(w), 0.05)
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    oc = om.ccds(example1_texts())
    orr = ODLWGD(lam, mu, [0.35, 0.15], 0.85, p)
    return m, c, om, oc, orr


def test_leaf_sampling_values(w, kat_p):
    m, c, om, oc, orr = kat_p
    ref = om.logpdf(oc, orr, per_family=True)
    refl = om.logpdf(oc, orr, per_family=True, precision="long")
    # p enters: the same families without p differ well beyond the tolerance
    r0 = ODLWGD(orr.lam, orr.mu, orr.q, orr.eta, [])
    assert np.min(np.abs(om.logpdf(oc, r0, per_family=True) - ref) / np.abs(ref)) > 1e-6
    for mode, want, tol in (("fp64", ref, 1e-12), ("fp64s", ref, 1e-12), ("fp32", refl, 1e-6)):
        got = w.logpdf_families(m, c, mode=mode)
        err = np.max(np.abs(got - want) / np.abs(want))
        print(mode, "max rel err with p != 0", err)
        assert err <= tol
    tot = om.logpdf(oc, orr)
    assert abs(w.logpdf(m, c) - tot) <= 1e-12 * abs(tot)


def test_leaf_sampling_gradient(w, kat_p):
    m, c, om, oc, orr = kat_p
    ov, og = om.gradient(oc, orr, nthreads=8)
    for mode, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        v, g, _ = w.gradient(m, c, mode=mode)
        err = np.max(np.abs(g - og)) / np.max(np.abs(og))
        print(mode, "grad rel err with p != 0", err)
        assert err <= tol


def test_kat3_gradient_nan_root_rates(w):
    """The root's λ and μ never enter the DP, so with NaN there the value is
    KAT-2's and the gradient is finite, exactly 0 at the root's entries and
    equal to the dual-number gradient elsewhere."""
    lam, mu = np.ones(17), np.ones(17)
    t = kat_tree(w)
    m = w.WhaleModel(w.DLWGD(λ=list(lam), μ=list(mu), q=[0.2, 0.1], η=0.9), t, 0.05)
    root = m.root().id
    lam[root - 1] = mu[root - 1] = math.nan
    m = m(w.DLWGD(λ=list(lam), μ=list(mu), q=[0.2, 0.1], η=0.9))
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    ov, og = om.gradient(om.ccds(example1_texts()), ODLWGD(list(lam), list(mu), [0.2, 0.1], 0.9), nthreads=8)
    for mode, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        v, g, _ = w.gradient(m, c, mode=mode)
        assert np.all(np.isfinite(g)), mode
        assert g[root - 1] == 0.0 and g[17 + root - 1] == 0.0, mode
        assert np.max(np.abs(g - og)) <= tol * np.max(np.abs(og)), mode


@pytest.mark.parametrize("p", [(0.0, 0.0), (0.2, 0.35)])
def test_mul_two_copies_hand_expanded(w, p):
    lam, mu, eta = 0.3, 0.2, 0.8
    want = mul_two_copies_by_hand(0.7, 0.4, lam, mu, eta, *p)
    t = w.readnw("(A:0.7,A:0.4);")
    r = w.DLWGD(λ=[math.log(lam)] * 3, μ=[math.log(mu)] * 3, p=list(p), η=eta)
    m = w.WhaleModel(r, t, 0.05)
    c = w.ccds_from_texts([TWO_GENES], m)
    for mode, tol in (("fp64", 1e-12), ("fp64s", 1e-12), ("fp32", 1e-6)):
        got = w.logpdf_families(m, c, mode=mode)[0]
        assert abs(got - want) <= tol * abs(want), (mode, got, want)


@pytest.mark.parametrize("cond", ["nonextinct", "root"])
def test_conditions_against_extinction_ode(w, cond):
    """The device's conditioned batch value Σℓ_i − N·c, with c from the
    extinction ODE, on a tree with a WGD and leaf sampling p ≠ 0."""
    rng = np.random.default_rng(7)
    tree, ins = "((A:0.6,B:0.4):0.5,C:0.9);", [("wgd_1", "A", "A")]
    om = OracleModel(tree, ins, 0.05, condition=cond)
    n = om.n_nonwgd
    lam = list(np.log(rng.uniform(0.15, 0.9, n)))
    mu = list(np.log(rng.uniform(0.15, 0.9, n)))
    q, eta, p = [0.4], 0.7, list(rng.uniform(0.05, 0.5, 3))
    ext_f, ext_g, ext_both = _extinction_by_ode(om, ODLWGD(lam, mu, q, eta, p))
    cval = math.log(1 - ext_both) if cond == "nonextinct" else math.log(1 - ext_f - ext_g + ext_both)
    t = w.readnw(tree)
    w.insertnode(w.getlca(t, "A", "A"), name="wgd_1")
    cls = w.NonExtinctCondition if cond == "nonextinct" else w.RootCondition
    m = w.WhaleModel(w.DLWGD(λ=lam, μ=mu, q=q, p=p, η=eta), t, 0.05, condition=cls())
    texts = [TWO_PARALOGS.replace("B_1", "C_1"), TWO_PARALOGS]
    c = w.ccds_from_texts(texts, m)
    fam = w.logpdf_families(m, c)
    tot = w.logpdf(m, c)
    want = fam.sum() - len(fam) * cval
    assert abs(tot - want) <= 1e-12 * abs(want), (tot, want)
