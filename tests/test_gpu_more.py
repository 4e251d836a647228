"""More HIP-path parity cases against the CPU oracle: WGD workloads with
branch-wise rates, MUL trees, the three conditions, constant rates, the
λ = μ branch of getα, ℓ download and the discretisation invariant."""
import math
import os
import re

import numpy as np
import pytest

from conftest import EXTREE, GOLDEN, KAT_INSERTS, MULTREE, example1_texts
from oracle.oracle import OracleModel, DLWGD as ODLWGD, ConstantDLWGD as OConst

pytestmark = pytest.mark.gpu


def orates(r):
    return ODLWGD(list(r.λ), list(r.μ), list(r.q), r.η, list(r.p))


@pytest.fixture(scope="module")
def cfg2(w):
    from whale_jl_amd import synth
    wl = synth.workload("config2", n_families=120)
    om = OracleModel(wl.newick, [], wl.dt)
    oc = om.ccds(wl.texts)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    return wl, om, oc, m, c


def test_config2_wgd_branch_rates_fp64(w, cfg2):
    wl, om, oc, m, c = cfg2
    ref = om.logpdf(oc, orates(wl.rates), per_family=True)
    got = w.logpdf_families(m, c)
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(got))
    assert np.max(np.abs(got[fin] - ref[fin]) / np.abs(ref[fin])) <= 1e-12
    assert abs(w.logpdf(m, c) - om.logpdf(oc, orates(wl.rates))) <= 1e-12 * abs(om.logpdf(oc, orates(wl.rates)))


def test_config2_scaled_forms_vs_long_double(w, cfg2):
    wl, om, oc, m, c = cfg2
    ref = om.logpdf(oc, orates(wl.rates), per_family=True, precision="long")
    for mode, tol in (("fp32", 1e-6), ("fp64s", 1e-12)):
        got = w.logpdf_families(m, c, mode=mode)
        err = np.max(np.abs(got - ref) / np.abs(ref))
        print(mode, "max rel err vs long double", err)
        assert err <= tol


def test_config2_gradient(w, cfg2):
    wl, om, oc, m, c = cfg2
    ov, og = om.gradient(oc, orates(wl.rates), nthreads=8)
    for mode, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        v, g, _ = w.gradient(m, c, mode=mode)
        err = np.max(np.abs(g - og)) / np.max(np.abs(og))
        print(mode, "grad rel err", err)
        assert err <= tol


@pytest.mark.parametrize("cond", ["root", "nonextinct", "none"])
def test_conditions(w, cond):
    cls = {"root": w.RootCondition, "nonextinct": w.NonExtinctCondition, "none": w.NoCondition}[cond]
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    rng = np.random.default_rng(11)
    lam = list(rng.normal(-0.5, 0.4, 17))
    mu = list(rng.normal(-0.7, 0.4, 17))
    r = w.DLWGD(λ=lam, μ=mu, q=[0.3, 0.6], η=0.8)
    m = w.WhaleModel(r, t, 0.05, condition=cls())
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05, condition=cond)
    oc = om.ccds(example1_texts())
    orr = ODLWGD(lam, mu, [0.3, 0.6], 0.8)
    ref = om.logpdf(oc, orr)
    assert abs(w.logpdf(m, c) - ref) <= 1e-12 * abs(ref)
    ov, og = om.gradient(oc, orr)
    v, g, _ = w.gradient(m, c)
    assert np.max(np.abs(g - og)) <= 1e-10 * np.max(np.abs(og))


def test_constant_rates_gradient(w):
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    r = w.ConstantDLWGD(λ=0.7, μ=0.4, q=[0.2, 0.1], η=0.8)
    m = w.WhaleModel(r, t, 0.05)
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    oc = om.ccds(example1_texts())
    ov, og = om.gradient(oc, OConst(0.7, 0.4, [0.2, 0.1], 0.8))
    v, g, _ = w.gradient(m, c)
    assert abs(v - ov) <= 1e-12 * abs(ov)
    assert np.max(np.abs(g - og)) <= 1e-10 * np.max(np.abs(og))


def test_equal_rates_branch_semantics(w):
    # at λ = μ the λΔ/(1+λΔ) branch of getα is differentiated, as ForwardDiff does
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    n = 17
    x = np.concatenate([np.ones(n), np.ones(n), [0.1, 0.2], [0.8]])
    r = w.DLWGD(λ=list(x[:n]), μ=list(x[n:2 * n]), q=[0.1, 0.2], η=0.8)
    m = w.WhaleModel(r, t, 0.05)
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    oc = om.ccds(example1_texts())
    ov, og = om.gradient(oc, ODLWGD(list(x[:n]), list(x[n:2 * n]), [0.1, 0.2], 0.8))
    v, g, _ = w.gradient(m, c)
    assert np.max(np.abs(g - og)) <= 1e-10 * np.max(np.abs(og))


def test_mul_tree_parity(w):
    t = w.readnw(MULTREE)
    n = len(w.postwalk(t))
    rng = np.random.default_rng(5)
    lam, mu = list(rng.normal(-1, 0.3, n)), list(rng.normal(-1, 0.3, n))
    m = w.WhaleModel(w.DLWGD(λ=lam, μ=mu, η=0.9), t, 0.05)
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    om = OracleModel(MULTREE, [], 0.05)
    oc = om.ccds(example1_texts())
    ref = om.logpdf(oc, ODLWGD(lam, mu, [], 0.9), per_family=True)
    for mode, tol in (("fp64", 1e-12), ("fp32", 1e-6), ("fp64s", 1e-12)):
        got = w.logpdf_families(m, c, mode=mode)
        assert np.max(np.abs(got - ref) / np.abs(ref)) <= tol
    ov, og = om.gradient(oc, ODLWGD(lam, mu, [], 0.9))
    v, g, _ = w.gradient(m, c)
    assert np.max(np.abs(g - og)) <= 1e-10 * np.max(np.abs(og))


def test_ell_download_matches_oracle(w):
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    r = w.DLWGD(λ=[0.1] * 17, μ=[-0.2] * 17, q=[0.2, 0.1], η=0.9)
    m = w.WhaleModel(r, t, 0.05)
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    oc = om.ccds(example1_texts())
    for mode, tol in (("fp64", 1e-13), ("fp64s", 1e-13), ("fp32", 1e-5)):
        w.logpdf_(m, c, mode=mode)
        for f in (0, 5, 11):
            ref = om.ell(oc, f, ODLWGD([0.1] * 17, [-0.2] * 17, [0.2, 0.1], 0.9))
            flat = np.concatenate([c[f].ell[i].T.ravel() for i in range(1, 20)])
            assert flat.shape == ref.shape
            scale = np.maximum(np.abs(ref), 1e-300)
            big = np.abs(ref) > 1e-30 * np.max(np.abs(ref))
            assert np.max(np.abs(flat - ref)[big] / scale[big]) <= tol


def test_discretization_convergence(w):
    t = open(os.path.join(GOLDEN, "example-5", "tree.nw")).readline().strip()
    t1 = re.sub(r":[0-9.eE+-]+", ":1", t).rstrip(";") + ":1;"
    rng = np.random.default_rng(7)
    for _ in range(2):
        tree = w.readnw(t1)
        n = len(w.postwalk(tree))
        rr = list(rng.standard_normal(n))
        vals = []
        for e in np.arange(-5, -0.99, 0.5):
            m = w.WhaleModel(w.DLWGD(λ=rr, μ=rr, η=0.9), w.readnw(t1), 10.0 ** e)
            c = w.read_ale(os.path.join(GOLDEN, "example-5", "OG0014587.ale"), m)
            vals.append(w.logpdf_(m, c))
            om = OracleModel(t1, [], 10.0 ** e)
            ref = om.logpdf(om.ccds([open(os.path.join(GOLDEN, "example-5", "OG0014587.ale")).read()]),
                            ODLWGD(rr, rr, [], 0.9))
            assert abs(vals[-1] - ref) <= 1e-11 * abs(ref)
        vals = np.array(vals)
        assert np.all(np.abs(vals - vals[0]) < 0.1)


def _paralog_case(w, tree, dt, others, lam, mu):
    from ale_util import ale_from_tree, balanced_paralogs
    adj, names = balanced_paralogs(64, "A", others)
    text = ale_from_tree(adj, names)
    om = OracleModel(tree, [], dt)
    oc = om.ccds([text])
    ref = om.logpdf(oc, ODLWGD(list(lam), list(mu), [], 0.9), per_family=True, precision="long")[0]
    m = w.WhaleModel(w.DLWGD(λ=list(lam), μ=list(mu), η=0.9), w.readnw(tree), dt)
    c = w.ccds_from_texts([text], m)
    got = w.logpdf_families(m, c, mode="fp32")[0]
    n_fb = w.device_batch(m, c).fallbacks()
    v, g, _ = w.gradient(m, c, mode="fp32")
    v64, g64, _ = w.gradient(m, c, mode="fp64s")
    return ref, got, n_fb, g, g64


def test_fp32_young_clade_renormalised_in_place(w):
    """64 in-paralogs on one long, finely sliced leaf branch: their clade values
    grow by ~2^360 after first becoming non-zero.  They had no values before
    this branch, so FP32 renormalises them in place; no FP64 recomputation."""
    ref, got, n_fb, g, g64 = _paralog_case(w, "(A:1.0,B:1.0);", 0.005, ("B_1", "B_2"),
                                           [np.log(0.8)] * 3, [np.log(0.3)] * 3)
    print("young expansion: ref", ref, "fp32", got, "fallbacks", n_fb)
    assert n_fb == 0
    assert abs(got - ref) <= 1e-6 * abs(ref)
    assert np.max(np.abs(g - g64)) <= 1e-5 * np.max(np.abs(g64))


def test_fp32_fallback_recomputes_in_fp64(w):
    """The same clade born at the top of a 6-slice leaf branch grows by ~2^354
    along a 200-slice parent branch (its sister loses genes readily): FP32
    cannot hold it, the family is recomputed in FP64 with per-clade exponents."""
    mu = [np.log(0.3)] * 5
    mu[1] = np.log(3.0)  # node 2 (B)
    ref, got, n_fb, g, g64 = _paralog_case(w, "((A:0.03,B:1.0):1.0,C:1.03);", 0.005, ("B_1", "C_1"),
                                           [np.log(0.8)] * 5, mu)
    print("grown expansion: ref", ref, "fp32", got, "fallbacks", n_fb)
    assert n_fb == 1
    assert abs(got - ref) <= 1e-9 * abs(ref)
    assert np.max(np.abs(g - g64)) <= 1e-9 * np.max(np.abs(g64))


def test_fp32_overflowing_family_is_routed_to_fp64(w):
    """A batch of ordinary families and the overflowing family above (first,
    middle and last).  The first FP32 call recomputes it in FP64 after the
    FP32 pass; later FP32 calls evaluate it in FP64 from the start, beside the
    FP32 size classes.  Every call must give the same values and gradient."""
    from ale_util import ale_from_tree, balanced_paralogs
    tree, dt = "((A:0.03,B:1.0):1.0,C:1.03);", 0.005
    lam, mu = [np.log(0.8)] * 5, [np.log(0.3)] * 5
    mu[1] = np.log(3.0)
    big = ale_from_tree(*balanced_paralogs(64, "A", ("B_1", "C_1")))
    small = [ale_from_tree(*balanced_paralogs(k, "A", ("B_1", "C_1"))) for k in (1, 2, 3, 5, 8)]
    texts = small[:2] + [big] + small[2:] + [big]
    om = OracleModel(tree, [], dt)
    ref = om.logpdf(om.ccds(texts), ODLWGD(lam, mu, [], 0.9), per_family=True, precision="long")
    m = w.WhaleModel(w.DLWGD(λ=lam, μ=mu, η=0.9), w.readnw(tree), dt)
    c = w.ccds_from_texts(texts, m)
    db = w.device_batch(m, c)
    v64, g64, f64 = w.gradient(m, c, mode="fp64s")
    runs = []
    for k in range(3):
        v, g, fam = w.gradient(m, c, mode="fp32")
        runs.append((v, g, fam, db.fallbacks()))
        print("call", k, "fallbacks", runs[-1][3])
    for v, g, fam, nfb in runs:
        assert nfb == 2
        assert np.max(np.abs(fam - ref) / np.abs(ref)) <= 1e-6
        assert abs(fam[2] - f64[2]) <= 1e-12 * abs(f64[2]) and abs(fam[-1] - f64[-1]) <= 1e-12 * abs(f64[-1])
        assert abs(v - v64) <= 1e-6 * abs(v64)
        assert np.max(np.abs(g - g64)) <= 1e-5 * np.max(np.abs(g64))
    # routed families never change the other modes
    fam = w.logpdf_families(m, c, mode="fp64")
    assert np.max(np.abs(fam - ref) / np.abs(ref)) <= 1e-12
    assert np.array_equal(runs[1][2], runs[2][2])
