"""Families whose stages do not fit one compute unit's LDS run with their
stage buffers in global memory (one wave per workgroup).  Checked against the
oracle on the reference's own largest inputs, which the reference evaluates
with no size limit (src/core.jl:83-158):

  example/example-4: cytp450.ale, 95 genes, 4,917 clades, 12,973 splits, on
      the 58-taxon PLAZA tree (every branch 1.0); its root stage needs more
      LDS than a compute unit has in the FP64 forms;
  docs/data/landplant: the tutorial's 100 families (up to 1,025 clades) on
      its 13-taxon tree, ConstantDLWGD(λ=0.1, μ=0.2, η=1/1.5), Δt = 0.01
      (docs/lit/tutorial.jl:88-124);

and with every family of the KAT and config2 workloads forced onto the
global-memory path, against the same oracle."""
import glob
import os

import numpy as np
import pytest

from conftest import EXTREE, GOLDEN, KAT_INSERTS, example1_texts
from oracle.oracle import OracleModel, ConstantDLWGD as OConst, DLWGD as ODLWGD

pytestmark = pytest.mark.gpu


def orates(r):
    return ODLWGD(list(r.λ), list(r.μ), list(r.q), r.η, list(r.p))


def check_values(w, m, c, om, oc, orr, tol32=1e-6):
    ref = om.logpdf(oc, orr, per_family=True)
    refl = om.logpdf(oc, orr, per_family=True, precision="long")
    got = w.logpdf_families(m, c)
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(got))
    if fin.any():
        assert np.max(np.abs(got[fin] - ref[fin]) / np.abs(ref[fin])) <= 1e-12
    for mode, tol in (("fp64s", 1e-12), ("fp32", tol32)):
        got = w.logpdf_families(m, c, mode=mode)
        err = np.max(np.abs(got - refl) / np.abs(refl))
        print(mode, "max rel err vs long double", err)
        assert err <= tol


def check_gradient(w, m, c, om, oc, orr):
    ov, og = om.gradient(oc, orr, nthreads=16)
    for mode, vtol, gtol in (("fp64s", 1e-12, 1e-10), ("fp32", 1e-6, 1e-5)):
        v, g, _ = w.gradient(m, c, mode=mode)
        err = np.max(np.abs(g - og)) / np.max(np.abs(og))
        print(mode, "grad rel err", err)
        assert abs(v - ov) <= vtol * abs(ov)
        assert err <= gtol


@pytest.mark.parametrize("dt", [0.1, 0.05])
def test_example4_cytp450(w, dt):
    tree = open(os.path.join(GOLDEN, "example-4", "tree.nw")).read().strip()
    ale = open(os.path.join(GOLDEN, "example-4", "cytp450.ale")).read()
    r = w.ConstantDLWGD(λ=0.1, μ=0.2, η=1 / 1.5)
    m = w.WhaleModel(r, w.readnw(tree), dt)
    c = w.ccds_from_texts([ale], m)
    assert len(c[0]) == 4917
    assert w.device_batch(m, c).hbm_families() == 1
    om = OracleModel(tree, [], dt)
    oc = om.ccds([ale])
    orr = OConst(0.1, 0.2, [], 1 / 1.5)
    check_values(w, m, c, om, oc, orr)
    check_gradient(w, m, c, om, oc, orr)
    # per-family launches take the same path
    got = w.logpdf_families(m, c, mode="fp64s", launch="per_family")
    ref = om.logpdf(oc, orr, per_family=True, precision="long")
    assert abs(got[0] - ref[0]) <= 1e-12 * abs(ref[0])


def test_example4_beside_other_families(w):
    """The large family in one batch with the 12 example-1 families: the
    others keep their LDS size classes, every value matches its own run."""
    tree = open(os.path.join(GOLDEN, "example-4", "tree.nw")).read().strip()
    ale = open(os.path.join(GOLDEN, "example-4", "cytp450.ale")).read()
    r = w.ConstantDLWGD(λ=0.15, μ=0.2, η=0.8)
    m = w.WhaleModel(r, w.readnw(tree), 0.1)
    texts = [ale] * 3
    c = w.ccds_from_texts(texts, m)
    b = w.device_batch(m, c)
    assert b.hbm_families() == 3
    om = OracleModel(tree, [], 0.1)
    oc = om.ccds(texts)
    orr = OConst(0.15, 0.2, [], 0.8)
    check_values(w, m, c, om, oc, orr)


def test_landplant_tutorial(w):
    tree = open(os.path.join(GOLDEN, "landplant", "speciestree.nw")).read().strip()
    files = sorted(glob.glob(os.path.join(GOLDEN, "landplant", "100fams", "*.ale")))
    texts = [open(f).read() for f in files]
    assert len(texts) == 100
    r = w.ConstantDLWGD(λ=0.1, μ=0.2, η=1 / 1.5)
    m = w.WhaleModel(r, w.readnw(tree), 0.01)
    c = w.read_ale(os.path.join(GOLDEN, "landplant", "100fams"), m)
    assert len(c) == 100
    om = OracleModel(tree, [], 0.01)
    oc = om.ccds(texts)
    orr = OConst(0.1, 0.2, [], 1 / 1.5)
    check_values(w, m, c, om, oc, orr)
    check_gradient(w, m, c, om, oc, orr)
    # the tutorial's reparameterisation model((λ=0.2, μ=0.1))
    m2 = m({"λ": 0.2, "μ": 0.1})
    check_values(w, m2, c, om, oc, OConst(0.2, 0.1, [], 1 / 1.5))


@pytest.fixture
def forced(w):
    """A KAT batch re-planned with no LDS for stages: every family on the
    global-memory path."""
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    rng = np.random.default_rng(3)
    lam, mu = list(rng.normal(-0.5, 0.3, 17)), list(rng.normal(-0.7, 0.3, 17))
    r = w.DLWGD(λ=lam, μ=mu, q=[0.3, 0.15], η=0.85)
    m = w.WhaleModel(r, t, 0.05)
    c = w.read_ale(os.path.join(GOLDEN, "example-1"), m)
    b = w.device_batch(m, c)
    b.set_lds_limit(1)
    assert b.hbm_families() == 12
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    yield m, c, om, om.ccds(example1_texts()), ODLWGD(lam, mu, [0.3, 0.15], 0.85)
    b.set_lds_limit(0)
    assert b.hbm_families() == 0


def test_forced_global_path_kat(w, forced):
    m, c, om, oc, orr = forced
    check_values(w, m, c, om, oc, orr)
    check_gradient(w, m, c, om, oc, orr)


def test_forced_global_path_config2(w):
    from whale_jl_amd import synth
    wl = synth.workload("config2", n_families=200)
    om = OracleModel(wl.newick, [], wl.dt)
    oc = om.ccds(wl.texts)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    b = w.device_batch(m, c)
    lim = 24 * 1024  # the larger families take the global path, the others stay in LDS
    b.set_lds_limit(lim)
    n_hbm = b.hbm_families()
    print("families on the global path:", n_hbm, "of", len(c))
    assert 0 < n_hbm < len(c)
    check_values(w, m, c, om, oc, orates(wl.rates))
    check_gradient(w, m, c, om, oc, orates(wl.rates))
    b.set_lds_limit(0)
