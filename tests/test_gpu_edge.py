"""Edge cases of the HIP path against the CPU oracle: the smallest families
(two genes, in the same species or across the root), a family that is all
in-paralogs, repeated families in one batch, and batch sizes around the
ticket and workgroup granularity."""
import numpy as np
import pytest

from ale_util import ale_from_tree, balanced_paralogs
from oracle.oracle import OracleModel, DLWGD as ODLWGD

pytestmark = pytest.mark.gpu

TREE = "((A:0.4,B:0.3):0.5,(C:0.6,D:0.2):0.3);"


def two_genes(a, b):
    return ale_from_tree({1: [2], 2: [1]}, {1: a, 2: b})


def rates(w, n=7, seed=3):
    rng = np.random.default_rng(seed)
    return w.DLWGD(λ=list(np.log(0.3) + 0.3 * rng.standard_normal(n)),
                   μ=list(np.log(0.2) + 0.3 * rng.standard_normal(n)), η=0.8)


def check(w, texts, tree=TREE, dt=0.02):
    r = rates(w)
    om = OracleModel(tree, [], dt)
    oc = om.ccds(texts)
    ro = ODLWGD(list(r.λ), list(r.μ), [], r.η)
    m = w.WhaleModel(r, w.readnw(tree), dt)
    c = w.ccds_from_texts(texts, m)
    ref = om.logpdf(oc, ro, per_family=True)
    refl = om.logpdf(oc, ro, per_family=True, precision="long")
    for mode, want, tol in (("fp64", ref, 1e-12), ("fp64s", refl, 1e-12), ("fp32", refl, 1e-6)):
        got = w.logpdf_families(m, c, mode=mode)
        assert np.array_equal(np.isfinite(got), np.isfinite(want)), mode
        fin = np.isfinite(want)
        err = np.max(np.abs(got[fin] - want[fin]) / np.abs(want[fin]))
        print(mode, "max rel err", err)
        assert err <= tol, mode
    ov, og = om.gradient(oc, ro, nthreads=4)
    v, g, _ = w.gradient(m, c)
    assert abs(v - ov) <= 1e-12 * abs(ov)
    assert np.max(np.abs(g - og)) <= 1e-10 * np.max(np.abs(og))
    return m, c


def test_two_genes_across_the_root(w):
    check(w, [two_genes("A_1", "C_1")])


def test_two_genes_in_one_species(w):
    check(w, [two_genes("B_1", "B_2")])


def test_in_paralogs_only(w):
    check(w, [ale_from_tree(*balanced_paralogs(12, "D", ("D_99",)))])


def test_repeated_families_and_batch_sizes(w):
    from whale_jl_amd import synth
    wl = synth.workload("config1", n_families=5, nthreads=4)
    base = wl.texts
    # batches of 1, 63, 65 and 130 families made of repeats of five families
    for nb in (1, 63, 65, 130):
        texts = [base[i % 5] for i in range(nb)]
        m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
        c = w.ccds_from_texts(texts, m)
        for mode in ("fp64", "fp32"):
            fam = w.logpdf_families(m, c, mode=mode)
            for i in range(nb):
                assert fam[i] == fam[i % 5], (nb, mode, i)  # one wave per family: same bits
        om = OracleModel(wl.newick, [], wl.dt)
        ref = om.logpdf(om.ccds(texts), ODLWGD(list(wl.rates.λ), list(wl.rates.μ), [], wl.rates.η))
        assert abs(w.logpdf(m, c) - ref) <= 1e-12 * abs(ref)
