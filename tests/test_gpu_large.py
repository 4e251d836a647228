"""HIP path vs the CPU oracle on shapes the small configurations do not reach:
stages of more than 64 segments (a 100-taxon tree), stages and root levels
with more items and records than the kernels hold in registers, and the
largest families of the benchmark workload."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleModel, DLWGD as ODLWGD

pytestmark = pytest.mark.gpu


def orates(r):
    return ODLWGD(list(r.λ), list(r.μ), list(r.q), r.η, list(r.p))


@pytest.fixture(scope="module")
def wide(w):
    """100 taxa + 2 WGDs, 24 families of 90–220 genes."""
    from whale_jl_amd import synth
    seed = synth.SEED + 77
    nw = synth.species_tree(seed, 100, 2)
    texts, st = synth.families(nw, seed, 24, ("uniform", 0, 0, 90, 220), 100, 100000, 1000000, nthreads=8)
    rng = np.random.default_rng(seed)
    n = 2 * 100 - 1
    rates = w.DLWGD(λ=list(math.log(0.2) + 0.5 * rng.standard_normal(n)),
                    μ=list(math.log(0.15) + 0.5 * rng.standard_normal(n)), q=[0.25, 0.4], η=0.9)
    om = OracleModel(nw, [], 0.02)
    oc = om.ccds(texts)
    m = w.WhaleModel(rates, w.readnw(nw), 0.02)
    c = w.ccds_from_texts(texts, m)
    print("wide: leaves", st[:, 0].min(), "..", st[:, 0].max(), "clades up to", st[:, 1].max())
    return rates, om, oc, m, c


def test_wide_tree_values(w, wide):
    rates, om, oc, m, c = wide
    ref = om.logpdf(oc, orates(rates), per_family=True)
    fin = np.isfinite(ref)
    got = w.logpdf_families(m, c)
    assert np.array_equal(fin, np.isfinite(got))
    assert np.max(np.abs(got[fin] - ref[fin]) / np.abs(ref[fin])) <= 1e-12
    refl = om.logpdf(oc, orates(rates), per_family=True, precision="long")
    for mode, tol in (("fp32", 1e-6), ("fp64s", 1e-12)):
        got = w.logpdf_families(m, c, mode=mode)
        err = np.max(np.abs(got - refl) / np.abs(refl))
        print(mode, "max rel err vs long double", err)
        assert err <= tol


def test_wide_tree_gradient(w, wide):
    rates, om, oc, m, c = wide
    ov, og = om.gradient(oc, orates(rates), nthreads=8)
    for mode, tol in (("fp64", 1e-10), ("fp64s", 1e-10), ("fp32", 1e-5)):
        v, g, _ = w.gradient(m, c, mode=mode)
        err = np.max(np.abs(g - og)) / np.max(np.abs(og))
        print(mode, "grad rel err", err)
        assert abs(v - ov) <= (1e-12 if mode != "fp32" else 1e-6) * abs(ov)
        assert err <= tol


@pytest.fixture(scope="module")
def big4(w):
    """The 16 costliest of the first 400 families of the 10,000-family
    benchmark workload (50 taxa + 4 WGDs, Δt 0.01)."""
    from whale_jl_amd import synth
    wl = synth.workload("config4", n_families=400, nthreads=8)
    pick = np.argsort(-(wl.clades.astype(np.int64) * 1000 + wl.splits))[:16]
    texts = [wl.texts[i] for i in pick]
    om = OracleModel(wl.newick, [], wl.dt)
    oc = om.ccds(texts)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(texts, m)
    print("big4: clades", wl.clades[pick].min(), "..", wl.clades[pick].max())
    return wl, om, oc, m, c


def test_benchmark_largest_families(w, big4):
    wl, om, oc, m, c = big4
    ref = om.logpdf(oc, orates(wl.rates), per_family=True, precision="long")
    for mode, tol in (("fp64", 1e-12), ("fp64s", 1e-12), ("fp32", 1e-6)):
        got = w.logpdf_families(m, c, mode=mode)
        fin = np.isfinite(ref)
        assert np.array_equal(fin, np.isfinite(got)), mode
        err = np.max(np.abs(got[fin] - ref[fin]) / np.abs(ref[fin]))
        print(mode, "max rel err vs long double", err)
        assert err <= tol
    v32, g32, _ = w.gradient(m, c, mode="fp32")
    v64, g64, _ = w.gradient(m, c, mode="fp64s")
    assert np.max(np.abs(g32 - g64)) <= 1e-5 * np.max(np.abs(g64))
