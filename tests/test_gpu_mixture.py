"""MixtureModel and ModelArray likelihoods (src/core.jl:66-79) and the
sampling conditions (src/condition.jl:11-29), against the CPU oracle: the
mixture must equal Σ_i logsumexp_j (ℓ_ij + log p_j − c_j) built from K
single-model oracle runs; the model array the sum of single-family runs."""
import math

import numpy as np
import pytest

from conftest import unpack_rates
from oracle.oracle import OracleModel, DLWGD as ODLWGD

pytestmark = pytest.mark.gpu


def orates(r):
    return ODLWGD(list(r.λ), list(r.μ), list(r.q), r.η, list(r.p))


@pytest.fixture(scope="module")
def cfg2(w):
    from whale_jl_amd import synth
    wl = synth.workload("config2", n_families=60, nthreads=8)
    om = OracleModel(wl.newick, [], wl.dt)
    oc = om.ccds(wl.texts)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    comps = synth.perturbed(wl.rates, 3, sigma=0.3)
    return wl, om, oc, m, c, comps


@pytest.mark.parametrize("cond", ["root", "nonextinct", "none"])
def test_condition_matches_oracle(w, cond):
    from whale_jl_amd import synth
    wl = synth.workload("config2", n_families=4, nthreads=4)
    C = {"root": w.RootCondition(), "nonextinct": w.NonExtinctCondition(), "none": w.NoCondition()}[cond]
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt, condition=C)
    om = OracleModel(wl.newick, [], wl.dt, condition=cond)
    ref = om.condition_value(om.ccds(wl.texts), orates(wl.rates))
    got = w.condition(m)
    print(cond, got, ref)
    assert abs(got - ref) <= 1e-13 * max(1.0, abs(ref))
    # the batch value is Σ ℓ_i − N·c
    c = w.ccds_from_texts(wl.texts, m)
    fam = w.logpdf_families(m, c)
    assert abs(w.logpdf(m, c) - (fam.sum() - len(fam) * got)) <= 1e-12 * abs(fam.sum())


def test_mixture_equals_single_model_runs(w, cfg2):
    wl, om, oc, m, c, comps = cfg2
    prior = [0.5, 0.3, 0.2]
    mm = w.MixtureModel([m(r) for r in comps], prior)
    cols = []
    for r, p in zip(comps, prior):
        cols.append(om.logpdf(oc, orates(r), per_family=True) + math.log(p) - om.condition_value(oc, orates(r)))
    L = np.stack(cols, axis=1)
    mx = L.max(axis=1)
    ref = float(np.sum(mx + np.log(np.exp(L - mx[:, None]).sum(axis=1))))
    got = mm.logpdf(c)
    print("mixture", got, ref)
    assert abs(got - ref) <= 1e-12 * abs(ref)
    assert abs(mm.logpdf(c, mode="fp32") - ref) <= 1e-6 * abs(ref)
    # one component with weight one is the model itself
    one = w.MixtureModel([m(comps[0])], [1.0])
    assert abs(one.logpdf(c) - w.logpdf(m(comps[0]), c)) <= 1e-12 * abs(ref)


def test_model_array_sums_single_families(w, cfg2):
    wl, om, oc, m, c, comps = cfg2
    n = 12
    models = [m(comps[i % 3]) for i in range(n)]
    fams = [c[i] for i in range(n)]
    got = w.ModelArray(models).logpdf(fams)
    ref = sum(om.logpdf(oc, orates(comps[i % 3]), per_family=True)[i] for i in range(n))
    print("model array", got, ref)
    assert abs(got - ref) <= 1e-12 * abs(ref)


def test_nowhere_extinct_batch_value_and_gradient(w):
    """The batch value with NowhereExtinctCondition: Σ ℓ_i − N·c, c from the
    tree's pgf on the host (the per-family values are unconditioned), and its
    gradient."""
    from whale_jl_amd import synth
    wl = synth.workload("config1", n_families=10, nthreads=4)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt, condition=w.NowhereExtinctCondition())
    c = w.ccds_from_texts(wl.texts, m)
    fam = w.logpdf_families(m, c)
    om = OracleModel(wl.newick, [], wl.dt)
    ref_fam = om.logpdf(om.ccds(wl.texts), orates(wl.rates), per_family=True)
    assert np.max(np.abs(fam - ref_fam) / np.abs(ref_fam)) <= 1e-12
    cond = w.condition(m)
    assert np.isfinite(cond) and cond < 0
    assert abs(w.logpdf(m, c) - (fam.sum() - len(fam) * cond)) <= 1e-12 * abs(fam.sum())
    # gradient: the device's unconditioned gradient minus N·∂c (complex step
    # on the host), against central differences of the conditioned FP64 value
    v, g, _ = w.gradient(m, c)
    assert abs(v - w.logpdf(m, c)) <= 1e-12 * abs(v)
    th = m.rates.pack()
    fd = np.zeros_like(th)
    h = 1e-5
    for i in range(len(th)):
        up, dn = th.copy(), th.copy()
        up[i] += h
        dn[i] -= h
        fd[i] = (w.logpdf(m(unpack_rates(m.rates, up)), c) - w.logpdf(m(unpack_rates(m.rates, dn)), c)) / (2 * h)
    err = np.max(np.abs(g - fd)) / np.max(np.abs(fd))
    print("NowhereExtinct gradient vs central differences: rel err", err)
    assert err <= 1e-5
