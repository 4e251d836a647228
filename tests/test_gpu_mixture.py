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


def test_family_weights_select_families(w, cfg2):
    """The reverse sweep's family weights: all ones give the plain gradient;
    a weight vector e_k gives the gradient of family k alone (conditioning
    included once); weights scale linearly."""
    wl, om, oc, m, c, comps = cfg2
    n = len(c)
    _, g64, _ = w.gradient(m, c, mode="fp64")
    for mode, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        v, g, fam = w.gradient(m, c, mode=mode)
        v1, g1, fam1 = w.gradient(m, c, mode=mode, weights=np.ones(n))
        assert v1 == v and np.array_equal(fam1, fam)
        # (FP32 accumulators are summed in a run-dependent order)
        assert np.max(np.abs(g1 - g)) <= (1e-12 if mode == "fp64" else 1e-6) * np.max(np.abs(g))
        for k in (0, 7, n - 1):
            e = np.zeros(n)
            e[k] = 1.0
            _, gk, _ = w.gradient(m, c, mode=mode, weights=e)
            _, gs, _ = w.gradient(m, w.ccds_from_texts([wl.texts[k]], m), mode=mode)
            err = np.max(np.abs(gk - gs)) / np.max(np.abs(gs))
            print(mode, "family", k, "weighted vs alone", err)
            assert err <= tol
        wt = np.linspace(0.0, 2.0, n)
        _, ga, _ = w.gradient(m, c, mode=mode, weights=wt)
        _, gb, _ = w.gradient(m, c, mode=mode, weights=3.0 * wt)
        assert np.max(np.abs(gb - 3.0 * ga)) <= 1e-6 * np.max(np.abs(gb))
    # weights never change later unweighted calls
    _, g2, _ = w.gradient(m, c, mode="fp64")
    assert np.max(np.abs(g2 - g64)) <= 1e-12 * np.max(np.abs(g64))


def test_mixture_gradient_against_oracle(w, cfg2):
    """d logpdf(MixtureModel)/dθ_j = Σ_i r_ij ∂(ℓ_ij − c_j) from per-family
    dual-number oracle gradients, and d/dprior_j = Σ_i r_ij / prior_j against
    central differences of the mixture value."""
    wl, om, oc, m, c, comps = cfg2
    nsub = 16
    texts = wl.texts[:nsub]
    cs = w.ccds_from_texts(texts, m)
    prior = [0.5, 0.3, 0.2]
    mm = w.MixtureModel([m(r) for r in comps], prior)
    tot, R = mm.responsibilities(cs)
    assert np.allclose(R.sum(axis=1), 1.0)
    for mode, tol in (("fp64", 1e-9), ("fp32", 1e-5)):
        v, grads, dprior = mm.gradient(cs, mode=mode)
        assert abs(v - tot) <= (1e-12 if mode == "fp64" else 1e-6) * abs(tot)
        for j, r in enumerate(comps):
            ref = np.zeros_like(grads[j])
            for i, t in enumerate(texts):
                _, gi = om.gradient(om.ccds([t]), orates(r))
                ref += R[i, j] * gi
            err = np.max(np.abs(grads[j] - ref)) / np.max(np.abs(ref))
            print(mode, "component", j, "mixture gradient rel err", err)
            assert err <= tol
    h = 1e-6
    for j in range(3):
        pp, pm = list(prior), list(prior)
        pp[j] += h
        pm[j] -= h
        fd = (w.MixtureModel(mm.components, pp).logpdf(cs) - w.MixtureModel(mm.components, pm).logpdf(cs)) / (2 * h)
        print("prior", j, dprior[j], fd)
        assert abs(dprior[j] - fd) <= 1e-6 * max(1.0, abs(fd))


def test_device_mixture_matches_per_component_path(w, cfg2):
    """whale_mixture_eval (all components in each size-class launch, the
    log-sum-exp and responsibilities on the device, one weighted reverse
    launch per class) against the per-component path (one evaluation per
    component combined on the host, one weighted reverse sweep each)."""
    wl, om, oc, m, c, comps = cfg2
    prior = [0.5, 0.3, 0.2]
    mm = w.MixtureModel([m(r) for r in comps], prior)
    assert mm.on_device(c)
    for mode, vtol, gtol in (("fp64", 1e-12, 1e-10), ("fp32", 1e-6, 1e-5)):
        F = mm.family_matrix(c, mode)
        tot_h, R_h = mm._host_resp(c, mode)
        tot, fam, R, g, dprior = mm._device_eval(c, mode, True)
        fam_ref = np.stack([w.logpdf_families(wm, c, mode=mode) for wm in mm.components], axis=1)
        assert np.max(np.abs(fam - fam_ref) / np.abs(fam_ref)) <= (1e-13 if mode == "fp64" else 1e-6)
        assert abs(tot - tot_h) <= vtol * abs(tot_h)
        assert np.max(np.abs(R - R_h)) <= (1e-10 if mode == "fp64" else 1e-5)
        v, grads, dp = mm.gradient(c, mode=mode)
        vh, grads_h, dp_h = mm.gradient(c, mode=mode, device=False)
        for j in range(3):
            err = np.max(np.abs(grads[j] - grads_h[j])) / np.max(np.abs(grads_h[j]))
            print(mode, "component", j, "device vs per-component gradient", err)
            assert err <= gtol
        assert np.max(np.abs(dp - dp_h) / np.abs(dp_h)) <= (1e-10 if mode == "fp64" else 1e-5)
    # one component with weight one is the model itself
    one = w.MixtureModel([m(comps[0])], [1.0])
    v1, g1, _ = one.gradient(c, mode="fp64")
    v0, g0, _ = w.gradient(m(comps[0]), c, mode="fp64")
    assert abs(v1 - v0) <= 1e-12 * abs(v0)
    assert np.max(np.abs(g1[0] - g0)) <= 1e-10 * np.max(np.abs(g0))


def test_device_mixture_fp32_overflow_and_routed(w):
    """Components on a batch holding the family FP32 cannot hold: the first
    mixture call recomputes it per component in FP64 (each component's own
    flagged list); after a single-model FP32 call has routed it to FP64, the
    mixture's routed launch covers it for every component."""
    from ale_util import ale_from_tree, balanced_paralogs
    tree, dt = "((A:0.03,B:1.0):1.0,C:1.03);", 0.005
    lam, mu = [np.log(0.8)] * 5, [np.log(0.3)] * 5
    mu[1] = np.log(3.0)
    big = ale_from_tree(*balanced_paralogs(64, "A", ("B_1", "C_1")))
    small = [ale_from_tree(*balanced_paralogs(k, "A", ("B_1", "C_1"))) for k in (1, 2, 3, 5, 8)]
    texts = small[:2] + [big] + small[2:] + [big]
    m = w.WhaleModel(w.DLWGD(λ=lam, μ=mu, η=0.9), w.readnw(tree), dt)
    c = w.ccds_from_texts(texts, m)
    lam2 = [x + 0.05 for x in lam]
    mm = w.MixtureModel([m, m(w.DLWGD(λ=lam2, μ=mu, η=0.85))], [0.6, 0.4])
    for call in range(2):
        v64, g64, d64 = mm.gradient(c, mode="fp64")
        v32, g32, d32 = mm.gradient(c, mode="fp32")
        print("call", call, v64, v32)
        assert abs(v32 - v64) <= 1e-6 * abs(v64)
        for a, b in zip(g32, g64):
            assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(b))
        assert np.max(np.abs(d32 - d64) / np.abs(d64)) <= 1e-5
        w.gradient(m, c, mode="fp32")  # flags the overflowing families: later calls route them to FP64
    assert w.device_batch(m, c).fallbacks() == 2
