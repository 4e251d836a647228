"""The benchmark workloads at their own sizes against the CPU oracle:

  config2  all 1,000 families in one batched launch (FP64 ≤ 1e-12, FP32 ≤ 1e-6,
           gradients against the dual-number oracle);
  config4  the gradient of 32 families (the 16 costliest of the first 400 and
           16 others) against the dual-number oracle, 203 parameters;
  config5  the sampler-style loop: 5,000 device-resident families evaluated at
           several log-normally perturbed parameter points back to back, each
           value and gradient against the oracle.
"""
import numpy as np
import pytest

from oracle.oracle import OracleModel, DLWGD as ODLWGD

pytestmark = pytest.mark.gpu

NCPU = 16


def orates(r):
    return ODLWGD(list(r.λ), list(r.μ), list(r.q), r.η, list(r.p))


def test_config2_full_batch(w):
    from whale_jl_amd import synth
    wl = synth.workload("config2")
    assert len(wl.texts) == 1000
    om = OracleModel(wl.newick, [], wl.dt)
    oc = om.ccds(wl.texts)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    ref = om.logpdf(oc, orates(wl.rates), per_family=True, nthreads=NCPU)
    refl = om.logpdf(oc, orates(wl.rates), per_family=True, precision="long", nthreads=NCPU)
    fin = np.isfinite(ref)
    got = w.logpdf_families(m, c)
    assert np.array_equal(fin, np.isfinite(got))
    assert np.max(np.abs(got[fin] - ref[fin]) / np.abs(ref[fin])) <= 1e-12
    for mode, tol in (("fp32", 1e-6), ("fp64s", 1e-12)):
        got = w.logpdf_families(m, c, mode=mode)
        err = np.max(np.abs(got - refl) / np.abs(refl))
        print(mode, "config2 1000 families, max rel err vs long double", err)
        assert err <= tol
    ov, og = om.gradient(oc, orates(wl.rates), nthreads=NCPU)
    for mode, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        v, g, _ = w.gradient(m, c, mode=mode)
        err = np.max(np.abs(g - og)) / np.max(np.abs(og))
        print(mode, "config2 grad rel err", err)
        assert abs(v - ov) <= (1e-12 if mode == "fp64" else 1e-6) * abs(ov)
        assert err <= tol


def test_config4_gradient_vs_dual_oracle(w):
    from whale_jl_amd import synth
    wl = synth.workload("config4", n_families=400, nthreads=8)
    cost = wl.clades.astype(np.int64) * 1000 + wl.splits
    big = list(np.argsort(-cost)[:16])
    rest = [i for i in range(0, 400, 23) if i not in big][:16]
    texts = [wl.texts[i] for i in big + rest]
    om = OracleModel(wl.newick, [], wl.dt)
    oc = om.ccds(texts)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(texts, m)
    ov, og = om.gradient(oc, orates(wl.rates), nthreads=NCPU)
    assert len(og) == 99 + 99 + 4 + 1
    for mode, vtol, gtol in (("fp64s", 1e-12, 1e-10), ("fp32", 1e-6, 1e-5)):
        v, g, _ = w.gradient(m, c, mode=mode)
        err = np.max(np.abs(g - og)) / np.max(np.abs(og))
        print(mode, "config4 32 families: logL rel err", abs(v - ov) / abs(ov), "grad rel err", err)
        assert abs(v - ov) <= vtol * abs(ov)
        assert err <= gtol


def test_config5_sampler_points(w):
    """5,000 device-resident families at three of the 256 sampler points: the
    per-family values of the whole batch and the value and gradient of a
    1,000-family sub-batch against the oracle (which parses only the
    sub-batch: its ALE ingest is the reference's quadratic one)."""
    from whale_jl_amd import synth
    wl = synth.workload("config5")
    assert len(wl.texts) == 5000
    sub = list(range(0, 5000, 5))
    om = OracleModel(wl.newick, [], wl.dt)
    oc = om.ccds([wl.texts[i] for i in sub])
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c_all = w.ccds_from_texts(wl.texts, m)
    c_sub = w.ccds_from_texts([wl.texts[i] for i in sub], m)
    points = synth.perturbed(wl.rates, 256)
    for i in (0, 97, 255):
        r = points[i]
        mi = m(r)
        ref = om.logpdf(oc, orates(r), per_family=True, nthreads=NCPU)
        refl = om.logpdf(oc, orates(r), per_family=True, precision="long", nthreads=NCPU)
        for mode, want, tol in (("fp64", ref, 1e-12), ("fp32", refl, 1e-6)):
            fam = w.logpdf_families(mi, c_all, mode=mode)[sub]
            err = np.max(np.abs(fam - want) / np.abs(want))
            print(f"config5 point {i} {mode}: 5000-family batch, max rel err on the sub-batch {err:.2e}")
            assert err <= tol
        ov, og = om.gradient(oc, orates(r), nthreads=NCPU)
        for mode, vtol, gtol in (("fp64", 1e-12, 1e-10), ("fp32", 1e-6, 1e-5)):
            v, g, _ = w.gradient(mi, c_sub, mode=mode)
            err = np.max(np.abs(g - og)) / np.max(np.abs(og))
            print(f"config5 point {i} {mode}: logL {v:.6f} oracle {ov:.6f} grad rel err {err:.2e}")
            assert abs(v - ov) <= vtol * abs(ov)
            assert err <= gtol
