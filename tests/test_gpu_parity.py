"""HIP path vs the CPU oracle on seeded synthetic families."""
import numpy as np
import pytest

from oracle.oracle import OracleModel, DLWGD as ODLWGD

pytestmark = pytest.mark.gpu


def orates(r):
    return ODLWGD(list(r.λ), list(r.μ), list(r.q), r.η, list(r.p))


@pytest.fixture(scope="module")
def cfg1(w):
    from whale_jl_amd import synth
    wl = synth.workload("config1")
    om = OracleModel(wl.newick, [], wl.dt)
    oc = om.ccds(wl.texts)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    return wl, om, oc, m, c


def test_config1_fp64_per_family(w, cfg1):
    wl, om, oc, m, c = cfg1
    ref = om.logpdf(oc, orates(wl.rates), per_family=True)
    got = w.logpdf_families(m, c)
    rel = np.abs(got - ref) / np.abs(ref)
    print("config1 fp64 max rel", rel.max())
    assert rel.max() <= 1e-12


def test_config1_batch_value(w, cfg1):
    wl, om, oc, m, c = cfg1
    ref = om.logpdf(oc, orates(wl.rates))
    got = w.logpdf(m, c)
    assert abs(got - ref) <= 1e-12 * abs(ref)


def test_config1_fp32_scaled(w, cfg1):
    wl, om, oc, m, c = cfg1
    ref = om.logpdf(oc, orates(wl.rates), per_family=True, precision="long")
    got = w.logpdf_families(m, c, mode="fp32")
    rel = np.abs(got - ref) / np.abs(ref)
    print("config1 fp32 max rel", rel.max())
    assert rel.max() <= 1e-6
    got = w.logpdf_families(m, c, mode="fp64s")
    assert (np.abs(got - ref) / np.abs(ref)).max() <= 1e-12


def test_config1_per_family_launch_matches(w, cfg1):
    wl, om, oc, m, c = cfg1
    a = w.logpdf_families(m, c)
    b = w.logpdf_families(m, c, launch="per_family")
    assert np.array_equal(a, b)


def test_config1_gradient_vs_dual(w, cfg1):
    wl, om, oc, m, c = cfg1
    ov, og = om.gradient(oc, orates(wl.rates), nthreads=8)
    for mode, tol in (("fp64", 1e-10), ("fp64s", 1e-10), ("fp32", 1e-5)):
        v, g, _ = w.gradient(m, c, mode=mode)
        err = np.max(np.abs(g - og)) / np.max(np.abs(og))
        print(mode, "value", v, ov, "grad rel err", err)
        assert abs(v - ov) <= (1e-12 if mode != "fp32" else 1e-6) * abs(ov)
        assert err <= tol


def test_graph_replays_match_eager(w, cfg1):
    """The batched evaluation is a captured HIP graph unless per-kernel timing is
    on; replays interleaved with eager runs must agree with them every time."""
    from whale_jl_amd import _lib
    wl, om, oc, m, c = cfg1
    ctx = w.context()
    results = []
    for rep in range(4):
        for timing in (0, 1):
            _lib.lib().whale_ctx_set_timing(ctx, timing)
            for mode in ("fp64", "fp32"):
                v, g, fam = w.gradient(m, c, mode=mode)
                results.append((rep, timing, mode, v, g, fam))
    _lib.lib().whale_ctx_set_timing(ctx, 0)
    # gradient sums are atomic (order varies run to run); FP32 runs accumulate
    # per-stage partials in FP32 before the FP64 sum, so they vary a little more
    for mode, gtol in (("fp64", 1e-9), ("fp32", 1e-7)):
        rs = [r for r in results if r[2] == mode]
        base = rs[0]
        for r in rs[1:]:
            assert np.array_equal(r[5], base[5])  # per-family logL: no atomics, bit-identical
            assert abs(r[3] - base[3]) <= 1e-12 * abs(base[3])
            assert np.max(np.abs(r[4] - base[4])) <= gtol * np.max(np.abs(base[4]))
