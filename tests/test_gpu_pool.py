"""The pooled launch (one launch per evaluation; each compute unit's LDS is a
pool of pages, every family takes the pages its stages need and runs both
directions by one wave) against the oracle and against the size-class
launches, on mixed family sizes: values and gradients in every form, ℓ kept
for backtracking, families the FP32 form cannot hold (recomputed in FP64, then
routed), families whose stages take the global-memory path beside the pool,
and repeated graph replays."""
import numpy as np
import pytest

from ale_util import ale_from_tree, balanced_paralogs
from oracle.oracle import OracleModel, DLWGD as ODLWGD
from test_gpu_hbm import check_gradient, check_values, orates

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="the pooled launch is switched off: its first GPU run (the FP32 gradient of "
                                      "test_pool_matches_classes) ended in a memory fault with no identified cause")]

POOL = 6  # layout index of the pooled launch in the tuning vector


def layout(db, lay):
    t = db.tuning()
    t[5] = lay
    db.set_tuning(t)
    if lay == POOL:
        assert all(db.pooled(mode) for mode in ("fp32", "fp64s", "fp64"))


@pytest.fixture(scope="module")
def cfg4(w):
    from whale_jl_amd import synth
    wl = synth.workload("config4", n_families=400)
    om = OracleModel(wl.newick, [], wl.dt)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    return wl, m, c, om, om.ccds(wl.texts)


def test_pool_matches_classes(w, cfg4):
    wl, m, c, om, oc = cfg4
    db = w.device_batch(m, c)
    res = {}
    for lay in (0, POOL):
        layout(db, lay)
        res[lay] = {mode: w.gradient(m, c, mode=mode) for mode in ("fp32", "fp64s", "fp64")}
        res[lay]["logpdf"] = w.logpdf_families(m, c, mode="fp32")
    layout(db, 0)
    for mode in ("fp32", "fp64s", "fp64"):
        (v0, g0, f0), (v1, g1, f1) = res[0][mode], res[POOL][mode]
        gerr = np.max(np.abs(g1 - g0)) / np.max(np.abs(g0))
        ferr = np.max(np.abs(f1 - f0) / np.abs(f0))
        print(mode, "pooled vs classes: family", ferr, "gradient", gerr)
        assert ferr <= (1e-7 if mode == "fp32" else 1e-13)
        assert abs(v1 - v0) <= 1e-12 * abs(v0)
        assert gerr <= (1e-6 if mode == "fp32" else 1e-12)
    assert np.max(np.abs(res[POOL]["logpdf"] - res[0]["logpdf"]) / np.abs(res[0]["logpdf"])) <= 1e-7


def test_pool_against_oracle(w, cfg4):
    wl, m, c, om, oc = cfg4
    db = w.device_batch(m, c)
    layout(db, POOL)
    try:
        check_values(w, m, c, om, oc, orates(wl.rates))
        # the dual-number gradient on a subset of the families
        sub = list(range(0, len(c), 25))
        cs = w.ccds_from_texts([wl.texts[i] for i in sub], m)
        layout(w.device_batch(m, cs), POOL)
        check_gradient(w, m, cs, om, om.ccds([wl.texts[i] for i in sub]), orates(wl.rates))
    finally:
        layout(db, 0)


def test_pool_replays_and_kept_ell(w, cfg4):
    wl, m, c, om, oc = cfg4
    db = w.device_batch(m, c)
    layout(db, 0)
    w.logpdf_(m, c, mode="fp32")
    ell0 = [db.ell(f) for f in range(0, len(c), 40)]
    t0 = w.backtrack(m, c, n=4, seed=7, families=range(0, 16))
    layout(db, POOL)
    try:
        runs = [w.gradient(m, c, mode="fp32") for _ in range(3)]
        for v, g, fam in runs[1:]:
            assert np.array_equal(fam, runs[0][2])  # the forward per family is deterministic
            assert np.max(np.abs(g - runs[0][1])) <= 1e-6 * np.max(np.abs(runs[0][1]))
        w.logpdf_(m, c, mode="fp32")
        ell1 = [db.ell(f) for f in range(0, len(c), 40)]
        for a, b in zip(ell0, ell1):
            assert np.allclose(a, b, rtol=1e-6, atol=0)
        t1 = w.backtrack(m, c, n=4, seed=7, families=range(0, 16))
        same = sum(x.key() == y.key() for fa, fb in zip(t0, t1) for x, y in zip(fa, fb))
        print("backtracked trees equal under both layouts:", same, "of", 64)
        assert same >= 60  # equal ℓ, equal draws (an FP32 rounding tie may flip a rare draw)
    finally:
        layout(db, 0)


def test_pool_fp32_overflow_recomputed_then_routed(w):
    """The family that FP32 cannot hold (a clade growing by ~2^354 along one
    branch): the pooled FP32 pass flags it, the FP64 pass recomputes it, and
    later calls route it to FP64 beside the pool."""
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
    layout(db, POOL)
    v64, g64, f64 = w.gradient(m, c, mode="fp64s")
    for k in range(3):
        v, g, fam = w.gradient(m, c, mode="fp32")
        print("call", k, "fallbacks", db.fallbacks())
        assert db.fallbacks() == 2
        assert np.max(np.abs(fam - ref) / np.abs(ref)) <= 1e-6
        assert abs(v - v64) <= 1e-6 * abs(v64)
        assert np.max(np.abs(g - g64)) <= 1e-5 * np.max(np.abs(g64))


def test_pool_beside_global_memory_families(w):
    from whale_jl_amd import synth
    wl = synth.workload("config2", n_families=200)
    om = OracleModel(wl.newick, [], wl.dt)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    b = w.device_batch(m, c)
    b.set_lds_limit(24 * 1024)
    layout(b, POOL)
    try:
        assert 0 < b.hbm_families() < len(c)
        check_values(w, m, c, om, om.ccds(wl.texts), orates(wl.rates))
        check_gradient(w, m, c, om, om.ccds(wl.texts), orates(wl.rates))
    finally:
        b.set_lds_limit(0)
        layout(b, 0)
