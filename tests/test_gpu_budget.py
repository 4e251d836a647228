"""Evaluations under a memory budget for stored ℓ (whale_batch_set_ell_budget):
each family's forward DP and reverse sweep run in one pass of one wave whose
slot holds only that family's rows, so ℓ memory no longer grows with the
batch.  At a forced small budget the value and gradient must equal the
stored form's and the oracle's, in every form, for families in LDS, in
global memory (4,917 clades), routed to FP64 and recomputed in FP64."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.oracle import OracleModel, DLWGD as ODLWGD, ConstantDLWGD as OConst

pytestmark = pytest.mark.gpu


def orates(r):
    return ODLWGD(list(r.λ), list(r.μ), list(r.q), r.η, list(r.p))


def test_budgeted_equals_stored_config2(w):
    from whale_jl_amd import synth
    wl = synth.workload("config2", n_families=300)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    b = w.device_batch(m, c)
    stored = {mode: (w.gradient(m, c, mode=mode), w.logpdf_families(m, c, mode=mode))
              for mode in ("fp64", "fp64s", "fp32")}
    full = b.ell_bytes()["stored_form"]
    budget = 4 << 20
    assert full > 4 * budget
    b.set_ell_budget(budget)
    om = OracleModel(wl.newick, [], wl.dt)
    oc = om.ccds(wl.texts)
    ov, og = om.gradient(oc, orates(wl.rates), nthreads=16)
    for mode in ("fp64", "fp64s", "fp32"):
        (sv, sg, sf), sfam = stored[mode]
        v, g, fam = w.gradient(m, c, mode=mode)
        lfam = w.logpdf_families(m, c, mode=mode)
        tol = 1e-6 if mode == "fp32" else 1e-12
        print(mode, "budgeted vs stored: grad", np.max(np.abs(g - sg)) / np.max(np.abs(sg)))
        assert np.array_equal(lfam, sfam) and np.array_equal(fam, sf)
        assert abs(v - sv) <= tol * abs(sv)
        assert np.max(np.abs(g - sg)) <= (1e-6 if mode == "fp32" else 1e-10) * np.max(np.abs(sg))
        assert np.max(np.abs(g - og)) <= (1e-5 if mode == "fp32" else 1e-10) * np.max(np.abs(og))
    eb = b.ell_bytes()
    print("bytes:", eb)
    assert eb["budgeted_slots"] <= 2 * budget and eb["held"] < full / 2
    # a call that keeps ℓ (logpdf!) stores every row again, then backtracking works
    w.logpdf_(m, c)
    assert len(w.backtrack(m, c, n=2, seed=1, families=[0, 1])) == 2
    b.set_ell_budget(0)
    v, g, _ = w.gradient(m, c, mode="fp64")
    assert np.max(np.abs(g - stored["fp64"][0][1])) <= 1e-10 * np.max(np.abs(g))


def test_budgeted_global_memory_family(w):
    tree = open(os.path.join(GOLDEN, "example-4", "tree.nw")).read().strip()
    ale = open(os.path.join(GOLDEN, "example-4", "cytp450.ale")).read()
    m = w.WhaleModel(w.ConstantDLWGD(λ=0.1, μ=0.2, η=1 / 1.5), w.readnw(tree), 0.1)
    c = w.ccds_from_texts([ale, ale], m)
    b = w.device_batch(m, c)
    assert b.hbm_families() == 2
    b.set_ell_budget(1 << 20)
    om = OracleModel(tree, [], 0.1)
    oc = om.ccds([ale, ale])
    ov, og = om.gradient(oc, OConst(0.1, 0.2, [], 1 / 1.5))
    for mode, tol in (("fp64s", 1e-10), ("fp32", 1e-5)):
        v, g, _ = w.gradient(m, c, mode=mode)
        assert abs(v - ov) <= tol * abs(ov)
        assert np.max(np.abs(g - og)) <= tol * np.max(np.abs(og))


def test_budgeted_routed_and_recomputed_families(w):
    """Families whose FP32 values leave the range: the first budgeted FP32
    call recomputes them in FP64, later calls route them; same results."""
    from ale_util import ale_from_tree, balanced_paralogs
    tree, dt = "((A:0.03,B:1.0):1.0,C:1.03);", 0.005
    lam, mu = [np.log(0.8)] * 5, [np.log(0.3)] * 5
    mu[1] = np.log(3.0)
    big = ale_from_tree(*balanced_paralogs(64, "A", ("B_1", "C_1")))
    small = [ale_from_tree(*balanced_paralogs(k, "A", ("B_1", "C_1"))) for k in (1, 2, 3, 5, 8)]
    texts = small[:2] + [big] + small[2:] + [big]
    m = w.WhaleModel(w.DLWGD(λ=lam, μ=mu, η=0.9), w.readnw(tree), dt)
    c = w.ccds_from_texts(texts, m)
    b = w.device_batch(m, c)
    v64, g64, f64 = w.gradient(m, c, mode="fp64s")
    b.set_ell_budget(64 << 10)
    for k in range(3):
        v, g, fam = w.gradient(m, c, mode="fp32")
        print("call", k, "fallbacks", b.fallbacks())
        assert b.fallbacks() == 2
        assert abs(fam[2] - f64[2]) <= 1e-12 * abs(f64[2])
        assert abs(v - v64) <= 1e-6 * abs(v64)
        assert np.max(np.abs(g - g64)) <= 1e-5 * np.max(np.abs(g64))
