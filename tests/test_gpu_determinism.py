"""Deterministic gradients (whale_batch_set_deterministic): repeated
logL + gradient evaluations on the same parameters are bit-identical, as an
HMC sampler's reversible trajectories need, in every form, under the ℓ
budget and for families in global memory; and they agree with the default
(atomic) accumulation and with the oracle's dual-number gradient."""
import os
import time

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.oracle import OracleModel, DLWGD as ODLWGD, ConstantDLWGD as OConst

pytestmark = pytest.mark.gpu


def orates(r):
    return ODLWGD(list(r.λ), list(r.μ), list(r.q), r.η, list(r.p))


@pytest.fixture(scope="module")
def cfg5(w):
    from whale_jl_amd import synth
    wl = synth.workload("config5", n_families=1500)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    return wl, m, c


def timed_grads(w, m, c, mode, k):
    out, t = [], time.perf_counter()
    for _ in range(k):
        out.append(w.gradient(m, c, mode=mode))
    return out, (time.perf_counter() - t) / k * 1e3


def test_bit_identical_gradients(w, cfg5):
    wl, m, c = cfg5
    b = w.device_batch(m, c)
    om = OracleModel(wl.newick, [], wl.dt)
    ov, og = om.gradient(om.ccds(wl.texts), orates(wl.rates), nthreads=16)
    for mode, tol in (("fp64s", 1e-10), ("fp32", 1e-5)):
        b.set_deterministic(False)
        free, t_free = timed_grads(w, m, c, mode, 4)
        b.set_deterministic(True)
        det, t_det = timed_grads(w, m, c, mode, 4)
        print(mode, f"default {t_free:.2f} ms, deterministic {t_det:.2f} ms per call;",
              "default calls bit-identical:", all(np.array_equal(free[0][1], x[1]) for x in free[1:]))
        for v, g, fam in det[1:]:
            assert v == det[0][0] and np.array_equal(g, det[0][1]) and np.array_equal(fam, det[0][2])
        assert np.max(np.abs(det[0][1] - free[0][1])) <= (1e-6 if mode == "fp32" else 1e-10) * np.max(np.abs(og))
        assert np.max(np.abs(det[0][1] - og)) <= tol * np.max(np.abs(og))
        assert abs(det[0][0] - ov) <= (1e-6 if mode == "fp32" else 1e-12) * abs(ov)
    b.set_deterministic(False)


def test_bit_identical_under_budget(w, cfg5):
    wl, m, c = cfg5
    b = w.device_batch(m, c)
    b.set_deterministic(True)
    b.set_ell_budget(8 << 20)
    g = [w.gradient(m, c, mode="fp32")[1] for _ in range(3)]
    assert all(np.array_equal(g[0], x) for x in g[1:])
    b.set_ell_budget(0)
    g2 = w.gradient(m, c, mode="fp32")[1]
    assert np.max(np.abs(g2 - g[0])) <= 1e-6 * np.max(np.abs(g2))
    b.set_deterministic(False)


def test_bit_identical_global_memory_and_routed_families(w):
    tree = open(os.path.join(GOLDEN, "example-4", "tree.nw")).read().strip()
    ale = open(os.path.join(GOLDEN, "example-4", "cytp450.ale")).read()
    m = w.WhaleModel(w.ConstantDLWGD(λ=0.1, μ=0.2, η=1 / 1.5), w.readnw(tree), 0.1)
    c = w.ccds_from_texts([ale, ale], m)
    b = w.device_batch(m, c)
    b.set_deterministic(True)
    assert b.hbm_families() == 2
    runs = [w.gradient(m, c, mode=mode) for mode in ("fp32", "fp32", "fp32", "fp64s", "fp64s")]
    assert np.array_equal(runs[1][1], runs[2][1]) and np.array_equal(runs[3][1], runs[4][1])
    om = OracleModel(tree, [], 0.1)
    ov, og = om.gradient(om.ccds([ale, ale]), OConst(0.1, 0.2, [], 1 / 1.5))
    for v, g, _ in runs:
        assert np.max(np.abs(g - og)) <= 1e-5 * np.max(np.abs(og))
