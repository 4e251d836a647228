"""The reference's known answers (test/runtests.jl:6-50) through the HIP path."""
import math

import numpy as np
import pytest

from conftest import EXTREE, KAT1, KAT2, KAT_INSERTS, MULTREE, GOLDEN, example1_texts

pytestmark = pytest.mark.gpu


def kat_model(w, maxn, lam=None, mu=None):
    t = w.readnw(EXTREE)
    n = len(w.postwalk(t))
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    lam = np.ones(n) if lam is None else lam
    mu = np.ones(n) if mu is None else mu
    r = w.DLWGD(λ=list(lam), μ=list(mu), q=[0.2, 0.1], η=0.9)
    return w.WhaleModel(r, t, 0.05, maxn=maxn), n


def test_kat1_single_family_unconditioned(w):
    m, _ = kat_model(w, maxn=5)
    ccd = w.read_ale(GOLDEN + "/example-1", m)
    l = w.logpdf(m, ccd[0])
    print("KAT-1", repr(l))
    assert abs(l - KAT1) <= 1e-12 * abs(KAT1)


def test_kat2_batch_conditioned(w):
    m, _ = kat_model(w, maxn=10000)
    ccd = w.read_ale(GOLDEN + "/example-1", m)
    a = w.logpdf_(m, ccd)
    b = w.logpdf(m, ccd)
    print("KAT-2", repr(a), repr(b))
    assert abs(a - KAT2) <= 1e-12 * abs(KAT2)
    assert abs(b - KAT2) <= 1e-12 * abs(KAT2)


def test_kat2_fp32_scaled(w):
    m, _ = kat_model(w, maxn=10000)
    ccd = w.read_ale(GOLDEN + "/example-1", m)
    v = w.logpdf(m, ccd, mode="fp32")
    print("KAT-2 fp32", repr(v), abs(v - KAT2) / abs(KAT2))
    assert abs(v - KAT2) <= 1e-6 * abs(KAT2)
    v = w.logpdf(m, ccd, mode="fp64s")
    assert abs(v - KAT2) <= 1e-12 * abs(KAT2)


def test_kat3_root_rates_irrelevant(w):
    m, n = kat_model(w, maxn=10000)
    lam = np.ones(n)
    mu = np.ones(n)
    root = m.root().id
    lam[root - 1] = math.nan
    mu[root - 1] = math.nan
    m3, _ = kat_model(w, maxn=10000, lam=lam, mu=mu)
    ccd = w.read_ale(GOLDEN + "/example-1", m3)
    v = w.logpdf_(m3, ccd)
    assert abs(v - KAT2) <= 1e-12 * abs(KAT2)


def test_kat_gradient_finite(w):
    m, n = kat_model(w, maxn=10000)
    ccd = w.read_ale(GOLDEN + "/example-1", m)
    x = np.concatenate([np.ones(n), np.ones(n), [0.1, 0.2], [0.8]])
    m2 = m(w.DLWGD(λ=list(x[:n]), μ=list(x[n:2 * n]), q=list(x[2 * n:2 * n + 2]), η=x[-1]))
    v, g, _ = w.gradient(m2, ccd)
    assert np.all(np.isfinite(g))
    assert g[n - 1] == 0.0 and g[2 * n - 1] == 0.0  # root rates have no effect


def test_mul_tree_finite(w):
    t = w.readnw(MULTREE)
    n = len(w.postwalk(t))
    r = w.DLWGD(λ=list(np.zeros(n) - 1), μ=list(np.zeros(n) - 1), η=0.9)
    m = w.WhaleModel(r, t, 0.05)
    ccd = w.read_ale(GOLDEN + "/example-1", m)
    assert math.isfinite(w.logpdf_(m, ccd))
