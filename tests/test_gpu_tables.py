"""Device slice tables (ε, φ, ψ per branch row, src/model.jl:162-191) against
the oracle's row-by-row recursion.  The device evaluates every row in closed
form (the slice map is a Möbius map with fixed points 1 and μ/λ), so the
regimes where a closed form can lose digits are the cases: critical branches
(λ = μ, ε → 1), λ within 1e-6 of μ (the reference's near-critical formula;
just outside that band its α = μ(e^{(λ−μ)t} − 1)/(λe^{(λ−μ)t} − μ) cancels,
so the last ulp of exp in any two libraries moves the table by ~1e-9),
strongly sub- and supercritical rates over long branches, and leaves whose
sampling probability p lies above μ/λ (ε falls towards μ/λ)."""
import numpy as np
import pytest

from conftest import EXTREE, KAT_INSERTS
from oracle.oracle import OracleModel, DLWGD as ODLWGD

pytestmark = pytest.mark.gpu

N = 17


def _cases():
    rng = np.random.default_rng(5)
    yield "moderate", rng.normal(-0.4, 0.3, N), rng.normal(-0.7, 0.3, N), []
    lc = np.log(rng.uniform(2.0, 9.0, N))
    yield "critical", lc, lc.copy(), []
    yield "near-critical", lc, np.log(np.exp(lc) + 5e-7), []  # |λ − μ| ≤ 1e-6: the λ = μ formula
    yield "subcritical", np.log(rng.uniform(0.01, 0.1, N)), np.log(rng.uniform(3.0, 8.0, N)), []
    yield "supercritical", np.log(rng.uniform(3.0, 8.0, N)), np.log(rng.uniform(0.001, 0.01, N)), []
    yield ("p above mu/lambda", np.log(rng.uniform(2.0, 6.0, N)), np.log(rng.uniform(0.001, 0.005, N)),
           list(rng.uniform(0.3, 0.95, 9)))


@pytest.mark.parametrize("name,lam,mu,p", list(_cases()), ids=[c[0] for c in _cases()])
def test_tables_match_recursion(w, name, lam, mu, p):
    t = w.readnw(EXTREE)
    for nm, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=nm)
    dt = 0.01  # ≈ 500 rows on the longest branch
    m = w.WhaleModel(w.DLWGD(λ=list(lam), μ=list(mu), q=[0.35, 0.15], p=p, η=0.85), t, dt)
    om = OracleModel(EXTREE, KAT_INSERTS, dt)
    orr = ODLWGD(list(lam), list(mu), [0.35, 0.15], 0.85, p)
    tabs = m.tables()
    worst = 0.0
    for i, (e, f, s) in tabs.items():
        oe, of, os_ = om.tables(orr, i, len(e))
        assert len(oe) == len(e)
        for got, want in ((e, oe), (f, of), (s, os_)):
            ok = want != 0
            assert np.all(got[~ok] == 0)
            worst = max(worst, float(np.max(np.abs(got[ok] - want[ok]) / np.abs(want[ok]), initial=0.0)))
    print(name, "max rel diff of the tables", worst)
    # the recursion itself drifts by about one ulp per row over ≈500 rows
    assert worst <= 2e-13
