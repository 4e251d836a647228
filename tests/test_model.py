"""Model-side host logic that needs no GPU: setsamplingp! (src/model.jl:248-255)
as the reference tests it (test/runtests.jl:81-93: each leaf's getθ gives the
p set for its name)."""
import numpy as np

from conftest import EXTREE


def test_setsamplingp_sets_each_leaf(w):
    t = w.readnw(EXTREE)
    n = len(w.postwalk(t))
    m = w.WhaleModel(w.DLWGD(λ=[0.0] * n, μ=[0.0] * n, η=0.9), t, 0.05)
    rng = np.random.default_rng(4)
    leaves = [x for x in m.s.nodes if x.isleaf()]
    d = {x.name: float(rng.random()) for x in leaves}
    m.setsamplingp(d)
    lam, mu, q, eta, p, _ = m.node_params()
    for x in leaves:
        assert p[x.id - 1] == d[x.name]
    # a leaf missing from the dictionary gets 0, inner nodes have none
    del d[leaves[0].name]
    m.setsamplingp(d)
    p = m.node_params()[4]
    assert p[leaves[0].id - 1] == 0.0
    assert all(p[x.id - 1] == 0.0 for x in m.s.nodes if not x.isleaf())
