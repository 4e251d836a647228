"""The C-ABI library: loads, exports every declared symbol, and its host-side
ALE ingest (whale_ccd_parse) agrees with the oracle's parser.  No GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import EXTREE, GOLDEN, KAT_INSERTS, MULTREE, ROOT, example1_texts
from oracle.oracle import OracleModel


def header_functions():
    text = open(os.path.join(ROOT, "include", "whale_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(whale_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function(w):
    from whale_jl_amd import _lib
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_last_error_is_empty_string_initially(w):
    from whale_jl_amd import _lib
    assert isinstance(_lib.lib().whale_last_error(), bytes)


def compare_parse(w, om, texts, spmap):
    oc = om.ccds(texts)
    for i, t in enumerate(texts):
        ours = w.CCD.from_text(t, spmap)
        ref = oc.export(i)
        assert np.array_equal(ours.size, ref["size"])
        leaf = ours.size == 1
        # oracle species are 1-based node ids
        assert np.array_equal(ours.leaf_species[leaf] + 1, ref["leaf_species"][leaf])
        assert np.array_equal(ours.split_off, ref["split_off"])
        assert np.array_equal(ours.g1 + 1, ref["g1"]) and np.array_equal(ours.g2 + 1, ref["g2"])
        assert np.array_equal(ours.p, ref["p"])  # same division, bit for bit


def kat_structure(w, tree=EXTREE, inserts=KAT_INSERTS):
    t = w.readnw(tree)
    for name, a, b in inserts:
        w.insertnode(w.getlca(t, a, b), name=name)
    n = len(w.postwalk(t))
    return w.WhaleModel(w.DLWGD(λ=[0.0] * n, μ=[0.0] * n, q=[0.2] * len(inserts), η=0.9), t, 0.05)


def test_parse_example1_matches_oracle(w):
    m = kat_structure(w)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    compare_parse(w, om, example1_texts(), m.s.spmap)


def test_parse_example1_on_mul_tree_matches_oracle(w):
    m = kat_structure(w, MULTREE, [])
    om = OracleModel(MULTREE, [], 0.05)
    compare_parse(w, om, example1_texts(), m.s.spmap)


def test_parse_synthetic_matches_oracle(w):
    from whale_jl_amd import synth
    for name, nf in (("config1", 20), ("config2", 20), ("config4", 40)):
        wl = synth.workload(name, n_families=nf)
        m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
        om = OracleModel(wl.newick, [], wl.dt)
        compare_parse(w, om, wl.texts, m.s.spmap)


def test_model_structure_matches_oracle(w):
    m = kat_structure(w)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    ref = {d["id"]: d for d in om.nodes()}
    for d in m.s.nodes:
        r = ref[d.id]
        assert d.name == r["name"] and d.k == r["k"] and d.wgdid == r["wgdid"]
        assert d.delta == r["delta"] and d.leafP == r["leafP"]
        assert d.parent == r["parent"]
    assert m.s.order == om.order()


def test_parse_errors_raise(w):
    m = kat_structure(w)
    with pytest.raises(ValueError):
        w.CCD.from_text("#observations\n1\n#END\n", m.s.spmap)
    bad = example1_texts()[0].replace("ATHA_", "XXXX_")
    with pytest.raises(ValueError):
        w.CCD.from_text(bad, m.s.spmap)


def test_parse_many_matches_one_by_one(w):
    """whale_ccd_parse_many, at any thread count, returns the families that
    whale_ccd_parse returns one by one, in order.  When texts do not parse,
    the error names the lowest such file and no family is returned."""
    from whale_jl_amd import synth
    from whale_jl_amd.ccd import parse_many
    wl = synth.workload("config4", n_families=40)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    texts = list(wl.texts) + [wl.texts[3]]  # and a repeated family
    names = [f"fam{i}" for i in range(len(texts))]
    one = [w.CCD.from_text(t, m.s.spmap) for t in texts]
    for nt in (0, 1, 3, 64):
        many = parse_many(texts, m.s.spmap, names, n_threads=nt)
        assert [c.fname for c in many] == names
        for a, b in zip(many, one, strict=True):
            for k in ("size", "leaf_species", "split_off", "g1", "g2", "p"):
                assert np.array_equal(getattr(a, k), getattr(b, k))
    assert parse_many([], m.s.spmap, []) == []
    bad = list(texts)
    bad[19] = bad[19].replace("#Bip_counts", "#Bip_count")
    bad[7] = bad[7].replace("#Dip_counts", "#Dip_count")
    with pytest.raises(ValueError, match=r"read_ale fam7: family 7: ALE section missing: Dip_counts"):
        parse_many(bad, m.s.spmap, names, n_threads=4)


def test_read_ale_directory_matches_one_by_one(w):
    """read_ale on a directory (files sorted by name, parsed together)
    equals from_text on each file."""
    m = kat_structure(w)
    d = os.path.join(GOLDEN, "example-1")
    files = sorted(os.listdir(d))
    got = w.read_ale(d, m)
    assert [c.fname for c in got] == files
    for c, f in zip(got, files, strict=True):
        with open(os.path.join(d, f)) as fh:
            ref = w.CCD.from_text(fh.read(), m.s.spmap)
        for k in ("size", "leaf_species", "split_off", "g1", "g2", "p"):
            assert np.array_equal(getattr(c, k), getattr(ref, k))


def test_workload_is_deterministic(w):
    from whale_jl_amd import synth
    a = synth.workload("config4", n_families=30, nthreads=3)
    b = synth.workload("config4", n_families=30, nthreads=7)
    assert a.newick == b.newick and a.texts == b.texts
    assert a.leaves.min() >= 4 and a.leaves.max() <= 200
