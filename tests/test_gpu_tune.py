"""The launch-shape autotuner (whale_batch_autotune): it searches the size
classes' LDS bounds, the streams' class layout and the waves per compute unit
on the batch itself, keeps results equal to the untuned ones, and caches its
choice in a file keyed by device and workload."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg5(w):
    from whale_jl_amd import synth
    wl = synth.workload("config5", n_families=2000)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    return m, c


def test_autotune_keeps_results_and_caches(w, cfg5, tmp_path):
    m, c = cfg5
    b = w.device_batch(m, c)
    default = b.tuning()
    assert default == [10240, 14336, 20480, 28672, 49152, 0, 8]
    f64 = w.logpdf_families(m, c, mode="fp64s")
    v0, g0, _ = w.gradient(m, c, mode="fp32")
    cache = str(tmp_path / "tuning.txt")
    t = b.autotune("fp32", gradient=True, reps=3, cache=cache)
    print("tuned:", t)
    assert t["candidates_timed"] > 10
    assert t["ms_after"] <= t["ms_before"]
    assert b.tuning() == t["class_bounds"] + [t["layout"], t["waves_per_cu"]]
    # the shape changes the schedule, not the values
    assert np.array_equal(w.logpdf_families(m, c, mode="fp64s"), f64)
    v1, g1, _ = w.gradient(m, c, mode="fp32")
    assert abs(v1 - v0) <= 1e-9 * abs(v0)
    assert np.max(np.abs(g1 - g0)) <= 1e-6 * np.max(np.abs(g0))
    # a second search on the same workload comes from the cache
    b.set_tuning(default)
    assert b.tuning() == default
    t2 = b.autotune("fp32", gradient=True, reps=2, cache=cache)
    assert t2["candidates_timed"] == 0
    assert t2["class_bounds"] == t["class_bounds"] and t2["layout"] == t["layout"]
    assert open(cache).read().count("\n") == 1
    b.set_tuning(default)


def test_tuning_shapes_give_equal_values(w, cfg5):
    """Every stream layout and a few class bounds evaluate the same values."""
    m, c = cfg5
    b = w.device_batch(m, c)
    ref = w.logpdf_families(m, c, mode="fp64s")
    for layout in range(6):
        b.set_tuning([8192, 12288, 16384, 24576, 40960, layout, 6 + layout % 3])
        assert np.array_equal(w.logpdf_families(m, c, mode="fp64s"), ref)
    with pytest.raises(Exception):
        b.set_tuning([8192, 8192, 16384, 24576, 40960, 0, 8])  # bounds must increase
    b.set_tuning([10240, 14336, 20480, 28672, 49152, 0, 8])
