"""whale_logpdf_noalloc: the batched logpdf with its workspace in the fixed
per-GPU device array instead of memory the library allocates.  Its results
must equal whale_logpdf's, call after call, on batches of different sizes
(the array keeps the previous call's contents), and a batch too large for the
array is refused."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import EXTREE, GOLDEN, KAT_INSERTS, example1_texts
from oracle.oracle import OracleModel, DLWGD as ODLWGD

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library itself is linked against (test-side
    allocations of the caller-owned output buffers)."""
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def noalloc(w, b, mode):
    """One whale_logpdf_noalloc call into caller-owned device buffers filled
    with NaN beforehand; returns (status, message, per-family, total)."""
    from whale_jl_amd import _lib
    hip = _hip()
    n = b.n
    nan = np.full(n + 1, math.nan)
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), 8 * (n + 1)) == 0
    try:
        assert hip.hipMemcpy(d, nan.ctypes.data, 8 * (n + 1), 1) == 0  # host to device
        rc = _lib.lib().whale_logpdf_noalloc(b.h, _lib.MODES[mode], C.c_void_p(d.value),
                                             C.c_void_p(d.value + 8 * n))
        msg = _lib.lib().whale_last_error().decode()
        assert hip.hipDeviceSynchronize() == 0
        out = np.zeros(n + 1)
        assert hip.hipMemcpy(out.ctypes.data, d, 8 * (n + 1), 2) == 0  # device to host
    finally:
        hip.hipFree(d)
    return rc, msg, out[:n], float(out[n])


def test_noalloc_matches_logpdf_on_successive_batches(w):
    t = w.readnw(EXTREE)
    for name, a, b in KAT_INSERTS:
        w.insertnode(w.getlca(t, a, b), name=name)
    rng = np.random.default_rng(17)
    lam, mu = list(rng.normal(-0.4, 0.3, 17)), list(rng.normal(-0.7, 0.3, 17))
    m = w.WhaleModel(w.DLWGD(λ=lam, μ=mu, q=[0.3, 0.15], η=0.85), t, 0.05)
    texts = example1_texts()
    big = w.ccds_from_texts(texts, m)
    small = w.ccds_from_texts(texts[:5], m)
    om = OracleModel(EXTREE, KAT_INSERTS, 0.05)
    orr = ODLWGD(lam, mu, [0.3, 0.15], 0.85)
    for mode in ("fp64", "fp32", "fp64s"):
        for ccds, ntx in ((big, 12), (small, 5), (big, 12)):  # the second call has a smaller input
            m._sync()
            b = w.device_batch(m, ccds)
            ref_fam = w.logpdf_families(m, ccds, mode=mode)
            ref_tot = w.logpdf(m, ccds, mode=mode)
            rc, msg, fam, tot = noalloc(w, b, mode)
            assert rc == 0, msg
            tol = 1e-12 if mode != "fp32" else 1e-6
            print(mode, ntx, "families: max rel diff", np.max(np.abs(fam - ref_fam) / np.abs(ref_fam)))
            assert np.all(np.abs(fam - ref_fam) <= tol * np.abs(ref_fam))
            assert abs(tot - ref_tot) <= tol * abs(ref_tot)
            # and against the oracle
            oc = om.ccds(texts[:ntx])
            want = om.logpdf(oc, orr, per_family=True, precision="long")
            assert np.all(np.abs(fam - want) <= max(tol, 1e-12) * np.abs(want))


def test_noalloc_global_memory_families(w):
    """A family whose stages run in global memory (4,917 clades) also takes
    its stage buffers from the fixed array."""
    tree = open(os.path.join(GOLDEN, "example-4", "tree.nw")).read().strip()
    ale = open(os.path.join(GOLDEN, "example-4", "cytp450.ale")).read()
    m = w.WhaleModel(w.ConstantDLWGD(λ=0.1, μ=0.2, η=1 / 1.5), w.readnw(tree), 0.1)
    c = w.ccds_from_texts([ale], m)
    b = w.device_batch(m, c)
    assert b.hbm_families() == 1
    for mode in ("fp64s", "fp32"):
        ref = w.logpdf_families(m, c, mode=mode)
        rc, msg, fam, tot = noalloc(w, b, mode)
        assert rc == 0, msg
        assert abs(fam[0] - ref[0]) <= 1e-12 * abs(ref[0])


def test_noalloc_refuses_batches_beyond_its_array(w):
    from whale_jl_amd import _lib, synth
    wl = synth.workload("config4", n_families=2000, nthreads=8)
    m = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    c = w.ccds_from_texts(wl.texts, m)
    b = w.device_batch(m, c)
    m._sync()
    assert 12 * b.stats()["ell_cells"] > (64 << 20)
    rc, msg, fam, tot = noalloc(w, b, "fp32")
    assert rc == 1 and "workspace" in msg  # WHALE_EINVAL
    assert np.all(np.isnan(fam)) and math.isnan(tot)
