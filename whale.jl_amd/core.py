"""logpdf / logpdf! / loglikelihood and the gradient, mirroring src/core.jl.

    logpdf(wm, ccd)          one family, unconditioned        (core.jl:39-43)
    logpdf(wm, ccds)         ℓhood(Σ ℓ_i − N·condition)       (core.jl:58-64)
    logpdf_(wm, ccd|ccds)    the same, keeping ℓ in ccd.ell   (logpdf!, core.jl:29-56)
    gradient(wm, ccds)       value and d value / dθ in the rates' own
                             parameterisation, from the reverse sweep
                             (the reference uses ForwardDiff, test/runtests.jl:36-38)

`mode` selects the arithmetic: "fp64" (the reference's Float64 recursion),
"fp32" (FP32 with per-clade power-of-two exponents) or "fp64s" (FP64 with
per-clade exponents; no underflow).  `launch="per_family"` runs one launch per
family instead of one batched launch.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence, Union

import numpy as np

from . import _lib
from ._lib import check, dblp, ptr
from .ccd import CCD, CCDSet, device_batch


def _mode(mode: str) -> int:
    try:
        return _lib.MODES[mode]
    except KeyError:
        raise ValueError(f"unknown mode {mode!r}; use one of {sorted(_lib.MODES)}")


def _launch(launch: str) -> int:
    return {"batched": _lib.BATCHED, "per_family": _lib.PER_FAMILY}[launch]


def _eval(wm, ccds, mode, launch, keep):
    h = wm._sync()
    b = device_batch(wm, ccds)
    fam = np.zeros(b.n)
    tot = C.c_double()
    m = _mode(mode) | (_lib.KEEP_ELL if keep else 0)
    check(_lib.lib().whale_logpdf(b.h, m, _launch(launch), ptr(fam, dblp), C.byref(tot)), "whale_logpdf")
    if keep:
        for i, c in enumerate(ccds):
            c.ell = _split_ell(wm, c, b.ell(i))
    total = tot.value
    if getattr(wm.s.condition, "host", False):
        # a condition applied on the host: the device returned Σ ℓ_i
        from .mixture import condition
        total = total - len(fam) * condition(wm)
        total = total if np.isfinite(total) else -np.inf
    return total, fam


def _split_ell(wm, ccd: CCD, flat: np.ndarray):
    """Per node id: the (k+1) x |compat| matrix of ℓ (column-major in flat)."""
    s = wm.s
    sp = _species_sets(ccd)
    out = {}
    k = 0
    for d in s.nodes:
        ncomp = sum(1 for x in sp if x <= d.clade) if not d.isroot() else len(ccd)
        R = d.k + 1
        n = R * ncomp
        out[d.id] = flat[k:k + n].reshape(ncomp, R).T.copy() if n else np.zeros((R, 0))
        k += n
    return out


def _species_sets(ccd: CCD):
    sets = []
    for c in range(len(ccd)):
        if ccd.size[c] == 1:
            sets.append(frozenset([int(ccd.leaf_species[c]) + 1]))
        else:
            a = ccd.split_off[c]
            sets.append(sets[ccd.g1[a]] | sets[ccd.g2[a]])
    return sets


def logpdf(wm, x: Union[CCD, Sequence[CCD]], mode: str = "fp64", launch: str = "batched") -> float:
    if isinstance(x, CCD):
        _, fam = _eval(wm, _single_set(x), mode, launch, False)
        return float(fam[0])
    tot, _ = _eval(wm, x, mode, launch, False)
    return tot


def logpdf_(wm, x: Union[CCD, Sequence[CCD]], mode: str = "fp64", launch: str = "batched") -> float:
    """logpdf! : as logpdf, and leaves ℓ in each CCD's .ell (for backtracking)."""
    if isinstance(x, CCD):
        _, fam = _eval(wm, _single_set(x), mode, launch, True)
        return float(fam[0])
    tot, _ = _eval(wm, x, mode, launch, True)
    return tot


loglikelihood = logpdf


def logpdf_families(wm, ccds, mode: str = "fp64", launch: str = "batched") -> np.ndarray:
    """Per-family unconditioned log-likelihoods."""
    return _eval(wm, ccds, mode, launch, False)[1]


def gradient(wm, ccds, mode: str = "fp64", launch: str = "batched", weights=None):
    """(value, gradient, per-family values): the conditioned batch
    log-likelihood and its derivative with respect to the rates' parameter
    vector (DLWGD: [λ; μ; q; η], log-scale λ/μ; ConstantDLWGD: [λ, μ, q; η]).
    With `weights` (one per family, ≥ 0) the gradient is that of
    Σ_i weights[i]·(ℓ_i − c) instead (the reverse sweep's family weights,
    whale_batch_set_seeds); the value and per-family values are unchanged."""
    if isinstance(ccds, CCD):
        ccds = _single_set(ccds)
    host = getattr(wm.s.condition, "host", False)
    if host:
        from .model import NowhereExtinctCondition
        if not isinstance(wm.s.condition, NowhereExtinctCondition):
            raise NotImplementedError(f"no gradient with {type(wm.s.condition).__name__}")
    h = wm._sync()
    b = device_batch(wm, ccds)
    n, nw = wm.s.n, wm.s.n_wgd
    fam = np.zeros(b.n)
    dl, dm = np.zeros(n), np.zeros(n)
    dq = np.zeros(max(1, nw))
    deta = C.c_double()
    tot = C.c_double()
    if weights is not None:
        b.set_seeds(weights)
    try:
        check(_lib.lib().whale_logpdf_grad(b.h, _mode(mode), _launch(launch), ptr(fam, dblp), C.byref(tot),
                                           ptr(dl, dblp), ptr(dm, dblp), ptr(dq, dblp), C.byref(deta)),
              "whale_logpdf_grad")
    finally:
        if weights is not None:
            b.set_seeds(None)
    total, dt = tot.value, deta.value
    if host:
        # a condition applied on the host: the device returned Σ ℓ_i and its
        # gradient; subtract N·c and N·∂c (natural scale, per node), N the
        # families' total weight for the gradient
        from .mixture import nowhere_extinct, nowhere_extinct_grad
        nf = len(fam)
        ng = nf if weights is None else float(np.sum(weights))
        cl, cm, cq, ce = nowhere_extinct_grad(wm)
        total -= nf * nowhere_extinct(wm)
        dl -= ng * cl
        dm -= ng * cm
        dq[:len(cq)] -= ng * cq
        dt -= ng * ce
    g = wm.chain_gradient(dl, dm, dq, dt)
    return total, g, fam


def _single_set(ccd: CCD) -> CCDSet:
    """The one-family set of `ccd`, kept on the CCD so its device batch is reused."""
    if ccd._one is None:
        ccd._one = CCDSet([ccd])
    return ccd._one
