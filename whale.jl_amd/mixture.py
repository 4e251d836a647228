"""MixtureModel and ModelArray likelihoods, mirroring src/core.jl:66-79, and
the sampling condition of a model (src/condition.jl:11-29).

    condition(wm)                       log P(sampling condition)
    MixtureModel(components, prior)     logpdf: ℓhood(Σ_i logsumexp_j (ℓ_ij + log p_j − c_j))
    ModelArray(models)                  logpdf: Σ_i logpdf(models[i], ccds[i])

When the components share one species-tree structure and a device-side
condition, the mixture is evaluated by whale_mixture_eval: every size-class
launch works for all components at once (a workgroup per component), the
log-sum-exp and responsibilities are formed on the device, and the gradient's
reverse sweeps of all components run as one launch per class with the
responsibilities as the families' weights.  Otherwise each component is
evaluated by the batched launches and the per-family values are combined on
the host.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import check, dblp, ptr
from .ccd import CCD, device_batch
from .core import _mode, gradient, logpdf, logpdf_families
from .model import NoCondition, NonExtinctCondition, NowhereExtinctCondition, RootCondition


def _geompgf(p, s):
    """pgf of the geometric number of lineages at the root (src/bdputil.jl:67)."""
    return p * s / (1.0 - (1.0 - p) * s)


def _bdp_pgf(lam, mu, t: float, s):
    """pgf of the linear birth-death process over time t (src/bdputil.jl:52-59);
    real or complex arguments (the branch test uses the real parts)."""
    if abs((lam - mu).real) <= 1e-6:
        return (1.0 - (lam * t - 1.0) * (s - 1.0)) / (1.0 - lam * t * (s - 1.0))
    rho = np.exp((mu - lam) * t)
    return (rho * (lam * s - mu) - mu * (s - 1.0)) / (rho * (lam * s - mu) - lam * (s - 1.0))


def _wgdpgf(q, s):
    return s * (1.0 - q + s * q)  # src/bdputil.jl:73


def _nowhere_c(nodes, lam, mu, q, eta):
    """P(no leaf family is empty) (src/condition.jl:31-36): one minus the
    signed sum of the tree's joint pgf at every all-binary argument but all
    ones (src/bdputil.jl:107-130), entries in the reference's order (first
    child's index fastest).  lam, mu per node id - 1, q per WGD id - 1;
    every operation is analytic, so complex arguments give complex-step
    derivatives."""
    def walk(d):
        if d.isroot():
            f = lambda x: _geompgf(eta, x)
        else:
            lam_e, mu_e, t = lam[d.id - 1], mu[d.id - 1], float(d.distance)
            f = lambda x: _bdp_pgf(lam_e, mu_e, t, x)
        if d.isleaf():
            return [f(0.0), f(1.0)], [1.0, -1.0]
        xs, sg = [1.0], [1.0]
        for c in d.children:
            v, s = walk(nodes[c - 1])
            xs = [a * b for b in v for a in xs]
            sg = [a * b for b in s for a in sg]
        if d.wgd:
            qq = q[d.wgdid - 1]
            return [f(_wgdpgf(qq, x)) for x in xs], sg
        return [f(x) for x in xs], sg

    root = next(d for d in nodes if d.isroot())
    p, sg = walk(root)
    return 1.0 - sum(a * b for a, b in zip(p[:-1], sg[:-1]))


def nowhere_extinct(wm) -> float:
    """log P(no leaf family is empty) (src/condition.jl:31-36)."""
    lam, mu, q, eta, _, _ = wm.node_params()
    c = float(np.real(_nowhere_c(wm.s.nodes, [float(x) for x in lam], [float(x) for x in mu],
                                 [float(x) for x in q], float(eta))))
    return math.log(c) if 0.0 < c < 1.0 else -math.inf


def nowhere_extinct_grad(wm):
    """∂ log P(no leaf family is empty) / ∂(λ_e, μ_e per node, q per WGD, η),
    natural scale: each parameter in turn gets an imaginary step ih, and
    ∂c = Im c(θ + ih) / h (complex-step differentiation, Squire and Trapp,
    SIAM Review 40, 1998), exact to rounding at h = 1e-30 since c is analytic
    in every parameter."""
    lam, mu, q, eta, _, _ = wm.node_params()
    nodes = wm.s.nodes
    base = [[complex(float(x)) for x in lam], [complex(float(x)) for x in mu], [complex(float(x)) for x in q],
            [complex(float(eta))]]
    c = float(np.real(_nowhere_c(nodes, base[0], base[1], base[2], base[3][0])))
    if not 0.0 < c < 1.0:
        raise ValueError("the NowhereExtinct condition is not a probability in (0, 1) at these parameters")
    h = 1e-30
    out = []
    for k, vec in enumerate(base):
        d = np.zeros(len(vec))
        for i in range(len(vec)):
            args = [list(v) for v in base]
            args[k][i] += 1j * h
            d[i] = np.imag(_nowhere_c(nodes, args[0], args[1], args[2], args[3][0])) / h / c
        out.append(d)
    dl, dm, dq, deta = out
    return dl, dm, dq, float(deta[0])


def condition(wm) -> float:
    """The model's sampling condition as a log-probability (src/condition.jl:11-36):
    Root and NonExtinct from the device slice tables (ε at the top of the root
    and its children), NowhereExtinct from the tree's pgf on the host."""
    cond = wm.s.condition
    if isinstance(cond, NoCondition):
        return 0.0
    if isinstance(cond, NowhereExtinctCondition):
        return nowhere_extinct(wm)
    tabs = wm.tables()
    eta = float(wm.node_params()[3])
    root = next(d for d in wm.s.nodes if d.isroot())
    eps_top = lambda i: float(tabs[i][0][-1])
    er = _geompgf(eta, eps_top(root.id))
    if isinstance(cond, NonExtinctCondition):
        return math.log(1.0 - er)
    if isinstance(cond, RootCondition):
        f, g = root.children
        p = 1.0 - _geompgf(eta, eps_top(f)) - _geompgf(eta, eps_top(g)) + er
        return math.log(p) if p > 0.0 else -math.inf
    raise ValueError(f"condition {type(cond).__name__} is not supported")


def _logsumexp_rows(a: np.ndarray) -> np.ndarray:
    m = np.max(a, axis=1)
    finite = np.isfinite(m)
    out = np.full(a.shape[0], -np.inf)
    mm = m[finite][:, None]
    out[finite] = m[finite] + np.log(np.sum(np.exp(a[finite] - mm), axis=1))
    return out


class MixtureModel:
    """A finite mixture of WhaleModels over families (Distributions.MixtureModel
    with WhaleModel components, src/core.jl:66-76): each family comes from
    component j with prior probability prior[j]."""

    def __init__(self, components: Sequence, prior: Sequence[float]):
        if len(components) != len(prior) or not components:
            raise ValueError("one prior weight per component")
        self.components = list(components)
        self.prior = np.asarray(prior, dtype=np.float64)

    def family_matrix(self, ccds, mode: str = "fp64") -> np.ndarray:
        """ℓ[i, j] = logpdf(component j, family i) + log prior[j] − condition(component j)."""
        cols = []
        for wm, pj in zip(self.components, self.prior):
            fam = logpdf_families(wm, ccds, mode=mode)
            cols.append(fam + math.log(pj) - condition(wm))
        return np.stack(cols, axis=1)

    def on_device(self, ccds=None) -> bool:
        """Whether whale_mixture_eval evaluates this mixture: the components
        share one structure and a device-side condition, at most 8 of them, and
        (given `ccds`) the batch has no family in global memory."""
        s = self.components[0].s
        ok = (len(self.components) <= 8 and all(wm.s is s for wm in self.components)
              and not getattr(s.condition, "host", False))
        if ok and ccds is not None and not isinstance(ccds, CCD):
            ok = device_batch(self.components[0], ccds).hbm_families() == 0
        return ok

    def _device_eval(self, ccds, mode: str, grad: bool):
        """whale_mixture_eval: (total, ℓ (n × K), r (n × K), per-component
        natural-scale gradients or None, d/dprior)."""
        comps = self.components
        K = len(comps)
        s = comps[0].s
        b = device_batch(comps[0], ccds)
        hs = s.mix_handles(K)
        for j, wm in enumerate(comps):
            hs[j][1] = wm._set_params(hs[j][0], hs[j][1])
        arr = (C.c_void_p * K)(*[h.value for h, _ in hs])
        prior = np.ascontiguousarray(self.prior, dtype=np.float64)
        n, nn, nw = b.n, s.n, s.n_wgd
        fam = np.zeros(n * K)
        resp = np.zeros(n * K)
        g = np.zeros(K * (2 * nn + nw + 1))
        dprior = np.zeros(K)
        tot = C.c_double()
        check(_lib.lib().whale_mixture_eval(b.h, arr, K, ptr(prior, dblp), _mode(mode), int(grad), C.byref(tot),
                                            ptr(fam, dblp), ptr(resp, dblp), ptr(g, dblp) if grad else None,
                                            ptr(dprior, dblp)), "whale_mixture_eval")
        return tot.value, fam.reshape(n, K), resp.reshape(n, K), (g.reshape(K, -1) if grad else None), dprior

    def logpdf(self, ccds, mode: str = "fp64") -> float:
        if self.on_device(ccds):
            tot = self._device_eval(ccds, mode, False)[0]
        else:
            tot = float(np.sum(_logsumexp_rows(self.family_matrix(ccds, mode))))
        return tot if math.isfinite(tot) else -math.inf

    def responsibilities(self, ccds, mode: str = "fp64"):
        """(logpdf, r): r[i, j] the posterior probability that family i comes
        from component j (rows of families no component can produce are 0)."""
        if self.on_device(ccds):
            tot, _, R, _, _ = self._device_eval(ccds, mode, False)
            return (tot if math.isfinite(tot) else -math.inf), R
        return self._host_resp(ccds, mode)

    def gradient(self, ccds, mode: str = "fp64", device: bool = True):
        """(value, [gradient of each component], d/dprior): the derivative of
        logpdf(MixtureModel, ccds) (src/core.jl:66-76) with respect to each
        component's rates vector (as `gradient` lays it out) and to the prior
        weights, i.e. Σ_i r_ij ∂(ℓ_ij − c_j) and Σ_i r_ij / prior[j].  On the
        device path all components run in the same launches; otherwise (or with
        device=False) one evaluation per component gives ℓ_ij and one reverse
        sweep per component, weighted by the responsibilities, the rest."""
        if device and self.on_device(ccds):
            tot, _, R, g, dprior = self._device_eval(ccds, mode, True)
            nn, nw = self.components[0].s.n, self.components[0].s.n_wgd
            grads = []
            for j, wm in enumerate(self.components):
                row = g[j]
                dq = row[2 * nn:2 * nn + nw].copy() if nw else np.zeros(1)
                grads.append(wm.chain_gradient(row[:nn].copy(), row[nn:2 * nn].copy(), dq, row[-1]))
            return (tot if math.isfinite(tot) else -math.inf), grads, dprior
        tot, R = self._host_resp(ccds, mode)
        grads = [gradient(wm, ccds, mode=mode, weights=R[:, j])[1] for j, wm in enumerate(self.components)]
        return tot, grads, R.sum(axis=0) / self.prior

    def _host_resp(self, ccds, mode):
        F = self.family_matrix(ccds, mode)
        lse = _logsumexp_rows(F)
        R = np.zeros_like(F)
        ok = np.isfinite(lse)
        R[ok] = np.exp(F[ok] - lse[ok, None])
        tot = float(np.sum(lse))
        return (tot if math.isfinite(tot) else -math.inf), R


class ModelArray:
    """One WhaleModel per family (src/model.jl:257-261, src/core.jl:78-79):
    logpdf is the sum of the families' unconditioned log-likelihoods."""

    def __init__(self, models: Sequence):
        self.models = list(models)

    def __len__(self):
        return len(self.models)

    def logpdf(self, ccds: Sequence[CCD], mode: str = "fp64") -> float:
        if len(ccds) != len(self.models):
            raise ValueError("one model per family")
        return float(sum(logpdf(m, x, mode=mode) for m, x in zip(self.models, ccds)))
