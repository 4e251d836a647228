"""Seeded synthetic workloads (libwhale_synth.so): species tree, gene
families as ALE v0.4 text, and parameters.

Five workloads, seed 20261202 + index:
  config1  9-taxon tree (height 1), λ=0.2 μ=0.15 η=0.9 on every branch,
           Δt=0.05, 100 families of 8–60 leaves, CCDs from 200 trees
  config2  12 taxa + 2 WGDs (q 0.2, 0.5), log λ_e ~ N(log 0.2, 0.5²),
           log μ_e ~ N(log 0.15, 0.5²), Δt=0.02, 1,000 families of 10–150 leaves
           (log-normal, median 30, σ 0.6), CCDs from 500 trees, ≤ 1,000 clades
  config3  the config1 model and families (the gradient workload)
  config4  50 taxa + 4 WGDs (q ~ U(0.1, 0.6)), rates as config2, Δt=0.01,
           10,000 families of 4–200 leaves (median 20, σ 0.9), CCDs from 200
           trees, ≤ 2,000 clades and 6,000 splits
  config5  20 taxa + 5 WGDs (q ~ U(0.1, 0.6)), rates as config2, Δt=0.02,
           5,000 families of 10–120 leaves (median 25, σ 0.5), 200 trees
Builder choices not fixed by the workload descriptions: η = 0.9 everywhere,
sampling probabilities 0, the family-size window and rate shift of the gene
family simulator (csrc/synth.cpp), and 200 trees per CCD for config4/5.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from .model import DLWGD

SEED = 20261202

CONFIGS = {
    "config1": dict(index=1, ntaxa=9, nwgd=0, nfam=100, size=("uniform", 0, 0, 8, 60), ntrees=200, dt=0.05,
                    rates="constant", q=None, max_clades=100000, max_splits=1000000),
    "config2": dict(index=2, ntaxa=12, nwgd=2, nfam=1000, size=("lognormal", 30, 0.6, 10, 150), ntrees=500, dt=0.02,
                    rates="lognormal", q=[0.2, 0.5], max_clades=1000, max_splits=1000000),
    "config3": dict(index=1, ntaxa=9, nwgd=0, nfam=100, size=("uniform", 0, 0, 8, 60), ntrees=200, dt=0.05,
                    rates="constant", q=None, max_clades=100000, max_splits=1000000),
    "config4": dict(index=4, ntaxa=50, nwgd=4, nfam=10000, size=("lognormal", 20, 0.9, 4, 200), ntrees=200,
                    dt=0.01, rates="lognormal", q="uniform", max_clades=2000, max_splits=6000),
    "config5": dict(index=5, ntaxa=20, nwgd=5, nfam=5000, size=("lognormal", 25, 0.5, 10, 120), ntrees=200,
                    dt=0.02, rates="lognormal", q="uniform", max_clades=100000, max_splits=1000000),
}


@dataclass
class Workload:
    name: str
    newick: str
    dt: float
    texts: List[str]
    rates: DLWGD
    n_nonwgd: int
    n_wgd: int
    leaves: np.ndarray = field(default_factory=lambda: np.zeros(0))
    clades: np.ndarray = field(default_factory=lambda: np.zeros(0))
    splits: np.ndarray = field(default_factory=lambda: np.zeros(0))


def species_tree(seed: int, ntaxa: int, nwgd: int) -> str:
    buf = C.create_string_buffer(1 << 20)
    n = _lib.synth().ws_species_tree(seed, ntaxa, nwgd, buf, len(buf))
    if n < 0:
        raise RuntimeError("species tree buffer too small")
    return buf.value.decode()


def families(newick: str, seed: int, nfam: int, size, ntrees: int, max_clades: int, max_splits: int,
             lam: float = 0.2, mu: float = 0.15, q: float = 0.3, nthreads: int = 16):
    kind, med, sig, lo, hi = size
    S = _lib.synth()
    h = S.ws_families(newick.encode(), seed, nfam, 0 if kind == "uniform" else 1, med, sig, lo, hi, ntrees,
                      max_clades, max_splits, lam, mu, q, nthreads)
    try:
        err = S.ws_error(h).decode()
        if err:
            raise RuntimeError(err)
        texts, st = [], np.zeros((nfam, 3), np.int64)
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        for i in range(nfam):
            texts.append(S.ws_text(h, i).decode())
            S.ws_stats(h, i, C.byref(a), C.byref(b), C.byref(c))
            st[i] = (a.value, b.value, c.value)
    finally:
        S.ws_free(h)
    return texts, st


def workload(name: str, n_families: Optional[int] = None, nthreads: int = 16) -> Workload:
    cfg = CONFIGS[name]
    seed = SEED + cfg["index"]
    nw = species_tree(seed, cfg["ntaxa"], cfg["nwgd"])
    nfam = cfg["nfam"] if n_families is None else n_families
    texts, st = families(nw, seed, nfam, cfg["size"], cfg["ntrees"], cfg["max_clades"], cfg["max_splits"],
                         nthreads=nthreads)
    n_nonwgd = 2 * cfg["ntaxa"] - 1
    rng = np.random.default_rng(seed)
    if cfg["rates"] == "constant":
        lam = np.full(n_nonwgd, math.log(0.2))
        mu = np.full(n_nonwgd, math.log(0.15))
    else:
        lam = math.log(0.2) + 0.5 * rng.standard_normal(n_nonwgd)
        mu = math.log(0.15) + 0.5 * rng.standard_normal(n_nonwgd)
    if cfg["q"] is None:
        q = []
    elif cfg["q"] == "uniform":
        q = list(rng.uniform(0.1, 0.6, cfg["nwgd"]))
    else:
        q = list(cfg["q"])
    rates = DLWGD(λ=list(lam), μ=list(mu), q=q, η=0.9)
    return Workload(name, nw, cfg["dt"], texts, rates, n_nonwgd, cfg["nwgd"], st[:, 0], st[:, 1], st[:, 2])


def perturbed(rates: DLWGD, n: int, sigma: float = 0.1, seed: int = SEED + 5) -> List[DLWGD]:
    """n parameter points around `rates`: log-normal perturbations of λ, μ, η
    and q (the sampler-style workload)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        lam = np.asarray(rates.λ) + sigma * rng.standard_normal(len(rates.λ))
        mu = np.asarray(rates.μ) + sigma * rng.standard_normal(len(rates.μ))
        q = np.clip(np.asarray(rates.q) * np.exp(sigma * rng.standard_normal(len(rates.q))), 1e-3, 0.999)
        eta = float(np.clip(rates.η * math.exp(sigma * rng.standard_normal()), 1e-3, 0.999))
        out.append(DLWGD(λ=list(lam), μ=list(mu), q=list(q), η=eta))
    return out
