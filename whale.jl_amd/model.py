"""WhaleModel, rate parameterisations and sampling conditions, mirroring
src/model.jl, src/rmodels.jl and src/condition.jl of Whale.jl.

The structure (node order, ids, WGD ids, MUL groups, slice counts) is built
here exactly as the reference builds it; the slice tables are built on the
GPU by the native library (whale_model_set_params).
"""
from __future__ import annotations

import ctypes as C
import math
from collections import Counter
from dataclasses import dataclass, field, replace
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, i32p, dblp, ptr
from .newick import Node, getleaves, postwalk


# ---------------------------------------------------------------- conditions
class SamplingCondition:
    code = -1


class RootCondition(SamplingCondition):
    """Non-extinction in both clades below the root (src/condition.jl:21-29)."""
    code = _lib.COND_ROOT


class NonExtinctCondition(SamplingCondition):
    """Non-extinction (src/condition.jl:15-19)."""
    code = _lib.COND_NONEXTINCT


class NoCondition(SamplingCondition):
    code = _lib.COND_NONE


class NowhereExtinctCondition(SamplingCondition):
    """Non-extinction at every leaf (src/condition.jl:5-9,31-36): inclusion–
    exclusion over the 2^L evaluations of the tree's pgf at all-binary
    arguments (src/bdputil.jl:107-130), so for trees with few leaves.  The
    device evaluates the unconditioned values; the condition and its
    gradient (complex-step differentiation of the same pgf sum) are applied
    on the host."""
    code = _lib.COND_NONE
    host = True


# ---------------------------------------------------------------- rates
@dataclass
class DLWGD:
    """Branch-wise duplication and loss rates on log scale, indexed by node id,
    WGD retention rates q (by WGD id), sampling probabilities p (by leaf id,
    optional) and the root's geometric parameter η (src/rmodels.jl:201-220)."""
    λ: Sequence[float]
    μ: Sequence[float]
    q: Sequence[float] = ()
    p: Sequence[float] = ()
    η: float = 0.66

    def __call__(self, θ: Dict) -> "DLWGD":
        return replace(self, **θ)

    def pack(self) -> np.ndarray:
        return np.concatenate([np.asarray(self.λ, float), np.asarray(self.μ, float),
                               np.asarray(self.q, float), [float(self.η)]])


@dataclass
class ConstantDLWGD:
    """Constant natural-scale rates (src/rmodels.jl:177-192)."""
    λ: float
    μ: float
    q: Sequence[float] = ()
    p: Sequence[float] = ()
    η: float = 0.66

    def __call__(self, θ: Dict) -> "ConstantDLWGD":
        return replace(self, **θ)

    def pack(self) -> np.ndarray:
        return np.concatenate([[float(self.λ), float(self.μ)], np.asarray(self.q, float), [float(self.η)]])


# ---------------------------------------------------------------- context
_CTX = None


def context():
    """The process-wide device context (device 0)."""
    global _CTX
    if _CTX is None:
        h = C.c_void_p()
        check(_lib.lib().whale_ctx_create(0, C.byref(h)), "whale_ctx_create")
        _CTX = h
    return _CTX


@dataclass
class ModelNode:
    id: int = 0
    name: str = ""
    distance: float = math.nan
    parent: int = 0                   # parent id, 0 for the root
    children: List[int] = field(default_factory=list)
    wgd: bool = False
    wgdid: int = 0
    leafP: float = 0.0
    k: int = 0                        # slices; rows = k + 1
    delta: float = 0.0
    clade: frozenset = frozenset()

    def isleaf(self):
        return not self.children

    def isroot(self):
        return self.parent == 0


class _Structure:
    """Discretised species tree shared by all reparameterisations of a model,
    with its native handle."""

    def __init__(self, tree: Node, dt: float, minn: int, maxn: int, condition: SamplingCondition):
        self.dt, self.minn, self.maxn, self.condition = dt, minn, maxn, condition
        leaves = getleaves(tree)
        counts = Counter(x.name for x in leaves)
        mulgroups = {k: v for k, v in counts.items() if v > 1}
        # walk: preorder WGD numbering, postorder push (src/model.jl:102-121)
        wgd_counter = 0
        post: List[Node] = []
        wgdid: Dict[int, int] = {}
        stack = [(tree, False)]
        while stack:
            x, done = stack.pop()
            if not done:
                if x.name.startswith("wgd"):
                    wgd_counter += 1
                    wgdid[id(x)] = wgd_counter
                stack.append((x, True))
                for c in reversed(x.children):
                    stack.append((c, False))
            else:
                post.append(x)
        order_nodes = list(leaves) + [x for x in post if not x.isleaf()]
        nonwgd = sum(1 for x in post if not x.name.startswith("wgd"))
        ids: Dict[int, int] = {}
        i, j = nonwgd + 1, 1
        for x in order_nodes:
            if x.name.startswith("wgd"):
                ids[id(x)] = i
                i += 1
            else:
                ids[id(x)] = j
                j += 1
        n = len(order_nodes)
        self.n, self.n_nonwgd, self.n_wgd = n, nonwgd, wgd_counter
        nodes = [ModelNode() for _ in range(n)]
        mulid: Dict[str, int] = {}
        for x in order_nodes:
            d = nodes[ids[id(x)] - 1]
            d.id = ids[id(x)]
            d.name = x.name
            d.distance = x.distance
            d.parent = ids[id(x.parent)] if x.parent is not None else 0
            d.children = [ids[id(c)] for c in x.children]
            d.wgd = x.name.startswith("wgd")
            d.wgdid = wgdid.get(id(x), 0)
            d.leafP = 1.0 / mulgroups[x.name] if x.name in mulgroups else (1.0 if x.isleaf() else 0.0)
            t = x.distance
            d.k = 0 if math.isnan(t) else min(maxn, max(minn, math.ceil(t / dt)))
            d.delta = 0.0 if d.k == 0 else t / d.k
            if not d.wgd and x.name in mulgroups:
                mulid[x.name] = d.id
            if not d.isleaf() and not d.wgd and len(d.children) != 2:
                raise ValueError(f"speciation node {d.id} must have two children")
            if d.wgd and len(d.children) != 1:
                raise ValueError(f"WGD node {d.name} must have one child")
        self.nodes = nodes
        self.order = [ids[id(x)] for x in order_nodes]
        self.root = self.order[-1]
        # setclade!: leaf species ids (MUL leaves share the representative id)
        for idx in self.order:
            d = nodes[idx - 1]
            if d.isleaf():
                d.clade = frozenset([mulid.get(d.name, d.id)])
            else:
                s = set()
                for c in d.children:
                    s |= nodes[c - 1].clade
                d.clade = frozenset(s)
        # spmap: leaf name -> id (later leaves win), as read_ale builds it
        self.spmap = {}
        for x in leaves:
            self.spmap[x.name] = ids[id(x)]
        self._h = None

    def handle(self):
        if self._h is None:
            self._h = self._create()
            self._current = None
        return self._h

    def mix_handles(self, k: int):
        """k further device models of this structure, each with its own
        parameters (the components of a device mixture evaluation): a list of
        [handle, key of the parameters it holds]."""
        mix = self.__dict__.setdefault("_mix", [])
        while len(mix) < k:
            mix.append([self._create(), None])
        return mix[:k]

    def _create(self):
        n = self.n
        par = np.array([d.parent - 1 for d in self.nodes], np.int32)
        kind = np.zeros(n, np.int32)
        c0 = np.full(n, -1, np.int32)
        c1 = np.full(n, -1, np.int32)
        for d in self.nodes:
            if d.isroot():
                kind[d.id - 1] = _lib.NODE_ROOT
            elif d.wgd:
                kind[d.id - 1] = _lib.NODE_WGD
            elif d.isleaf():
                kind[d.id - 1] = _lib.NODE_LEAF
            else:
                kind[d.id - 1] = _lib.NODE_SPECIATION
            if d.children:
                c0[d.id - 1] = d.children[0] - 1
            if len(d.children) > 1:
                c1[d.id - 1] = d.children[1] - 1
        ks = np.array([d.k for d in self.nodes], np.int32)
        dl = np.array([d.delta for d in self.nodes], float)
        lp = np.array([d.leafP for d in self.nodes], float)
        wi = np.array([d.wgdid - 1 if d.wgd else -1 for d in self.nodes], np.int32)
        ls = np.array([(min(d.clade) - 1) if d.isleaf() else -1 for d in self.nodes], np.int32)
        order = np.array([i - 1 for i in self.order], np.int32)
        h = C.c_void_p()
        check(_lib.lib().whale_model_create(context(), n, ptr(par, i32p), ptr(kind, i32p), ptr(c0, i32p),
                                            ptr(c1, i32p), ptr(ks, i32p), ptr(dl, dblp), ptr(lp, dblp),
                                            ptr(wi, i32p), ptr(ls, i32p), n, ptr(order, i32p),
                                            self.condition.code, C.byref(h)), "whale_model_create")
        return h

    def __del__(self):
        if _lib._lib is None:
            return
        for h, _ in self.__dict__.get("_mix", []):
            _lib.lib().whale_model_destroy(h)
        self.__dict__["_mix"] = []
        if getattr(self, "_h", None) is not None:
            _lib.lib().whale_model_destroy(self._h)
            self._h = None


class WhaleModel:
    """WhaleModel(rates, tree, Δt; minn=5, maxn=10000, condition=RootCondition())
    (src/model.jl:96-145).  Calling a model with new parameters returns a model
    that shares the discretised tree (src/model.jl:147-160)."""

    def __init__(self, rates, tree: Optional[Node], dt: float = 0.05, minn: int = 5, maxn: int = 10000,
                 condition: SamplingCondition = None, _structure: Optional[_Structure] = None):
        self.rates = rates
        self.s = _structure or _Structure(tree, dt, minn, maxn, condition or RootCondition())

    def __call__(self, θ) -> "WhaleModel":
        rates = self.rates(θ) if isinstance(θ, dict) else θ
        return WhaleModel(rates, None, _structure=self.s)

    # -- reference accessors
    @property
    def order(self) -> List[ModelNode]:
        return [self.s.nodes[i - 1] for i in self.s.order]

    def __len__(self):
        return self.s.n

    def __getitem__(self, i: int) -> ModelNode:
        return self.s.nodes[i - 1]

    def root(self) -> ModelNode:
        return self.s.nodes[self.s.root - 1]

    @property
    def condition(self):
        return self.s.condition

    def setsamplingp(self, d: Dict[str, float]):
        """setsamplingp!(model, dict) (src/model.jl:248-255): every leaf's
        sampling probability p from `d` by leaf name (0 for a leaf not in
        `d`), stored in the rates by leaf id."""
        leaves = [x for x in self.s.nodes if x.isleaf()]
        p = list(self.rates.p) if len(self.rates.p) else [0.0] * len(leaves)
        for x in leaves:
            p[x.id - 1] = float(d.get(x.name, 0.0))
        self.rates.p = p

    # -- getθ (src/rmodels.jl:31-33, 55-64) on the natural scale, per node
    def node_params(self):
        s, r = self.s, self.rates
        n = s.n
        lam = np.full(n, np.nan)
        mu = np.full(n, np.nan)
        p = np.zeros(n)
        rate_index = np.full(n, -1, np.int64)  # index into r.λ the node's rate comes from
        pv = np.asarray(r.p, float)
        for d in s.nodes:
            e = d.id - 1
            if d.wgd:
                c = d
                while c.wgd:
                    c = s.nodes[c.children[0] - 1]
                ri = c.id
            else:
                ri = d.id
                if len(pv) > 0 and d.isleaf():
                    p[e] = pv[d.id - 1]
            if isinstance(r, ConstantDLWGD):
                lam[e], mu[e] = float(r.λ), float(r.μ)
                rate_index[e] = 0
            elif ri <= len(r.λ):
                lam[e], mu[e] = math.exp(r.λ[ri - 1]), math.exp(r.μ[ri - 1])
                rate_index[e] = ri - 1
        q = np.asarray(r.q, float)
        if len(q) < s.n_wgd:
            raise ValueError(f"the model has {s.n_wgd} WGD nodes but q has {len(q)} values")
        return lam, mu, q, float(r.η), p, rate_index

    def _sync(self):
        """Build this parameter point's slice tables on the device (once)."""
        h = self.s.handle()
        self.s._current = self._set_params(h, self.s._current)
        return h

    def _set_params(self, h, current):
        """This parameter point's tables on device model `h` unless `current`
        (the key of what h holds) is already this point; returns the key."""
        lam, mu, q, eta, p, _ = self.node_params()
        q = np.ascontiguousarray(q if len(q) else np.zeros(1))
        key = (lam.tobytes(), mu.tobytes(), q.tobytes(), eta, p.tobytes())
        if current != key:
            check(_lib.lib().whale_model_set_params(h, ptr(lam, dblp), ptr(mu, dblp), ptr(q, dblp), len(q), eta,
                                                    ptr(p, dblp)), "whale_model_set_params")
        return key

    def tables(self):
        """Device slice tables, per node id: (ε, φ, ψ) arrays of k+1 rows."""
        h = self._sync()
        n = self.s.n
        off = np.zeros(n + 1, np.int32)
        check(_lib.lib().whale_model_table_offsets(h, ptr(off, i32p)), "tables")
        T = int(off[-1])
        e, f, s = np.zeros(T), np.zeros(T), np.zeros(T)
        check(_lib.lib().whale_model_tables(h, ptr(e, dblp), ptr(f, dblp), ptr(s, dblp)), "tables")
        return {i + 1: (e[off[i]:off[i + 1]], f[off[i]:off[i + 1]], s[off[i]:off[i + 1]]) for i in range(n)}

    def chain_gradient(self, d_lam, d_mu, d_q, d_eta) -> np.ndarray:
        """Maps per-node natural-scale derivatives to the rates' own parameter
        vector (DLWGD: [λ; μ; q; η] on log scale; ConstantDLWGD: [λ, μ, q; η])."""
        lam, mu, q, eta, p, ri = self.node_params()
        r = self.rates
        # A node whose rates do not enter the likelihood (the root: one row,
        # no slices) has an exact zero derivative; its rate may be NaN
        # (test/runtests.jl:44-49), so zero terms are skipped rather than
        # multiplied, which keeps NaN·0 out of the sums as ForwardDiff does.
        if isinstance(r, ConstantDLWGD):
            gl = float(np.sum(d_lam[d_lam != 0]))
            gm = float(np.sum(d_mu[d_mu != 0]))
            return np.concatenate([[gl, gm], d_q[:len(r.q)], [d_eta]])
        nl = len(r.λ)
        gl = np.zeros(nl)
        gm = np.zeros(nl)
        for e in range(self.s.n):
            if ri[e] >= 0:
                if d_lam[e] != 0:
                    gl[ri[e]] += lam[e] * d_lam[e]
                if d_mu[e] != 0:
                    gm[ri[e]] += mu[e] * d_mu[e]
        return np.concatenate([gl, gm, d_q[:len(r.q)], [d_eta]])

    def __repr__(self):
        s = self.s
        lines = [f"WhaleModel\n{'—' * 10}\n⋅Parameterization:\n{self.rates}\n⋅Condition:\n{type(s.condition).__name__}\n",
                 f"⋅Model structure:\n{s.n} nodes ({sum(d.isleaf() for d in s.nodes)} leaves, {s.n_wgd} WGD nodes)",
                 "node_id,wgd_id,distance,Δt,n"]
        for d in self.order:
            lines.append(f"{d.id},{d.wgdid},{round(d.k * d.delta, 4)},{round(d.delta, 4)},{d.k}")
        return "\n".join(lines)
