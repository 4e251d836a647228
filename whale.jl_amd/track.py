"""Reconciled trees by stochastic backtracking (src/track.jl:188-414) from the
ℓ that a `logpdf_` (logpdf!) or gradient call leaves on the device, sampled
on the device by whale_backtrack, and the per-gene-pair reconciliation
summary of src/rectree.jl:53-103 (getpairs).

    logpdf_(wm, ccds)                       # stores ℓ, as logpdf! does
    trees = backtrack(wm, ccds, n=100, seed=1)   # trees[i][s]: family i, sample s
    pp = pair_posteriors(trees[0], wm)      # {(γ1, γ2): {"e_label": freq}}
"""
from __future__ import annotations

import ctypes as C
from collections import Counter, defaultdict
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check, i32p, ptr
from .ccd import device_batch

# node labels of src/track.jl:455
LABELS = ["loss", "wgd", "wgdloss", "duplication", "sploss", "speciation"]


class RecTree:
    """One reconciled tree: its nodes in creation order, each (clade 0-based or
    -1 for a loss, species node id - 1, slice row t, parent or -1)."""

    __slots__ = ("nodes", "_children")

    def __init__(self, nodes: np.ndarray):
        self.nodes = np.asarray(nodes, np.int32).reshape(-1, 4)
        self._children = None

    def __len__(self):
        return len(self.nodes)

    def children(self) -> List[List[int]]:
        if self._children is None:
            ch: List[List[int]] = [[] for _ in range(len(self.nodes))]
            for i in range(1, len(self.nodes)):
                ch[int(self.nodes[i, 3])].append(i)
            self._children = ch
        return self._children

    def leaves(self, i: int = 0) -> List[int]:
        """Node indices of the leaves below node i (losses included)."""
        ch = self.children()
        out, stack = [], [i]
        while stack:
            x = stack.pop()
            if ch[x]:
                stack.extend(reversed(ch[x]))
            else:
                out.append(x)
        return out

    def genes(self) -> List[int]:
        """Clades of the non-loss leaves (each gene once)."""
        return sorted(int(self.nodes[i, 0]) for i in self.leaves() if self.nodes[i, 0] >= 0)

    def label(self, i: int, wm) -> str:
        """getlabel (src/track.jl:457-478)."""
        g, e = int(self.nodes[i, 0]), int(self.nodes[i, 1])
        ch = self.children()[i]
        if g < 0:
            return "loss"
        wgd = wm.s.nodes[e].wgd
        dup = len(ch) == 2 and all(int(self.nodes[c, 1]) == e for c in ch)
        loss = any(self.nodes[c, 0] < 0 for c in ch)
        if len(ch) == 1 and wgd:
            return "wgdloss"
        if not dup and not loss and wgd:
            return "wgd"
        if dup:
            return "duplication"
        if loss:
            return "sploss"
        return "speciation"

    def key(self) -> bytes:
        """Identity of the reconciliation (the nodes in creation order)."""
        return self.nodes.tobytes()


def backtrack(wm, ccds, n: int = 1, seed: int = 0, families: Optional[Sequence[int]] = None,
              cap: int = 4096) -> List[List[RecTree]]:
    """n reconciled trees per family from the ℓ the last evaluation stored
    (run logpdf_(wm, ccds) first, as the reference runs logpdf! before
    backtrack).  families: a contiguous range of family indices (default all)."""
    b = device_batch(wm, ccds)
    if families is None:
        f0, nf = 0, b.n
    else:
        fams = list(families)
        f0, nf = fams[0], len(fams)
        if fams != list(range(f0, f0 + nf)):
            raise ValueError("families must be a contiguous range")
    counts = np.zeros(nf * n, np.int32)
    nodes = np.zeros(nf * n * cap * 4, np.int32)
    check(_lib.lib().whale_backtrack(b.h, f0, nf, n, C.c_uint64(seed), cap, ptr(counts, i32p), ptr(nodes, i32p)),
          "whale_backtrack")
    if np.any(counts < 0):
        raise RuntimeError(f"a reconciled tree has more than cap={cap} nodes; call again with a larger cap")
    nodes = nodes.reshape(nf * n, cap, 4)
    out = []
    for i in range(nf):
        out.append([RecTree(nodes[i * n + s, :counts[i * n + s]].copy()) for s in range(n)])
    return out


def pair_labels(wm) -> List[str]:
    """Column labels of getpairs (src/rectree.jl:53-63): per node in wm.order,
    duplication, then wgd and wgdloss for WGD nodes or speciation."""
    out = []
    for d in wm.order:
        out.append(f"{d.id}_duplication")
        out.extend([f"{d.id}_wgd", f"{d.id}_wgdloss"] if d.wgd else [f"{d.id}_speciation"])
    return out


def pair_posteriors(trees: Sequence[RecTree], wm) -> Dict[Tuple[int, int], Dict[str, float]]:
    """For each pair of genes, the fraction of the trees that reconcile their
    last common ancestor as each event on each branch (_getpairs!,
    src/rectree.jl:93-103; node ids as the reference's)."""
    d: Dict[Tuple[int, int], Counter] = defaultdict(Counter)
    f = 1.0 / len(trees)
    for t in trees:
        ch = t.children()
        for i in range(len(t)):
            if len(ch[i]) != 2:
                continue  # leaves and degree-1 nodes
            lab = f"{int(t.nodes[i, 1]) + 1}_{t.label(i, wm)}"
            a = [int(t.nodes[x, 0]) for x in t.leaves(ch[i][0]) if t.nodes[x, 0] >= 0]
            b = [int(t.nodes[x, 0]) for x in t.leaves(ch[i][1]) if t.nodes[x, 0] >= 0]
            for x in a:
                for y in b:
                    d[(min(x, y), max(x, y))][lab] += f
    return {k: dict(v) for k, v in d.items()}
