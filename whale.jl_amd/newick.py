"""The part of NewickTree.jl (0.3.1) the likelihood path uses: readnw, node
ids/names/distances, getleaves/postwalk/prewalk, getlca and insertnode!.

Distances absent from the Newick string are NaN, as in NewickTree; the
reference turns a NaN branch (normally the root's) into a branch without
slices (src/model.jl:18).
"""
from __future__ import annotations

import math
from typing import List, Optional


class Node:
    __slots__ = ("id", "name", "distance", "children", "parent")

    def __init__(self, id: int, name: str = "", distance: float = math.nan, parent: Optional["Node"] = None):
        self.id = id
        self.name = name
        self.distance = distance
        self.children: List[Node] = []
        self.parent = parent

    def isleaf(self) -> bool:
        return not self.children

    def isroot(self) -> bool:
        return self.parent is None

    def __getitem__(self, i: int) -> "Node":  # n[1] is the first child, as in Julia
        return self.children[i - 1]

    def __repr__(self):
        return f"Node({self.id}, {self.name!r}, {self.distance})"


def readnw(s: str) -> Node:
    s = s.strip()
    pos = 0
    counter = [0]

    def ws():
        nonlocal pos
        while pos < len(s) and s[pos].isspace():
            pos += 1

    def node(parent):
        nonlocal pos
        ws()
        counter[0] += 1
        n = Node(counter[0], parent=parent)
        if pos < len(s) and s[pos] == "(":
            pos += 1
            while True:
                n.children.append(node(n))
                ws()
                if pos < len(s) and s[pos] == ",":
                    pos += 1
                    continue
                if pos < len(s) and s[pos] == ")":
                    pos += 1
                    break
                raise ValueError("newick: expected ',' or ')' at position %d" % pos)
        ws()
        st = pos
        while pos < len(s) and s[pos] not in ",():;" and not s[pos].isspace():
            pos += 1
        n.name = s[st:pos]
        ws()
        if pos < len(s) and s[pos] == ":":
            pos += 1
            ws()
            st = pos
            while pos < len(s) and s[pos] not in ",();" and not s[pos].isspace():
                pos += 1
            n.distance = float(s[st:pos])
        return n

    root = node(None)
    ws()
    if pos >= len(s) or s[pos] != ";":
        raise ValueError("newick: missing ';'")
    return root


def postwalk(n: Node) -> List[Node]:
    out: List[Node] = []
    stack = [(n, False)]
    while stack:
        x, done = stack.pop()
        if done or x.isleaf():
            out.append(x)
            continue
        stack.append((x, True))
        for c in reversed(x.children):
            stack.append((c, False))
    return out


def prewalk(n: Node) -> List[Node]:
    out: List[Node] = []
    stack = [n]
    while stack:
        x = stack.pop()
        out.append(x)
        stack.extend(reversed(x.children))
    return out


def getleaves(n: Node) -> List[Node]:
    return [x for x in prewalk(n) if x.isleaf()]


def getroot(n: Node) -> Node:
    while n.parent is not None:
        n = n.parent
    return n


def getlca(t: Node, a: str, b: str) -> Node:
    leaves = {x.name: x for x in reversed(getleaves(t))}  # first leaf of a name
    x, y = leaves[a], leaves[b]
    anc = set()
    while x is not None:
        anc.add(id(x))
        x = x.parent
    while y is not None:
        if id(y) in anc:
            return y
        y = y.parent
    raise ValueError("no common ancestor")


def insertnode(n: Node, name: str = "", dist: Optional[float] = None) -> Node:
    """insertnode!(n; dist=distance(n)/2, name): a new node above n, taking
    n's place among its parent's children."""
    if n.parent is None:
        raise ValueError("cannot insert a node above the root")
    d = n.distance
    dist = d / 2 if dist is None else dist
    p = n.parent
    m = Node(max(x.id for x in prewalk(getroot(n))) + 1, name=name, distance=dist, parent=p)
    p.children[p.children.index(n)] = m
    m.children.append(n)
    n.parent = m
    n.distance = d - dist
    return m


def deepcopy_tree(n: Node) -> Node:
    m = Node(n.id, n.name, n.distance)
    for c in n.children:
        cc = deepcopy_tree(c)
        cc.parent = m
        m.children.append(cc)
    return m


def nwstr(n: Node) -> str:
    def go(x):
        s = ""
        if x.children:
            s = "(" + ",".join(go(c) for c in x.children) + ")"
        s += x.name
        if not math.isnan(x.distance):
            s += ":%r" % x.distance
        return s
    return go(n) + ";"
