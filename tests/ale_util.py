"""Writes the ALE v0.4 text of a single unrooted gene tree observed `obs`
times (every directed edge contributes a clade and its split, as ALEobserve
counts them).  Used to build small hand-made families for tests."""
from collections import defaultdict


def ale_from_tree(adj, names, obs=1):
    """adj: node -> list of neighbours (leaves have degree 1); names: leaf node -> gene name."""
    leaves = sorted(names)
    lid = {x: i + 1 for i, x in enumerate(leaves)}
    memo = {}

    def side(u, v):  # leaves on v's side of edge (u, v)
        if (u, v) in memo:
            return memo[(u, v)]
        if v in names:
            r = frozenset([lid[v]])
        else:
            r = frozenset().union(*[side(v, w) for w in adj[v] if w != u])
        memo[(u, v)] = r
        return r

    clades, splits = {}, defaultdict(set)
    for u in adj:
        for v in adj[u]:
            c = side(u, v)
            clades[c] = 1
            if v not in names:
                a, b = [side(v, w) for w in adj[v] if w != u]
                splits[c].add((a, b))
    sid = {}
    for c in sorted(clades, key=lambda c: (len(c), sorted(c))):
        if len(c) == 1:
            sid[c] = next(iter(c))
    k = len(leaves)
    for c in sorted(clades, key=lambda c: (len(c), sorted(c))):
        if len(c) > 1:
            k += 1
            sid[c] = k
    out = ["#constructor_string", "(x);", "#observations", str(obs), "#Bip_counts"]
    out += [f"{sid[c]}\t{obs}" for c in sorted(clades, key=lambda c: sid[c]) if len(c) > 1]
    out += ["#Bip_bls"] + [f"{sid[c]}\t1" for c in sorted(clades, key=lambda c: sid[c])]
    out += ["#Dip_counts"]
    for c in sorted(splits, key=lambda c: sid[c]):
        for a, b in sorted(splits[c], key=lambda ab: (sid[ab[0]], sid[ab[1]])):
            out.append(f"{sid[c]}\t{sid[a]}\t{sid[b]}\t{obs}")
    out += ["#last_leafset_id", str(len(clades)), "#leaf-id"]
    out += [f"{names[x]}\t{lid[x]}" for x in leaves]
    out += ["#set-id"]
    out += [f"{sid[c]}\t:\t" + "\t".join(str(i) for i in sorted(c)) for c in sorted(clades, key=lambda c: sid[c])]
    out += ["#END", ""]
    return "\n".join(out)


def balanced_paralogs(n_a, species_a="A", other=("B_1",)):
    """Unrooted tree: a balanced tree of n_a genes of species A, joined to the
    genes in `other` (one edge)."""
    adj = defaultdict(list)
    names = {}
    nxt = [0]

    def new():
        nxt[0] += 1
        return nxt[0]

    def build(m):
        x = new()
        if m == 1:
            names[x] = f"{species_a}_{len([1 for v in names.values() if v.startswith(species_a)]) + 1}"
            return x
        a, b = build(m // 2), build(m - m // 2)
        adj[x] += [a, b]
        adj[a].append(x)
        adj[b].append(x)
        return x

    top = build(n_a)
    prev = top
    for g in other:
        y = new()
        names[y] = g
        j = new()
        adj[j] += [prev, y]
        adj[prev].append(j)
        adj[y].append(j)
        prev = j
    # suppress the degree-2 node if any (the join node has degree 2: merge into an edge)
    for x in list(adj):
        if x not in names and len(adj[x]) == 2:
            a, b = adj[x]
            adj[a][adj[a].index(x)] = b
            adj[b][adj[b].index(x)] = a
            del adj[x]
    return dict(adj), names
