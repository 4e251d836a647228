"""ctypes front for the CPU oracle (oracle/whale_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference values the HIP path is compared
with (and, in bench.py, the CPU baseline that is timed).  The product package
never imports this module.

The interface follows the reference's Julia API closely enough for tests to
read like test/runtests.jl:

    m = OracleModel(extree, inserts=[("wgd_1", "ATHA", "ATHA"),
                                     ("wgd_2", "ATHA", "ATRI")], dt=0.05, maxn=5)
    fam = m.ccds(texts)
    m.logpdf(fam, DLWGD(lam, mu, q, eta))             # conditioned batch value
    m.logpdf(fam, ..., per_family=True)               # unconditioned per family
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

COND = {"root": 0, "nonextinct": 1, "none": 2}


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.orc_last_error.restype = C.c_char_p
        L.orc_model_new.restype = C.c_void_p
        L.orc_model_new.argtypes = [C.c_char_p, C.c_char_p, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_model_free.argtypes = [C.c_void_p]
        L.orc_model_counts.argtypes = [C.c_void_p, ip, ip, ip, ip]
        L.orc_model_node.argtypes = [C.c_void_p, C.c_int, ip, ip, ip, dp, dp, ip, C.c_char_p, C.c_int]
        L.orc_model_order.argtypes = [C.c_void_p, ip]
        L.orc_ccds_new.restype = C.c_void_p
        L.orc_ccds_new.argtypes = [C.c_void_p, C.POINTER(C.c_char_p), C.c_int]
        L.orc_ccds_free.argtypes = [C.c_void_p]
        L.orc_ccd_nclades.argtypes = [C.c_void_p, C.c_int]
        L.orc_ccd_nsplits.argtypes = [C.c_void_p, C.c_int]
        L.orc_ccd_export.argtypes = [C.c_void_p, C.c_int, ip, ip, ip, ip, ip, dp]
        L.orc_logpdf.restype = C.c_double
        L.orc_logpdf.argtypes = [C.c_void_p, C.c_void_p, C.c_int, dp, C.c_int, C.c_int, dp, C.c_int,
                                 C.c_int, C.c_int, dp, dp]
        L.orc_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_int, dp, C.c_int, C.c_int, dp, C.c_int,
                               C.c_int, dp, dp]
        L.orc_tables.argtypes = [C.c_void_p, C.c_int, dp, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp, dp]
        L.orc_getslice.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.orc_backtrack.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, C.c_int,
                                    C.c_uint64, C.c_int, C.c_int, ip, ip]
        L.orc_ell.restype = C.c_long
        L.orc_ell.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, C.c_int,
                              dp, C.c_long]
        _LIB = L
    return _LIB


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


@dataclass
class DLWGD:
    """Branch rates on log scale, indexed by reference node id (rmodels.jl:201-218)."""
    lam: Sequence[float]
    mu: Sequence[float]
    q: Sequence[float] = ()
    eta: float = 0.66
    p: Sequence[float] = ()

    def pack(self) -> Tuple[int, np.ndarray, int, int, np.ndarray]:
        nl, nq = len(self.lam), len(self.q)
        th = np.concatenate([np.asarray(self.lam, float), np.asarray(self.mu, float),
                             np.asarray(self.q, float), [float(self.eta)]])
        return 0, th, nl, nq, np.asarray(self.p, float)


@dataclass
class ConstantDLWGD:
    """Constant natural-scale rates (rmodels.jl:177-192)."""
    lam: float
    mu: float
    q: Sequence[float] = ()
    eta: float = 0.66
    p: Sequence[float] = ()

    def pack(self):
        nq = len(self.q)
        th = np.concatenate([[float(self.lam), float(self.mu)], np.asarray(self.q, float), [float(self.eta)]])
        return 1, th, 0, nq, np.asarray(self.p, float)


class OracleCCDs:
    def __init__(self, model: "OracleModel", texts: Sequence[str]):
        L = lib()
        arr = (C.c_char_p * len(texts))(*[t.encode() for t in texts])
        h = L.orc_ccds_new(model.h, arr, len(texts))
        if not h:
            raise ValueError(L.orc_last_error().decode())
        self.h = h
        self.n = len(texts)
        self.model = model

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ccds_free(self.h)
            self.h = None

    def __len__(self):
        return self.n

    def nclades(self, i: int) -> int:
        return lib().orc_ccd_nclades(self.h, i)

    def export(self, i: int):
        """Flattened clades of family i (1-based clade ids, as the reference)."""
        L = lib()
        nc, ns = L.orc_ccd_nclades(self.h, i), L.orc_ccd_nsplits(self.h, i)
        size = np.zeros(nc, np.int32); sp = np.zeros(nc, np.int32); off = np.zeros(nc + 1, np.int32)
        g1 = np.zeros(ns, np.int32); g2 = np.zeros(ns, np.int32); pr = np.zeros(ns)
        L.orc_ccd_export(self.h, i, _ip(size), _ip(sp), _ip(off), _ip(g1), _ip(g2), _dp(pr))
        return dict(size=size, leaf_species=sp, split_off=off, g1=g1, g2=g2, p=pr)


class OracleModel:
    def __init__(self, newick: str, inserts: Sequence[Tuple[str, str, str]] = (), dt: float = 0.05,
                 minn: int = 5, maxn: int = 10000, condition: str = "root"):
        L = lib()
        ins = ";".join(f"{n}={a},{b}" for n, a, b in inserts)
        h = L.orc_model_new(newick.encode(), ins.encode(), dt, minn, maxn, COND[condition])
        if not h:
            raise ValueError(L.orc_last_error().decode())
        self.h = h
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        L.orc_model_counts(h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        self.n_nodes, self.n_nonwgd, self.n_wgd, self.n_leaves = a.value, b.value, c.value, d.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_model_free(self.h)
            self.h = None

    def nodes(self):
        """Per reference id: dict(parent, wgdid, k, delta, leafP, leaf, wgd, root, name)."""
        L = lib()
        out = []
        for i in range(1, self.n_nodes + 1):
            p, w, k, fl = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            dl, lp = C.c_double(), C.c_double()
            nm = C.create_string_buffer(256)
            L.orc_model_node(self.h, i, C.byref(p), C.byref(w), C.byref(k), C.byref(dl), C.byref(lp),
                             C.byref(fl), nm, 256)
            out.append(dict(id=i, parent=p.value, wgdid=w.value, k=k.value, delta=dl.value, leafP=lp.value,
                            leaf=bool(fl.value & 1), wgd=bool(fl.value & 2), root=bool(fl.value & 4),
                            name=nm.value.decode()))
        return out

    def order(self) -> List[int]:
        o = np.zeros(self.n_nodes, np.int32)
        lib().orc_model_order(self.h, _ip(o))
        return [int(x) for x in o]

    def ccds(self, texts: Sequence[str]) -> OracleCCDs:
        return OracleCCDs(self, texts)

    def logpdf(self, ccds: OracleCCDs, rates, precision: str = "double", nthreads: int = 1,
               per_family: bool = False):
        kind, th, nl, nq, p = rates.pack()
        fam = np.zeros(len(ccds))
        cond = C.c_double()
        tot = lib().orc_logpdf(self.h, ccds.h, kind, _dp(th), nl, nq, _dp(p), len(p),
                               1 if precision == "long" else 0, nthreads, _dp(fam), C.byref(cond))
        if np.isnan(tot) and lib().orc_last_error():
            raise ValueError(lib().orc_last_error().decode())
        if per_family:
            return fam
        return tot

    def condition_value(self, ccds: OracleCCDs, rates) -> float:
        kind, th, nl, nq, p = rates.pack()
        cond = C.c_double()
        lib().orc_logpdf(self.h, ccds.h, kind, _dp(th), nl, nq, _dp(p), len(p), 0, 1, None, C.byref(cond))
        return cond.value

    def gradient(self, ccds: OracleCCDs, rates, nthreads: int = 1):
        """ForwardDiff-style (chunk 12) gradient of the conditioned batch value
        with respect to the packed parameter vector [λ; μ; q; η]."""
        kind, th, nl, nq, p = rates.pack()
        g = np.zeros(len(th))
        v = C.c_double()
        rc = lib().orc_grad(self.h, ccds.h, kind, _dp(th), nl, nq, _dp(p), len(p), nthreads, C.byref(v), _dp(g))
        if rc != 0:
            raise ValueError(lib().orc_last_error().decode())
        return v.value, g

    def tables(self, rates, node_id: int, nrows: int):
        kind, th, nl, nq, p = rates.pack()
        e, f, s = np.zeros(nrows), np.zeros(nrows), np.zeros(nrows)
        n = lib().orc_tables(self.h, kind, _dp(th), nl, nq, _dp(p), len(p), node_id, _dp(e), _dp(f), _dp(s))
        return e[:n], f[:n], s[:n]

    def backtrack(self, ccds: OracleCCDs, fam: int, rates, seed: int, n_samples: int, cap: int = 4096):
        """Reconciled trees sampled by stochastic backtracking (src/track.jl:188-414)
        from the oracle's ℓ: per sample an (n, 4) array of nodes in creation
        order, [clade (0-based, -1 loss), species node (0-based), slice row, parent]."""
        kind, th, nl, nq, p = rates.pack()
        counts = np.zeros(n_samples, np.int32)
        nodes = np.zeros(n_samples * cap * 4, np.int32)
        rc = lib().orc_backtrack(self.h, ccds.h, fam, kind, _dp(th), nl, nq, _dp(p), len(p), seed, n_samples, cap,
                                 _ip(counts), _ip(nodes))
        if rc != 0:
            raise ValueError(lib().orc_last_error().decode())
        nodes = nodes.reshape(n_samples, cap, 4)
        return [nodes[i, :counts[i]].copy() if counts[i] >= 0 else None for i in range(n_samples)]

    def ell(self, ccds: OracleCCDs, fam: int, rates) -> np.ndarray:
        kind, th, nl, nq, p = rates.pack()
        n = lib().orc_ell(self.h, ccds.h, fam, kind, _dp(th), nl, nq, _dp(p), len(p), None, 0)
        out = np.zeros(n)
        lib().orc_ell(self.h, ccds.h, fam, kind, _dp(th), nl, nq, _dp(p), len(p), _dp(out), n)
        return out


def getslice(lam: float, mu: float, t: float, eps: float) -> np.ndarray:
    out = np.zeros(3)
    lib().orc_getslice(lam, mu, t, eps, _dp(out))
    return out
