"""CCDs and read_ale, mirroring src/ccd.jl.  Parsing is done by the native
library (whale_ccd_parse, whale_ccd_parse_many); a CCD holds the flattened clades in the
reference's order (sorted by size then ALE set id; leaves first, the
ubiquitous clade last), with 0-based clade ids."""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, i32p, i64p, dblp, ptr


class CCD:
    __slots__ = ("size", "leaf_species", "split_off", "g1", "g2", "p", "fname", "ell", "_one", "__weakref__")

    def __init__(self, size, leaf_species, split_off, g1, g2, p, fname: str = ""):
        self.size = np.asarray(size, np.int32)
        self.leaf_species = np.asarray(leaf_species, np.int32)
        self.split_off = np.asarray(split_off, np.int64)
        self.g1 = np.asarray(g1, np.int32)
        self.g2 = np.asarray(g2, np.int32)
        self.p = np.asarray(p, float)
        self.fname = fname
        self.ell = None  # per node id: (k+1) x |compat| matrix, set by logpdf_
        self._one = None

    def __len__(self):
        return len(self.size)

    @property
    def nsplits(self) -> int:
        return len(self.g1)

    def __repr__(self):
        return f"CCD(Γ={len(self)}, 𝓛={int(np.sum(self.size == 1))})"

    @staticmethod
    def from_text(text: str, spmap, fname: str = "") -> "CCD":
        names, ids, arr = _species_args(spmap)
        raw = text.encode()
        h = C.c_void_p()
        check(_lib.lib().whale_ccd_parse(raw, len(raw), arr, ptr(ids, i32p), len(names), C.byref(h)),
              f"read_ale {fname}")
        return _take(h, fname)


def _species_args(spmap):
    names = list(spmap.keys())
    ids = np.array([spmap[k] - 1 for k in names], np.int32)
    arr = (C.c_char_p * len(names))(*[k.encode() for k in names])
    return names, ids, arr


def _take(h, fname: str) -> CCD:
    """Export a parsed family (whale_ccd_export) and free its handle."""
    try:
        nc, ns = C.c_int32(), C.c_int32()
        check(_lib.lib().whale_ccd_shape(h, C.byref(nc), C.byref(ns)), "whale_ccd_shape")
        size = np.zeros(nc.value, np.int32)
        sp = np.zeros(nc.value, np.int32)
        off = np.zeros(nc.value + 1, np.int32)
        g1 = np.zeros(ns.value, np.int32)
        g2 = np.zeros(ns.value, np.int32)
        p = np.zeros(ns.value)
        check(_lib.lib().whale_ccd_export(h, ptr(size, i32p), ptr(sp, i32p), ptr(off, i32p), ptr(g1, i32p),
                                          ptr(g2, i32p), ptr(p, dblp)), "whale_ccd_export")
    finally:
        _lib.lib().whale_ccd_destroy(h)
    return CCD(size, sp, off, g1, g2, p, fname)


def parse_many(texts: Sequence, spmap, fnames: Sequence[str], n_threads: int = 0) -> List[CCD]:
    """ALE texts (str or bytes) to CCDs, parsed on host threads
    (whale_ccd_parse_many); the error of a text that does not parse names
    its file."""
    n = len(texts)
    if n == 0:
        return []
    raw = [t if isinstance(t, bytes) else t.encode() for t in texts]
    names, ids, arr = _species_args(spmap)
    tarr = (C.c_char_p * n)(*raw)
    lens = (C.c_size_t * n)(*[len(b) for b in raw])
    hs = (C.c_void_p * n)()
    bad = C.c_int32(-1)
    rc = _lib.lib().whale_ccd_parse_many(tarr, lens, n, arr, ptr(ids, i32p), len(names), n_threads, hs,
                                         C.byref(bad))
    if rc != _lib.WHALE_OK:
        check(rc, f"read_ale {fnames[bad.value]}" if 0 <= bad.value < n else "read_ale")
    out = []
    try:
        for i in range(n):
            h, hs[i] = hs[i], None
            out.append(_take(h, fnames[i]))
    finally:
        for i in range(len(out), n):  # handles left by a failed export
            if hs[i]:
                _lib.lib().whale_ccd_destroy(hs[i])
    return out


class CCDSet(list):
    """A list of CCDs (what read_ale returns) that keeps its device batches."""

    def __init__(self, it=()):
        super().__init__(it)
        self._batches = {}


def read_ale(path: str, wm) -> CCDSet:
    """read_ale(path, wm) (src/ccd.jl:126-137): one .ale file, a directory of
    them (sorted), or a text file listing them."""
    if not os.path.exists(path):
        raise ValueError(f"Not a file nor directory `{path}`")
    spmap = wm.s.spmap
    if os.path.isfile(path) and path.endswith(".ale"):
        files = [path]
    else:
        if os.path.isfile(path):
            files = [l.rstrip("\n") for l in open(path)]
        else:
            files = [os.path.join(path, f) for f in sorted(os.listdir(path))]
        files = [f for f in files if not f.startswith("#")]
    texts = []
    for f in files:
        with open(f) as fh:  # text mode: line endings normalised
            texts.append(fh.read())
    return CCDSet(parse_many(texts, spmap, [os.path.basename(f) for f in files]))


def ccds_from_texts(texts: Sequence[str], wm, names: Optional[Sequence[str]] = None) -> CCDSet:
    fnames = list(names) if names else [f"family{i}" for i in range(len(texts))]
    return CCDSet(parse_many(texts, wm.s.spmap, fnames))


class DeviceBatch:
    """A device-resident batch of families for one model structure
    (whale_batch_create)."""

    def __init__(self, wm, ccds: Sequence[CCD]):
        self.s = wm.s
        self.n = len(ccds)
        nc = np.array([len(c) for c in ccds], np.int64)
        fam_off = np.zeros(self.n + 1, np.int64)
        np.cumsum(nc, out=fam_off[1:])
        size = np.concatenate([c.size for c in ccds]).astype(np.int32)
        sp = np.concatenate([c.leaf_species for c in ccds]).astype(np.int32)
        g1 = np.concatenate([c.g1 for c in ccds]).astype(np.int32)
        g2 = np.concatenate([c.g2 for c in ccds]).astype(np.int32)
        p = np.concatenate([c.p for c in ccds]).astype(float)
        offs = []
        base = 0
        for c in ccds:
            offs.append(c.split_off[:-1].astype(np.int64) + base)
            base += c.nsplits
        split_off = np.concatenate(offs + [np.array([base], np.int64)])
        h = C.c_void_p()
        check(_lib.lib().whale_batch_create(wm.s.handle(), self.n, ptr(fam_off, i64p), ptr(size, i32p),
                                            ptr(sp, i32p), ptr(split_off, i64p), ptr(g1, i32p), ptr(g2, i32p),
                                            ptr(p, dblp), C.byref(h)), "whale_batch_create")
        self.h = h

    def stats(self) -> dict:
        o = np.zeros(12, np.int64)
        check(_lib.lib().whale_batch_stats(self.h, ptr(o, i64p)), "whale_batch_stats")
        keys = ["families", "clades", "splits", "cells", "split_terms", "sum_compat", "items", "records", "stages",
                "row_steps", "max_stage_items", "ell_cells"]
        return dict(zip(keys, (int(x) for x in o)))

    def fallbacks(self) -> int:
        """Families the last FP32 call recomputed in FP64 (their FP32 values left the range)."""
        n = C.c_int32()
        check(_lib.lib().whale_batch_fallbacks(self.h, C.byref(n)), "whale_batch_fallbacks")
        return n.value

    def hbm_families(self) -> int:
        """Families whose stages run in global memory (they do not fit one compute unit's LDS)."""
        n = C.c_int32()
        check(_lib.lib().whale_batch_hbm_families(self.h, C.byref(n)), "whale_batch_hbm_families")
        return n.value

    def set_lds_limit(self, nbytes: int):
        """Re-plans with at most `nbytes` of LDS per family's stages (0: the device's
        limit); families above it run with their stage buffers in global memory."""
        check(_lib.lib().whale_batch_set_lds_limit(self.h, int(nbytes)), "whale_batch_set_lds_limit")

    def autotune(self, mode: str = "fp32", gradient: bool = True, reps: int = 5, cache: Optional[str] = None) -> dict:
        """Searches this batch's launch shape (whale_batch_autotune) and applies it;
        `cache` is a text file of earlier results, keyed by device and workload."""
        o = np.zeros(8, np.int64)
        ms = np.zeros(2)
        check(_lib.lib().whale_batch_autotune(self.h, _lib.MODES[mode], int(gradient), int(reps),
                                              cache.encode() if cache else None, ptr(o, i64p), ptr(ms, dblp)),
              "whale_batch_autotune")
        return dict(class_bounds=[int(x) for x in o[:5]], layout=int(o[5]), waves_per_cu=int(o[6]),
                    candidates_timed=int(o[7]), ms_before=float(ms[0]), ms_after=float(ms[1]))

    def set_ell_budget(self, nbytes: int):
        """Bounds the memory of stored ℓ rows (0: none); above it evaluations run
        each family's forward and reverse sweep in one pass (whale_batch_set_ell_budget)."""
        check(_lib.lib().whale_batch_set_ell_budget(self.h, int(nbytes)), "whale_batch_set_ell_budget")

    def set_deterministic(self, on: bool = True):
        """Bit-identical gradients across calls (whale_batch_set_deterministic)."""
        check(_lib.lib().whale_batch_set_deterministic(self.h, int(bool(on))), "whale_batch_set_deterministic")

    def ell_bytes(self) -> dict:
        o = np.zeros(3, np.int64)
        check(_lib.lib().whale_batch_ell_bytes(self.h, ptr(o, i64p)), "whale_batch_ell_bytes")
        return dict(stored_form=int(o[0]), budgeted_slots=int(o[1]), held=int(o[2]))

    def tuning(self) -> list:
        o = np.zeros(7, np.int64)
        check(_lib.lib().whale_batch_get_tuning(self.h, ptr(o, i64p)), "whale_batch_get_tuning")
        return [int(x) for x in o]

    def set_seeds(self, seeds=None):
        """Per-family weights of the reverse sweep for later gradient calls
        (whale_batch_set_seeds); None restores 1."""
        if seeds is None:
            check(_lib.lib().whale_batch_set_seeds(self.h, None), "whale_batch_set_seeds")
            return
        a = np.ascontiguousarray(seeds, dtype=np.float64)
        if a.shape != (self.n,):
            raise ValueError(f"one seed per family ({self.n})")
        check(_lib.lib().whale_batch_set_seeds(self.h, ptr(a, dblp)), "whale_batch_set_seeds")

    def pooled(self, mode: str = "fp32") -> bool:
        """Whether batched evaluations of `mode` run the pooled launch (layout 6)."""
        n = C.c_int32()
        check(_lib.lib().whale_batch_pooled(self.h, _lib.MODES[mode], C.byref(n)), "whale_batch_pooled")
        return bool(n.value)

    def set_tuning(self, t: Sequence[int]):
        a = np.asarray(t, np.int64)
        check(_lib.lib().whale_batch_set_tuning(self.h, ptr(a, i64p)), "whale_batch_set_tuning")

    def launches(self, mode: str = "fp32") -> list:
        """Launch shape of each batched kernel: forward and reverse sweep per size class."""
        ncls = _lib.SIZE_CLASSES
        o = np.zeros(16 * ncls, np.int64)
        check(_lib.lib().whale_batch_launches(self.h, _lib.MODES[mode], ptr(o, i64p)), "whale_batch_launches")
        keys = ["families", "per_workgroup", "tables_in_lds", "grid", "lds_bytes", "shared_lds_bytes", "reserved",
                "family_lds_bytes"]
        out = []
        for d, name in enumerate(("forward", "backward")):
            for k in range(ncls):
                r = dict(zip(keys, (int(x) for x in o[8 * (ncls * d + k):8 * (ncls * d + k + 1)])))
                r["kernel"], r["class"] = name, k
                out.append(r)
        return out

    def ell(self, fam: int) -> np.ndarray:
        n = C.c_int64()
        check(_lib.lib().whale_download_ell(self.h, fam, None, 0, C.byref(n)), "whale_download_ell")
        out = np.zeros(n.value)
        check(_lib.lib().whale_download_ell(self.h, fam, ptr(out, dblp), n.value, C.byref(n)), "whale_download_ell")
        return out

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib.lib().whale_batch_destroy(self.h)
            self.h = None


def device_batch(wm, ccds) -> DeviceBatch:
    """The cached device batch of `ccds` for `wm`'s structure."""
    if isinstance(ccds, CCDSet):
        b = ccds._batches.get(id(wm.s))
        if b is None or b.s is not wm.s:
            b = DeviceBatch(wm, ccds)
            ccds._batches[id(wm.s)] = b
        return b
    return DeviceBatch(wm, ccds)
