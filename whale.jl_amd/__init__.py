"""whale.jl_amd: MI355X (gfx950) amalgamated-likelihood reconciliation DP with
Whale.jl's API for its likelihood path.

    from whale_jl_amd import readnw, insertnode, getlca, WhaleModel, DLWGD, read_ale, logpdf, logpdf_
    t = readnw(newick); insertnode(getlca(t, "ATHA", "ATHA"), name="wgd_1")
    w = WhaleModel(DLWGD(λ=..., μ=..., q=[0.2, 0.1], η=0.9), t, 0.05)
    ccd = read_ale("example/ale", w)
    logpdf(w, ccd)                 # conditioned batch log-likelihood (FP64)
    logpdf(w, ccd, mode="fp32")    # FP32 with per-clade exponents
    gradient(w, ccd)               # value and d/dθ by the reverse sweep

The computation runs in libwhale_amd.so (HIP kernels behind the C ABI of
include/whale_amd.h); there is no CPU fallback.
"""
from .newick import Node, readnw, insertnode, getlca, getleaves, postwalk, prewalk, nwstr, deepcopy_tree
from .model import (WhaleModel, DLWGD, ConstantDLWGD, RootCondition, NonExtinctCondition, NoCondition,
                    NowhereExtinctCondition, SamplingCondition, context)
from .ccd import CCD, CCDSet, read_ale, ccds_from_texts, DeviceBatch, device_batch
from .core import logpdf, logpdf_, loglikelihood, logpdf_families, gradient
from .mixture import MixtureModel, ModelArray, condition
from .track import RecTree, backtrack, pair_posteriors, pair_labels

__all__ = ["Node", "readnw", "insertnode", "getlca", "getleaves", "postwalk", "prewalk", "nwstr", "deepcopy_tree",
           "WhaleModel", "DLWGD", "ConstantDLWGD", "RootCondition", "NonExtinctCondition", "NoCondition",
           "SamplingCondition", "CCD", "CCDSet", "read_ale", "ccds_from_texts", "DeviceBatch", "device_batch",
           "logpdf", "logpdf_", "loglikelihood", "logpdf_families", "gradient", "context",
           "MixtureModel", "ModelArray", "condition", "NowhereExtinctCondition", "RecTree", "backtrack",
           "pair_posteriors", "pair_labels"]

EXTREE = ("((MPOL:4.752,PPAT:4.752):0.292,(SMOE:4.457,(((OSAT:1.555,(ATHA:0.5548,CPAP:0.5548):1.0002):0"
          ".738,ATRI:2.293):1.225,(GBIL:3.178,PABI:3.178):0.34):0.939):0.587);")
"""The reference's example species tree (src/Whale.jl:34-37)."""
