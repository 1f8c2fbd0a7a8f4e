"""The default ClustalW run (Clustal::align, CPU/clustalw-2.0.11/src/Clustal.cpp:104-230)
on the library: full-DP pairwise distances on the device, the NJ guide tree
(.dnd text), weights / similarities / steps read back from it, the progressive
profile alignment on the device, and the CLUSTAL (.aln) text.
"""
from __future__ import annotations

import time

import numpy as np

from . import lib


def distance_matrix(n: int, pairs: np.ndarray, dist: np.ndarray) -> np.ndarray:
    """(n+1)x(n+1) DistMatrix from per-pair float distances (SetAt twice,
    FullPairwiseAlign.cpp:182-183); row/column 0 unused."""
    D = np.zeros((n + 1, n + 1), dtype=np.float64)
    D[pairs[:, 0] + 1, pairs[:, 1] + 1] = dist.astype(np.float64)
    D[pairs[:, 1] + 1, pairs[:, 0] + 1] = dist.astype(np.float64)
    return D


def align(ctx: lib.Context, names, seqs, dna: bool, timings: dict | None = None,
          guide_dnd: str | None = None):
    """(dnd_text, aln_text, aligned rows, 1-based output order) of one run.
    guide_dnd: a guide tree to use instead of the NJ tree of the distances
    (Clustal::doAlignUseOldTree, Clustal.cpp:780-900); a tree without branch
    lengths leaves UserParameters::distanceTree false (tree/Tree.cpp:564), which
    fixes the profile matrix of every step (SubMatrix.cpp:565)."""
    n = len(seqs)
    t0 = time.perf_counter()
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(n, np.int64)
    if n > 1:
        offs[1:] = np.cumsum(lens[:-1])
    codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if n else np.zeros(0, np.uint8)
    ctx.load_sequences(codes, offs, lens)
    if guide_dnd is None:
        params = lib.pair_params(dna=dna, max_len=max(int(lens.max()) if n else 1, 1))
        ctx.pairwise_range(params, 0, n, 0, n)
        t1 = time.perf_counter()
        # the guide tree from the distances as they stand in device memory
        dnd = ctx.nj_tree_from_run(names)
    else:
        t1 = time.perf_counter()
        dnd = guide_dnd
    w, sim, steps, dtree = lib.guide_from_tree(dnd, names, with_distance_tree=True)
    t2 = time.perf_counter()
    rows, order, stats = ctx.msa(seqs, sim, w, steps, params=lib.msa_params(dna, distance_tree=dtree),
                                 dna=dna)
    t3 = time.perf_counter()
    aln = lib.clustal_text(names, rows, order, dna=dna)
    t4 = time.perf_counter()
    if timings is not None:
        timings.update(pairwise_s=t1 - t0, tree_s=t2 - t1, msa_s=t3 - t2, output_s=t4 - t3,
                       msa_stats=stats)
    return dnd, aln, rows, order
