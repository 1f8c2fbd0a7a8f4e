"""The neighbour-joining guide tree built on the device (cw_nj, behind
cwhip_nj_tree_from_run and cwhip_nj_tree_on_device) against the oracle's
NJTree::generateTree (tree/NJTree.cpp:161-751) and the host code, text for text:

* the reference's own data sets: the tree of the device distances, joined in
  device memory, equals the reference's .dnd fixture;
* matrices whose distances are heavily tied (values on a coarse grid, all
  equal, all 1.0, zeros from identical sequences), where the pair chosen at
  each join is decided by the reference's scan order alone;
* the smallest sizes the kernel takes (4 sequences) and the host-code sizes
  either side of it (2, 3), and up to 1000 sequences.
"""
import numpy as np
import pytest

import cwhip
import oracle as O
from cwhip import pipeline, seqgen
from fixtures import load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cwhip.Context(0)
    yield c
    c.close()


def names_of(n):
    return ["seq%d" % i for i in range(n)]


def full_matrix(d):
    n = d.shape[0]
    D = np.zeros((n + 1, n + 1))
    D[1:, 1:] = d
    return D


def symmetric(rng, n, grid=None, zero_diag=True):
    x = rng.random((n, n))
    d = (x + x.T) * 0.5
    if grid:
        d = np.round(d * grid) / grid
    d = d.astype(np.float32).astype(np.float64)       # DistMatrix values are floats
    if zero_diag:
        np.fill_diagonal(d, 0.0)
    return d


@pytest.mark.parametrize("setname", ["100_97", "100_498", "1000_97"])
def test_tree_of_device_distances_equals_reference_dnd(ctx, setname):
    print("nj from run", setname, flush=True)
    names, codes, offs, lens, dna, dnd = load(setname)
    n = len(names)
    ctx.load_sequences(codes, offs, lens)
    params = cwhip.pair_params(dna=dna, max_len=int(lens.max()))
    pairs, _, dist = ctx.pairwise_range(params, 0, n, 0, n)
    got = ctx.nj_tree_from_run(names)
    assert got == dnd
    assert got == cwhip.nj_tree(pipeline.distance_matrix(n, pairs, dist), names)


@pytest.mark.parametrize("seed,n,grid", [(1, 4, 4), (2, 5, None), (3, 17, 3), (4, 64, 8), (5, 130, 10),
                                         (6, 257, 6), (7, 500, None), (8, 500, 20)])
def test_matrix_on_device_equals_oracle(ctx, seed, n, grid):
    print("nj matrix", seed, n, grid, flush=True)
    rng = np.random.default_rng(seed)
    D = full_matrix(symmetric(rng, n, grid))
    names = names_of(n)
    got = ctx.nj_tree_on_device(D, names)
    assert got == O.nj_dnd(D, names)
    assert got == cwhip.nj_tree(D, names)


@pytest.mark.parametrize("case", ["all_equal", "all_one", "identical_pairs", "two_values"])
def test_degenerate_matrices(ctx, case):
    print("nj degenerate", case, flush=True)
    n = 40
    rng = np.random.default_rng(11)
    if case == "all_equal":
        d = np.full((n, n), 0.25)
    elif case == "all_one":
        d = np.ones((n, n))
    elif case == "identical_pairs":                   # zero distances: dmin <= 0 path
        d = symmetric(rng, n, 5)
        for i in range(0, n - 1, 2):
            d[i, i + 1] = d[i + 1, i] = 0.0
    else:
        d = np.where(symmetric(rng, n) < 0.5, 0.125, 0.75)
        d = np.triu(d, 1)
        d = d + d.T
    np.fill_diagonal(d, 0.0)
    D = full_matrix(d)
    names = names_of(n)
    got = ctx.nj_tree_on_device(D, names)
    assert got == O.nj_dnd(D, names)


@pytest.mark.parametrize("n", [2, 3, 4])
def test_smallest_sizes(ctx, n):
    rng = np.random.default_rng(20 + n)
    D = full_matrix(symmetric(rng, n, 8))
    names = names_of(n)
    assert ctx.nj_tree_on_device(D, names) == O.nj_dnd(D, names)


def test_from_run_needs_a_run_over_all_pairs(ctx):
    fam = seqgen.family(5, 12, 60, name_prefix="nj")
    codes, offs, lens = fam.packed()
    ctx.load_sequences(codes, offs, lens)
    with pytest.raises(cwhip.CwhipError):
        ctx.nj_tree_from_run(fam.names)               # loaded, never run
    params = cwhip.pair_params(max_len=int(lens.max()))
    ctx.pairwise_range(params, 0, 6, 0, 12)           # a part of the pairs
    with pytest.raises(cwhip.CwhipError):
        ctx.nj_tree_from_run(fam.names)
    pairs, _, dist = ctx.pairwise_range(params, 0, 12, 0, 12)
    assert ctx.nj_tree_from_run(fam.names) == O.nj_dnd(O.distance_matrix(12, pairs, dist), fam.names)


def test_thousand_sequences_from_run(ctx):
    """prog1002's 1000 x ~1002 family: the device tree of the device distances
    against the oracle's NJ of the same distances."""
    fam = seqgen.workload("prog1002")
    codes, offs, lens = fam.packed()
    n = len(lens)
    print("nj prog1002", n, flush=True)
    ctx.load_sequences(codes, offs, lens)
    pairs, _, dist = ctx.pairwise_range(cwhip.pair_params(max_len=int(lens.max())), 0, n, 0, n)
    got = ctx.nj_tree_from_run(fam.names)
    assert got == O.nj_dnd(O.distance_matrix(n, pairs, dist), fam.names)
