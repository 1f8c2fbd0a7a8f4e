"""The library's host stages either side of the profile DP, against the
reference's own files: the NJ guide tree text (.dnd), the weights /
similarities / steps read back from it, and the CLUSTAL (.aln) writer.
None of these needs a device."""
import numpy as np
import pytest

import cwhip
import oracle as O
from cwhip import fasta
from fixtures import load, parse_aln, text

SETS = ["100_97", "100_498", "1000_97", "100_1002", "100_1523"]


@pytest.mark.parametrize("setname", ["100_97", "100_498", "100_1002", "100_1523"])
def test_nj_tree_reproduces_reference_dnd(setname):
    names, codes, offs, lens, dna, dnd = load(setname)
    n = len(names)
    pairs = O.range_pairs(n, 0, n, 0, n)
    cnt, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring(dna=dna), nthreads=8)
    assert cwhip.nj_tree(O.distance_matrix(n, pairs, dist), names) == dnd


@pytest.mark.parametrize("n,levels", [(300, 0), (700, 12), (1000, 40)])
def test_nj_tree_large_matches_oracle(n, levels):
    """Sizes where the minimum search runs on several host threads: the tree text
    must equal the oracle's single-threaded restatement, also when distances are
    quantised to a few values so that the scan order decides many ties."""
    rng = np.random.default_rng(n)
    x = rng.random((n, 6))
    d = np.zeros((n + 1, n + 1))
    d[1:, 1:] = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    if levels:
        d = np.round(d * levels) / levels
    d = d.astype(np.float32).astype(np.float64)
    np.fill_diagonal(d, 0.0)
    names = ["s%d" % i for i in range(n)]
    assert cwhip.nj_tree(d, names) == O.nj_dnd(d, names)


def test_nj_tree_two_and_three_sequences():
    for n in (2, 3):
        names = ["s%d" % i for i in range(n)]
        d = np.zeros((n + 1, n + 1))
        rng = np.random.default_rng(n)
        for i in range(1, n + 1):
            for j in range(i + 1, n + 1):
                d[i, j] = d[j, i] = float(np.float32(rng.uniform(0.1, 0.9)))
        assert cwhip.nj_tree(d, names) == O.nj_dnd(d, names)


@pytest.mark.parametrize("setname", SETS)
def test_guide_from_tree_matches_oracle(setname):
    names, codes, offs, lens, dna, dnd = load(setname)
    w, sim, steps = cwhip.guide_from_tree(dnd, names)
    wo, simo, stepso = O.guide_from_dnd(dnd, names)
    np.testing.assert_array_equal(w, wo)
    np.testing.assert_array_equal(sim.view(np.uint64), simo.view(np.uint64))
    np.testing.assert_array_equal(steps, stepso)


def _aln_rows(setname):
    names, codes, offs, lens, dna, dnd = load(setname)
    order, rows = parse_aln(text(setname + ".aln"))
    idx = {nm: k for k, nm in enumerate(names)}
    codes_rows = [np.array([30 if ch == "-" else fasta.ALPHABET.index(ch) for ch in rows[nm]],
                           np.uint8) for nm in names]
    return names, codes_rows, np.array([idx[nm] + 1 for nm in order], np.int32), dna


@pytest.mark.parametrize("setname", SETS + ["1000_498"])
def test_clustal_text_reproduces_reference_aln(setname):
    """The rows and order of the reference .aln written back out give the
    reference file byte for byte (name column, line length, conservation)."""
    names, rows, order, dna = _aln_rows(setname)
    assert cwhip.clustal_text(names, rows, order, dna=dna) == text(setname + ".aln")


def test_guide_from_tree_distance_tree_flag():
    """readTree leaves distanceTree true for a tree with branch lengths and false
    for a rooted one without (Tree.cpp:145, :561-565); without lengths every
    weight is INT_SCALE/n (Tree.cpp:177-182) and the steps equal the oracle's."""
    names, codes, offs, lens, dna, dnd = load("100_97")
    assert cwhip.guide_from_tree(dnd, names, with_distance_tree=True)[3] == 1

    def build(lo, hi):
        if hi - lo == 1:
            return names[lo]
        mid = (lo + hi) // 2
        return "(%s,%s)" % (build(lo, mid), build(mid, hi))
    tree = build(0, len(names)) + ";\n"
    w, sim, steps, dtree = cwhip.guide_from_tree(tree, names, with_distance_tree=True)
    wo, simo, stepso = O.guide_from_dnd(tree, names)
    assert dtree == 0
    assert (w == 1000 // len(names)).all()
    np.testing.assert_array_equal(w, wo)
    np.testing.assert_array_equal(sim.view(np.uint64), simo.view(np.uint64))
    np.testing.assert_array_equal(steps, stepso)
