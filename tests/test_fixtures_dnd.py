"""The oracle reproduces the reference's own guide trees.

test_data/correct/*.dnd of the reference were written by clustalw-2.0.11 from
its pairwise DistMatrix through NJ (tree/NJTree.cpp) and the PHYLIP writer
(tree/ClusterTreeOutput.cpp) with 5 decimals.  FASTA -> oracle pairwise
distances -> NJ restatement must give the same bytes; this pins the pairwise
tie rules (forward end cell, reverse start cell, Myers-Miller midpoints) with
the reference's data, for protein (Gonnet) and DNA (IUB) scoring.
"""
import numpy as np
import pytest

import oracle as O
from fixtures import load


@pytest.mark.parametrize("setname", ["100_97", "100_498", "1000_97", "100_1002", "100_1523"])
def test_oracle_reproduces_reference_guide_tree(setname):
    names, codes, offs, lens, dna, dnd = load(setname)
    assert dna == setname.endswith(("1002", "1523"))
    n = len(names)
    pairs = O.range_pairs(n, 0, n, 0, n)
    _, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring(dna=dna), nthreads=8)
    assert O.nj_dnd(O.distance_matrix(n, pairs, dist), names) == dnd


def test_nj_two_sequences():
    d = np.zeros((3, 3))
    d[1, 2] = d[2, 1] = 0.25
    assert O.nj_dnd(d, ["a", "b"]) == "(a:0.125,b:0.125);\n"
