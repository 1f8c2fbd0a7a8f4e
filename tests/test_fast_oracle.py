"""The oracle's k-tuple distances (FastPairwiseAlign, the -QUICKTREE path)
against what the algorithm fixes by itself.  The reference ships no k-tuple
output, so the restatement is otherwise unpinned; these closed forms catch a
wrong diagonal index, a dropped chain step, a wrong DNA mapping or range."""
import numpy as np
import pytest

import oracle as O


def _rand(rng, n, alphabet):
    return np.array(rng.choice(alphabet, size=n), np.int32)


def test_identical_sequences_score_their_length():
    # the main-diagonal chain adds ktup per k-tuple step: score = L, percent 100
    rng = np.random.default_rng(1)
    for L in (1, 2, 7, 60, 300):
        s = _rand(rng, L, np.arange(25))
        p, d = O.fast_pairwise([s, s.copy()])
        assert d[0] == 0.0, L
        p, d = O.fast_pairwise([s, s.copy()], percent=0)
        assert d[0] == (100.0 - L) / 100.0


def test_no_common_ktuple_is_distance_one():
    rng = np.random.default_rng(2)
    a = _rand(rng, 80, np.arange(0, 10))
    b = _rand(rng, 90, np.arange(10, 20))
    p, d = O.fast_pairwise([a, b])
    assert d[0] == 1.0
    # gaps and codes above maxAA break every k-tuple
    g = np.full(50, 30, np.int32)
    p, d = O.fast_pairwise([g, g.copy()])
    assert d[0] == 1.0


def test_dna_mapping():
    # A C G T (U) only; other codes break a k-tuple, so a DNA run of them is empty
    rng = np.random.default_rng(3)
    acgt = np.array([0, 2, 6, 18], np.int32)
    s = _rand(rng, 120, acgt)
    u = s.copy()
    u[u == 18] = 19                                  # U reads as T
    p, d = O.fast_pairwise([s, u], dna=True)
    assert d[0] == 0.0
    other = _rand(rng, 120, np.array([1, 3, 4, 5, 7, 8], np.int32))
    p, d = O.fast_pairwise([other, other.copy()], dna=True)
    assert d[0] == 1.0


def test_partial_alignment_keeps_the_first_fragment():
    # max_aln = 1: the second fragment hits 2 x max_aln and the pass stops with
    # the first fragment alone, whose score is ktup
    rng = np.random.default_rng(4)
    s = _rand(rng, 40, np.arange(25))
    for ktup in (1, 2):
        p, d = O.fast_pairwise([s, s.copy()], max_aln=1, ktup=ktup)
        assert d[0] == (100.0 - (ktup / 40.0) * 100.0) / 100.0


def test_range_order_and_checks():
    rng = np.random.default_rng(5)
    seqs = [_rand(rng, 30, np.arange(25)) for _ in range(6)]
    p, d = O.fast_pairwise(seqs)
    want = [(i, j) for i in range(6) for j in range(i + 1, 6)]
    assert [tuple(x) for x in p] == want
    # FastPairwiseAlign.cpp:64-80: i = iStart+1 .. iEnd, j = max(jStart+2, i+1) .. jEnd
    p, d = O.fast_pairwise(seqs, ranges=(1, 4, 2, 6))
    want = [(i - 1, j - 1) for i in range(2, 5) for j in range(max(4, i + 1), 7)]
    assert [tuple(x) for x in p] == want
    for bad in ((-1, 3, 0, 3), (3, 2, 0, 3), (0, 3, -1, 3), (0, 3, 3, 2)):
        with pytest.raises(ValueError):
            O.fast_pairwise(seqs, ranges=bad)


def test_distances_in_unit_range_on_a_family():
    from fixtures import load
    names, codes, offs, lens, dna, dnd = load("100_97")
    seqs = [codes[offs[i]:offs[i] + lens[i]].astype(np.int32) for i in range(40)]
    p, d = O.fast_pairwise(seqs)
    assert len(d) == 40 * 39 // 2
    assert d.min() >= 0.0 and d.max() <= 1.0
