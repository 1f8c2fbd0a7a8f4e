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


# --- chains of k-tuple blocks (FastPairwiseAlign.cpp:171-403) -----------------
# Blocks of distinct letters in fillers that the other sequence lacks: the only
# k-tuple matches are the blocks' own, each block on one diagonal.  A block of
# b letters is a chain of b matches that adds ktup, then 1 per step: score b.
# A block B2 after B1 in both sequences on another flagged diagonal chains
# from B1's best fragment at the window-gap price: b1 - wind_gap + ktup at its
# first match, then 1 per step, so b1 + b2 - wind_gap; on the same diagonal
# there is no price (b1 + b2).  Only the top diagonals and those within
# `window` of them are searched at all.

B1 = np.arange(0, 8, dtype=np.int32)          # 8 distinct letters
B2 = np.arange(8, 13, dtype=np.int32)         # 5 more
F1, F2 = 13, 14                                # filler of sequence 1 / 2 only


def _two_blocks(gap1, gap2, lead1=10, lead2=10, tail=10):
    s1 = np.concatenate([np.full(lead1, F1), B1, np.full(gap1, F1), B2, np.full(tail, F1)]).astype(np.int32)
    s2 = np.concatenate([np.full(lead2, F2), B1, np.full(gap2, F2), B2, np.full(tail, F2)]).astype(np.int32)
    return s1, s2


def _score(s1, s2, **kw):
    p, d = O.fast_pairwise([s1, s2], percent=0, **kw)
    return round((1.0 - d[0]) * 100.0, 6)      # percent off: d = (100 - score) / 100


def test_single_block_scores_its_length():
    s1 = np.concatenate([np.full(20, F1), B1, np.full(15, F1)]).astype(np.int32)
    s2 = np.concatenate([np.full(5, F2), B1, np.full(30, F2)]).astype(np.int32)
    assert _score(s1, s2) == 8
    for ktup in (1, 2):
        assert _score(s1, s2, ktup=ktup) == 8
    # percent on: score / min length x 100
    p, d = O.fast_pairwise([s1, s2])
    assert d[0] == (100.0 - (8.0 / 43.0) * 100.0) / 100.0


def test_two_blocks_chain_with_the_window_gap():
    # B2 three diagonals off B1: chained, charged wind_gap once
    s1, s2 = _two_blocks(4, 7)
    assert _score(s1, s2) == 8 + 5 - 3
    assert _score(s1, s2, wind_gap=1) == 8 + 5 - 1
    assert _score(s1, s2, wind_gap=6) == 8                      # chaining would cost more
    # the same diagonal: the chain goes on at no price across the filler
    s1, s2 = _two_blocks(4, 4)
    assert _score(s1, s2) == 8 + 5
    # B2 before B1 in sequence 2: no chain, the longer block alone
    s1 = np.concatenate([np.full(10, F1), B1, np.full(4, F1), B2, np.full(10, F1)]).astype(np.int32)
    s2 = np.concatenate([np.full(10, F2), B2, np.full(4, F2), B1, np.full(10, F2)]).astype(np.int32)
    assert _score(s1, s2) == 8


def test_only_top_diagonals_and_their_window_are_searched():
    # desQuickSort leaves the last of the tl1 diagonals in place (its top
    # partition ends at tl1 - 1, :584-640) and the top-SIGNIF loop starts at
    # that unsorted last one (:236-251): signif 2 flags the one best diagonal,
    # signif 1 only the corner diagonal l1 + l2 - 1, which no block is on
    s1, s2 = _two_blocks(4, 7)                     # B2 three diagonals off B1
    assert _score(s1, s2, signif=1) == 0
    assert _score(s1, s2, signif=2) == 10          # inside window 5 of B1's diagonal
    assert _score(s1, s2, signif=2, window=3) == 10
    assert _score(s1, s2, signif=2, window=2) == 8  # outside: B2 adds nothing
    s1, s2 = _two_blocks(4, 16)                    # twelve diagonals off
    assert _score(s1, s2, signif=2) == 8
    assert _score(s1, s2, signif=3) == 10          # B2's diagonal is a top one itself
    assert _score(s1, s2, signif=2, window=12) == 10
    assert _score(s1, s2, signif=2, window=11) == 8
