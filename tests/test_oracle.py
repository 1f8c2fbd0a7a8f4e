"""Pins of the CPU restatement (oracle/cw_oracle.c) that do not reuse its DP.

* closed forms: matAvgScore of Gonnet 250, the gap-open table (values derived
  from the matrix text and libm, not from the DP), the float distance;
* exhaustive enumeration of every alignment path for tiny inputs: the forward
  pass's maxScore and its first row-major end cell, and the optimum that
  diff() returns for the sub-rectangle (free end-gap opening, as diff() is
  called with tb = te = 0);
* the oracle's own edit script re-scored run by run equals diff()'s value,
  and the identity count of the script equals the sum over the M==1 leaves.
The reference's own fixtures (.dnd files) pin the tie-breaking; see
tests/test_fixtures_dnd.py.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O

SC = O.Scoring()


def test_gonnet_matavg_and_scale():
    assert SC.mat_avg == 82
    assert SC.int_scale == 10
    m = SC.mat
    assert m[0][0] == 240 and m[21][21] == 1420 and m[2][2] == 1150
    # symmetric, gap rows zero, O and U rows zero
    assert (m == m.T).all()
    assert (m[30] == 0).all() and (m[31] == 0).all()
    assert (m[13] == 0).all() and (m[19] == 0).all()


@pytest.mark.parametrize("L,go", [(97, 2390), (400, 2622), (1002, 2773), (1515, 2840),
                                  (1523, 2841), (3000, 2953)])
def test_gap_open_closed_form(L, go):
    assert SC.gaps(L, L + 7) == (go, 1)
    assert SC.gaps(L + 7, L) == (go, 1)


def test_gap_open_table_margins():
    # int(2*82*(10 + ln L)) in double; float evaluation would flip L=1515
    vals = [int(2 * 82 * (10.0 + math.log(L)) * 1.0) for L in range(1, 5001)]
    assert len(set(vals)) == 724
    assert vals[1514] == 2840
    # 2*82*(10 + ln 1515) = 2840.99999774: double rounding must be kept
    assert 2841 - 2 * 82 * (10.0 + math.log(1515)) < 3e-6
    tab = SC.gap_table(5000)
    assert list(tab[1:]) == vals


def test_dna_gaps():
    d = O.Scoring(dna=True)
    assert d.gaps(1000, 900) == (3000, 666)
    assert d.mat[0][0] == 1000 and d.mat[0][4] == 500    # A/G transition weight 0.5


@pytest.mark.parametrize("count,l1,l2", [(0, 10, 20), (10, 10, 20), (37, 97, 96), (1, 3, 7),
                                         (1200, 1515, 1600), (0, 0, 5)])
def test_distance(count, l1, l2):
    mn = min(l1, l2)
    if mn == 0:
        want = np.float32(1.0)
    else:
        s = np.float32(100.0 * np.float32(count)) / np.float32(mn)
        want = (np.float32(100.0) - s) / np.float32(100.0)
    got = O.lib().cwo_distance(count, l1, l2)
    assert np.float32(got) == want


# ---------------------------------------------------------------- enumeration

def paths(M, N):
    """All monotone paths from (0,0) to (M,N): steps 'M' (1,1), 'D' (1,0), 'I' (0,1)."""
    if M == 0 and N == 0:
        yield ()
        return
    if M > 0 and N > 0:
        for p in paths(M - 1, N - 1):
            yield p + ("M",)
    if M > 0:
        for p in paths(M - 1, N):
            yield p + ("D",)
    if N > 0:
        for p in paths(M, N - 1):
            yield p + ("I",)


def score_path(p, a, b, i0, j0, go, ge, tb=None, te=None):
    """Affine score of path p starting at node (i0, j0) over codes a, b (0-based
    arrays).  Gap runs cost go + ge*k; with tb/te given, a run touching the
    start node costs tb + ge*k and one touching the end node te + ge*k."""
    s, i, j = 0, i0, j0
    runs = [(k, len(list(g))) for k, g in itertools.groupby(p)]
    pos = 0
    for idx, (kind, ln) in enumerate(runs):
        if kind == "M":
            for _ in range(ln):
                s += SC.mat[a[i]][b[j]]
                i += 1
                j += 1
        else:
            if tb is not None and idx == 0:
                s -= tb + ge * ln
            elif te is not None and idx == len(runs) - 1:
                s -= te + ge * ln
            else:
                s -= go + ge * ln
            if kind == "D":
                i += ln
            else:
                j += ln
        pos += ln
    return s


def local_H(a, b, go, ge):
    n, m = len(a), len(b)
    H = np.zeros((n + 1, m + 1), dtype=np.int64)
    for i in range(n + 1):
        for j in range(m + 1):
            best = 0
            for i0 in range(i + 1):
                for j0 in range(j + 1):
                    for p in paths(i - i0, j - j0):
                        if p:
                            best = max(best, score_path(p, a, b, i0, j0, go, ge))
            H[i, j] = best
    return H


ALPHAS = {"full": [0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 17, 18, 20, 21, 23],
          "two": [0, 6], "three": [0, 17, 18], "one": [10]}


def cases(nc, maxlen, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nc):
        alpha = ALPHAS[list(ALPHAS)[k % len(ALPHAS)]]
        n = int(rng.integers(1, maxlen + 1))
        m = int(rng.integers(1, maxlen + 1))
        a = np.array(rng.choice(alpha, n), dtype=np.uint8)
        b = np.array(rng.choice(alpha, m), dtype=np.uint8)
        out.append((a, b))
    return out


@pytest.mark.parametrize("a,b", cases(24, 4, 11))
@pytest.mark.parametrize("go,ge", [(2390, 1), (60, 7), (5, 1)])
def test_forward_pass_by_enumeration(a, b, go, ge):
    H = local_H(a, b, go, ge)
    best = int(H.max())
    sc = O.Scoring()
    r = O.PairResult()
    A = np.zeros(len(a) + 1, np.int32); A[1:] = a
    B = np.zeros(len(b) + 1, np.int32); B[1:] = b
    O.lib().cwo_align_pair(A, len(a), B, len(b), sc.mat, go, ge, 30, 31, None, r)
    assert r.max_score == best
    if best > 0:
        idx = np.argwhere(H == best)[0]          # first in row-major order
        assert (r.se1, r.se2) == (int(idx[0]), int(idx[1]))
    else:
        assert (r.se1, r.se2) == (0, 0)


def global_opt(a, b, go, ge):
    return max(score_path(p, a, b, 0, 0, go, ge, tb=0, te=0) for p in paths(len(a), len(b)))


def rescore_displ(displ, a, b, i0, j0, go, ge):
    """Score of an edit script run by run (adjacent same-kind entries form one run)."""
    steps = []
    for v in displ:
        steps.extend(["M"] if v == 0 else (["I"] * v if v > 0 else ["D"] * (-v)))
    return score_path(tuple(steps), a, b, i0, j0, go, ge, tb=0, te=0), steps


@pytest.mark.parametrize("a,b", cases(30, 5, 23) + cases(8, 7, 29))
@pytest.mark.parametrize("go,ge", [(2390, 1), (40, 9), (3, 2)])
def test_myers_miller_optimal_and_consistent(a, b, go, ge):
    A = np.zeros(len(a) + 1, np.int32); A[1:] = a
    B = np.zeros(len(b) + 1, np.int32); B[1:] = b
    r = O.PairResult()
    d = np.zeros(2 * (len(a) + len(b)) + 2, np.int32)
    O.lib().cwo_align_pair(A, len(a), B, len(b), SC.mat, go, ge, 30, 31, d.ctypes.data, r)
    displ = d[1:r.n_displ + 1]
    assert r.count == r.count_leaves
    if r.max_score == 0:
        assert r.count == 0 and r.n_displ == 0
        return
    sa = a[r.sb1 - 1:r.se1]
    sb = b[r.sb2 - 1:r.se2]
    assert r.mm_score == global_opt(sa, sb, go, ge)
    s, steps = rescore_displ(displ, a, b, r.sb1 - 1, r.sb2 - 1, go, ge)
    assert s == r.mm_score
    # the path spans the sub-rectangle exactly and its identities are the count
    assert steps.count("M") + steps.count("D") == len(sa)
    assert steps.count("M") + steps.count("I") == len(sb)
    i, j, cnt = r.sb1 - 1, r.sb2 - 1, 0
    for st in steps:
        if st == "M":
            cnt += int(a[i] == b[j])
            i += 1; j += 1
        elif st == "D":
            i += 1
        else:
            j += 1
    assert cnt == r.count


def test_identical_and_empty():
    rng = np.random.default_rng(5)
    s = np.array(rng.choice(ALPHAS["full"], 80), dtype=np.uint8)
    r = O.align_pair(s, s.copy(), SC)
    assert r.count == 80 and (r.sb1, r.sb2, r.se1, r.se2) == (1, 1, 80, 80)
    codes = np.concatenate([s, s]).astype(np.uint8)
    offs = np.array([0, 80, 160], np.int64)
    lens = np.array([80, 80, 0], np.int32)
    cnt, dist = O.pairwise_pairs(codes, offs, lens, np.array([[0, 1], [0, 2], [2, 1]]), SC)
    assert list(cnt) == [80, -1, -1]
    assert list(dist) == [0.0, 1.0, 1.0]
