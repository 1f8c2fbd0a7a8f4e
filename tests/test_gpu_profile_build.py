"""Profile construction on the device (cwhip_step_profiles): the two profiles
of a profileAlign step -- gap coefficients (GAPCOL/LENCOL), the residue columns
of profile 1 (with the substitution matrix) and of profile 2 (standard) -- bit
for bit against the oracle's builders, on groups cut from the reference
alignments and under every switch of calcGapCoeff."""
import numpy as np
import pytest

import oracle as O
import cwhip
from cwhip.lib import msa_params, step_profiles
from test_msa_oracle import aligned_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cwhip.Context(0)
    yield c
    c.close()


def _check(ctx, seqs, sim, w, group, dna, **kw):
    n = len(seqs)
    aligned = np.ones(n + 1, np.int32)
    ref = O.step_profiles([np.asarray(s, np.int32) for s in seqs], sim, w, group, aligned,
                          O.msa_params(dna, **kw))
    got = step_profiles(ctx, [np.asarray(s, np.uint8) for s in seqs], sim, w, group, aligned,
                        msa_params(dna, **kw))
    if ref is None:
        assert got is None
        return
    p1, l1, p2, l2 = ref
    q1, m1, q2, m2 = got
    assert (m1, m2) == (l1, l2)
    np.testing.assert_array_equal(q1, p1, err_msg="profile 1")
    np.testing.assert_array_equal(q2, p2, err_msg="profile 2")


def _random_groups(setname, count, seed, kmax=20, trim=True):
    """Random groups of a fixture's aligned rows; the second group's rows cut
    at both ends so the profile lengths differ and rows start or end early."""
    names, seqs, dna, dnd = aligned_rows(setname)
    n = len(seqs)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    rng = np.random.default_rng(seed)
    for _ in range(count):
        perm = rng.permutation(n)
        k1 = int(rng.integers(1, min(kmax, n - 1)))
        k2 = int(rng.integers(1, min(kmax, n - k1) + 1))
        group = np.zeros(n + 1, np.int32)
        group[perm[:k1] + 1] = 1
        group[perm[k1:k1 + k2] + 1] = 2
        s2 = [s.copy() for s in seqs]
        if trim:
            cut = int(rng.integers(0, 60))
            for i in perm[k1:k1 + k2]:
                s2[i] = s2[i][cut // 2:len(s2[i]) - cut // 2]
        yield s2, sim, w, group, dna


@pytest.mark.parametrize("setname,count,kmax", [("100_97", 12, 20), ("100_498", 12, 40),
                                                 ("1000_97", 6, 300), ("1000_498", 4, 600)])
def test_step_profiles_match_oracle(ctx, setname, count, kmax):
    print("step profiles", setname, flush=True)
    for s2, sim, w, group, dna in _random_groups(setname, count, 5, kmax):
        _check(ctx, s2, sim, w, group, dna)


@pytest.mark.parametrize("setname", ["100_1002", "100_1523"])
@pytest.mark.parametrize("egp", [0, 1])
def test_step_profiles_dna(ctx, setname, egp):
    print("step profiles dna", setname, egp, flush=True)
    for s2, sim, w, group, dna in _random_groups(setname, 5, 6):
        _check(ctx, s2, sim, w, group, dna, end_gap_penalties=egp)


@pytest.mark.parametrize("kw", [dict(no_hyd_penalties=1), dict(no_pref_penalties=1),
                                dict(use_end_gaps=1), dict(end_gap_penalties=1),
                                dict(gap_dist=0), dict(gap_dist=1), dict(gap_dist=9),
                                dict(neg_matrix=1), dict(no_hyd_penalties=1, no_pref_penalties=1)])
def test_step_profiles_switches(ctx, kw):
    print("step profiles", kw, flush=True)
    for s2, sim, w, group, dna in _random_groups("100_498", 6, 7):
        _check(ctx, s2, sim, w, group, dna, **kw)


def test_step_profiles_variable_penalties(ctx):
    """Two-row groups more than 60% identical with variable penalties on: the
    +-5 window weights replace the hydrophilic and residue-specific ones and the
    gap-distance windows are off."""
    print("step profiles vw", flush=True)
    names, seqs, dna, dnd = aligned_rows("100_498")
    n = len(seqs)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    pairs = []
    for i in range(n):
        for j in range(i + 1, n):
            a, b = seqs[i], seqs[j]
            res = a < 24
            ident = 100.0 * np.sum(res & (a == b)) / max(1, np.sum(res))
            if ident > 62:
                pairs.append((i, j))
    assert len(pairs) >= 4
    rng = np.random.default_rng(8)
    for t in range(6):
        i, j = pairs[int(rng.integers(0, len(pairs)))]
        k, m = pairs[int(rng.integers(0, len(pairs)))]
        if len({i, j, k, m}) < 4:
            continue
        group = np.zeros(n + 1, np.int32)
        group[[i + 1, j + 1]] = 1
        group[[k + 1, m + 1]] = 2
        _check(ctx, seqs, sim, w, group, dna, no_var_penalties=0)
        _check(ctx, seqs, sim, w, group, dna, no_var_penalties=0, end_gap_penalties=1)


def test_step_profiles_degenerate(ctx):
    """Zero weights (a standard profile of zeros), empty rows, all-gap columns,
    one-row groups and an empty group."""
    print("step profiles degenerate", flush=True)
    names, seqs, dna, dnd = aligned_rows("100_97")
    n = len(seqs)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    g = np.zeros(n + 1, np.int32)
    g[[1, 2, 3]] = 1
    g[[4, 5]] = 2
    _check(ctx, seqs, sim, np.zeros(n, np.int32), g, dna)          # sum of weights 0
    s2 = [s.copy() for s in seqs]
    s2[3] = s2[3][:0]                                               # an empty row
    _check(ctx, s2, sim, w, g, dna)
    s3 = [np.concatenate([s[:10], np.full(7, 30, np.int32), s[10:]]) for s in seqs]
    _check(ctx, s3, sim, w, g, dna)                                 # all-gap columns
    s4 = [np.concatenate([np.full(9, 30, np.int32), s, np.full(5, 31, np.int32)]) for s in seqs]
    _check(ctx, s4, sim, w, g, dna)                                 # end gaps of both codes
    _check(ctx, s4, sim, w, g, dna, use_end_gaps=1)
    g1 = np.zeros(n + 1, np.int32)
    g1[7], g1[9] = 1, 2                                             # one row each
    _check(ctx, seqs, sim, w, g1, dna)
    g2 = np.zeros(n + 1, np.int32)
    g2[[1, 2]] = 1                                                  # group 2 empty
    _check(ctx, seqs, sim, w, g2, dna)
    s5 = [s.copy() for s in seqs]
    s5[6] = s5[6][:0]
    s5[8] = s5[8][:0]
    g3 = np.zeros(n + 1, np.int32)
    g3[7], g3[9] = 1, 2
    _check(ctx, s5, sim, w, g3, dna)                                # empty row against a row
