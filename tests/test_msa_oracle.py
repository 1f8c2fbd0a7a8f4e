"""The oracle's progressive stage against the reference's own fixtures.

The chain FASTA -> reference .dnd -> Tree::readTree / calcSeqWeights /
calcSimilarities / createSets (oracle/cw_guide.c) -> MSA::multiSeqAlign with
MyersMillerProfileAlign::profileAlign (oracle/cw_msa.c) must reproduce every
row of the reference's .aln files, gaps included, in the reference's output
order.  One wrong tie (progDiff midpoint, base case, progAdd/progDel order),
one wrong float rounding in a profile builder, or one wrong gap coefficient
moves a gap and fails the row comparison.
"""
import os

import numpy as np
import pytest

import oracle as O
from fixtures import load, parse_aln, row_text, text

SETS = ["100_97", "100_498", "1000_97", "1000_498", "100_1002", "100_1523"]


def _seqs(codes, offs, lens):
    return [codes[offs[i]:offs[i] + lens[i]].astype(np.int32) for i in range(len(lens))]


@pytest.mark.parametrize("setname", SETS)
def test_oracle_msa_reproduces_reference_aln(setname):
    names, codes, offs, lens, dna, dnd = load(setname)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    rows, order, info = O.msa(_seqs(codes, offs, lens), sim, w, steps, O.msa_params(dna))
    ref_order, ref_rows = parse_aln(text(setname + ".aln"))
    assert [names[i - 1] for i in order] == ref_order
    for i, nm in enumerate(names):
        assert row_text(rows[i]) == ref_rows[nm], nm
    assert len(info) == len(names) - 1


@pytest.mark.parametrize("setname", ["100_97", "1000_97", "100_1002"])
def test_guide_invariants(setname):
    names, codes, offs, lens, dna, dnd = load(setname)
    n = len(names)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    # n-1 steps; each step splits a set of sequences into two non-empty groups
    assert steps.shape == (n - 1, n + 1)
    assert (steps[:, 0] == 0).all()
    for s in steps:
        assert (s == 1).any() and (s == 2).any()
    # the root step covers everything; every leaf appears in the root step
    assert (steps[-1][1:] != 0).all()
    # weights are normalised to INT_SCALE_FACTOR (1000) up to integer truncation
    assert (w >= 1).all() and 1000 - n <= w.sum() <= 1000 + n
    # similarities: 100*(1 - tree distance), symmetric, zero diagonal, in [0, 99]
    assert np.array_equal(sim, sim.T)
    assert (np.diag(sim) == 0).all()
    off = sim[1:, 1:][~np.eye(n, dtype=bool)]
    assert off.min() >= 0.0 and off.max() <= 99.0 + 1e-9


def test_guide_rejects_mismatched_names():
    names, codes, offs, lens, dna, dnd = load("100_97")
    with pytest.raises(ValueError):
        O.guide_from_dnd(dnd, names[:-1] + ["not_in_tree"])


def _table(name):
    import re
    src = open(__file__.replace("tests/test_msa_oracle.py",
                                "cuda-clustalw_amd/csrc/cw_matrices.inc")).read()
    body = src[src.index("cw_%s[25 * 25] = {" % name):]
    body = body[body.index("{") + 1:body.index("};")]
    vals = [int(v) for v in re.findall(r"-?\d+", re.sub(r"/\*.*?\*/", "", body))]
    return np.array(vals, np.int64).reshape(25, 25)


@pytest.mark.parametrize("pcid,min_len,name,scale", [
    (90.0, 500, "gonnet80", 0.1875), (30.0, 500, "gonnet120", 0.375),
    (30.0, 50, "gonnet250", 0.375), (10.0, 500, "gonnet160", 0.375),
    (10.0, 50, "gonnet350", 0.375), (35.0, 500, "gonnet120", 0.375),
    (25.0, 500, "gonnet160", 0.375)])
def test_profile_matrix_series_closed_form(pcid, min_len, name, scale):
    """getProfileAlignMatrix: the gonnet series by pcid/minLen, scale x10,
    made non-negative by subtracting the minimum over the 23 table letters
    (not O and U), gap rows zero; matAvgScore from the unshifted table."""
    r, m, sc, avg = O.profile_matrix(False, pcid, min_len)
    t = _table(name) * 10
    letters = [c for c in range(25) if c not in (13, 19)]
    lo = min(t[i, j] for i in range(25) for j in range(1, i + 1))
    lo = min(lo, t[0, 0])
    exp = np.zeros((32, 32), np.int64)
    exp[:25, :25] = t
    for i in letters:
        for j in letters:
            exp[i, j] -= lo
    assert np.array_equal(m, exp)
    assert abs(sc - scale) < 1e-7
    off = sum(t[i, j] for i in range(25) for j in range(i))
    assert avg == -int(off / ((22 * 22 - 22) / 2.0))
    assert r == 24


def test_profile_matrix_dna_transitions():
    r, m, sc, avg = O.profile_matrix(True, 50.0, 500)
    t = _table("iub_dna") * 100
    A, C, G, T, U = 0, 2, 6, 18, 19
    assert m[A, A] == t[A, A] + 900          # IUB minimum is -900 after x100
    assert m[A, G] == m[G, A] == int(0.5 * m[A, A])
    assert m[C, T] == m[T, C] == m[C, U] == m[U, C] == int(0.5 * m[A, A])
    assert m[30, 30] == 0 and m[A, 30] == 0
    assert sc == 1.0


def test_prog_diff_score_equals_rescored_edit_script():
    """progDiff's returned value is the score of the alignment it emits under
    the same position-specific gap model (checked on a leaf-pair step: two
    single sequences, where the model reduces to the textbook affine one
    with per-position open/extend)."""
    names, codes, offs, lens, dna, dnd = load("100_498")
    seqs = _seqs(codes, offs, lens)
    n = len(seqs)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    p = O.msa_params(False)
    for a, b in [(0, 1), (5, 40), (17, 99)]:
        group = np.zeros(n + 1, np.int32)
        group[a + 1], group[b + 1] = 1, 2
        aligned = np.ones(n + 1, np.int32)
        p1, l1, p2, l2 = O.step_profiles(seqs, sim, w, group, aligned, p)
        score, displ, st = O.prog_diff(p1, l1, p2, l2)
        # columns: M + N - (aligned pairs); the script consumes both profiles
        used1 = sum(-k for k in displ if k < 0) + sum(1 for k in displ if k == 0)
        used2 = sum(k for k in displ if k > 0) + sum(1 for k in displ if k == 0)
        assert used1 == l1 and used2 == l2
        assert st.fwd_cells + st.rev_cells > 0
        assert _rescore(p1, l1, p2, l2, displ) == score


def _rescore(p1, l1, p2, l2, displ):
    """Score of an edit script: prfScore on aligned columns; a run of k
    profile-2 columns against gaps inside profile 1 at i costs
    open1(i, j0) + sum ext1 (gapPenalty1), a run of profile-1 columns
    open2 + sum ext2 (gapPenalty2); nothing at the profile ends."""
    def open1(i, j):
        return 0 if i in (0, l1) else p2[j, 32] + p1[i, 32]

    def open2(i, j):
        return 0 if j in (0, l2) else p1[i, 32] + p2[j, 32]

    def prf(i, j):
        s = int(np.dot(p1[i, :25].astype(np.int64), p2[j, :25]))
        s += int(p1[i, 30]) * int(p2[j, 30]) + int(p1[i, 31]) * int(p2[j, 31])
        return int(s / 10)

    i = j = 0
    total = 0
    for k in displ:
        if k == 0:
            i += 1
            j += 1
            total += prf(i, j)
        elif k > 0:                        # profile-2 columns j+1..j+k against gaps
            if i not in (0, l1):
                total -= open1(i, j + 1) + sum(int(p2[j + 1 + x, 33]) for x in range(k)
                                                if j + 1 + x < l2)
            j += k
        else:                              # profile-1 columns against gaps
            k = -k
            if j not in (0, l2):
                total -= open2(i + 1, j) + sum(int(p1[i + 1 + x, 33]) for x in range(k)
                                                if i + 1 + x < l1)
            i += k
    return total


def aligned_rows(setname):
    """The reference's aligned rows of a fixture as code arrays (gap -> 30)."""
    from cwhip import fasta
    names, codes, offs, lens, dna, dnd = load(setname)
    _, rows = parse_aln(text(setname + ".aln"))
    out = [np.array([30 if ch == "-" else fasta.ALPHABET.index(ch) for ch in rows[nm]], np.int32)
           for nm in names]
    return names, out, dna, dnd


def random_profile_pairs(setname, count, seed, egp=0):
    """Profile pairs built from random groups of a fixture's aligned rows
    (realistic internal gaps), the second group trimmed so the lengths differ."""
    names, seqs, dna, dnd = aligned_rows(setname)
    n = len(seqs)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    rng = np.random.default_rng(seed)
    p = O.msa_params(dna, end_gap_penalties=egp)
    out = []
    for _ in range(count):
        perm = rng.permutation(n)
        k1 = int(rng.integers(1, 20))
        k2 = int(rng.integers(1, 20))
        group = np.zeros(n + 1, np.int32)
        group[perm[:k1] + 1] = 1
        group[perm[k1:k1 + k2] + 1] = 2
        s2 = [s.copy() for s in seqs]
        cut = int(rng.integers(0, 60))
        for i in perm[k1:k1 + k2]:
            s2[i] = s2[i][cut // 2:len(s2[i]) - cut // 2]
        out.append(O.step_profiles(s2, sim, w, group, np.ones(n + 1, np.int32), p))
    return out


@pytest.mark.parametrize("setname", ["100_498", "100_1002"])
def test_prog_diff_score_equals_rescored_group_profiles(setname):
    for p1, l1, p2, l2 in random_profile_pairs(setname, 8, 3):
        score, displ, st = O.prog_diff(p1, l1, p2, l2)
        assert _rescore(p1, l1, p2, l2, displ) == score


def divergent_set(seed):
    """Two seeded families, one close and one divergent, with their NJ guide
    tree: (seqs int32, sim, weights, steps).  Its steps span every matrix of
    the gonnet series, unlike the fixtures (all above 90% identity)."""
    from cwhip import seqgen
    a = seqgen.family(seed * 13 + 5, 12, 220, branch_mean=0.8, name_prefix="da")
    b = seqgen.family(seed * 13 + 6, 8, 180, branch_mean=2.0, name_prefix="db")
    seqs, names = a.seqs + b.seqs, a.names + b.names
    n = len(seqs)
    codes = np.concatenate(seqs)
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(n, np.int64)
    offs[1:] = np.cumsum(lens[:-1])
    pairs = O.range_pairs(n, 0, n, 0, n)
    _, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring())
    dnd = O.nj_dnd(O.distance_matrix(n, pairs, dist), names)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    return [s.astype(np.int32) for s in seqs], sim, w, steps


def test_oracle_distance_tree_false_fixes_the_matrix():
    """distanceTree false: every step takes gon250 (SubMatrix.cpp:565);
    with it true the series follows pcid/minLen and the rows differ."""
    seqs, sim, w, steps = divergent_set(0)
    rows0, _, info0 = O.msa(seqs, sim, w, steps, O.msa_params(False, distance_tree=0))
    rows1, _, info1 = O.msa(seqs, sim, w, steps, O.msa_params(False))
    g250 = O.profile_matrix(False, 90.0, 500, distance_tree=0)[3]
    ran = [i for i in info0 if i.nseqs1 > 0]            # steps whose groups were not empty
    assert ran and {i.mat_avg for i in ran} == {g250}
    assert len({i.mat_avg for i in info1 if i.nseqs1 > 0}) > 1
    assert any(not np.array_equal(a, b) for a, b in zip(rows0, rows1))


def test_oracle_matrix_callback_plumbing():
    """A caller series that returns the built-in choice gives the default
    alignment; one that fails ends multiSeqAlign with -1 (MSA.cpp:147-152)."""
    names, codes, offs, lens, dna, dnd = load("100_97")
    w, sim, steps = O.guide_from_dnd(dnd, names)
    rows, order, _ = O.msa(_seqs(codes, offs, lens), sim, w, steps, O.msa_params(dna))
    p = O.set_profile_matrix(O.msa_params(dna), lambda pcid, ml: O.profile_matrix(False, pcid, ml))
    rows2, order2, _ = O.msa(_seqs(codes, offs, lens), sim, w, steps, p)
    np.testing.assert_array_equal(order, order2)
    for a, b in zip(rows, rows2):
        np.testing.assert_array_equal(a, b)
    bad = O.set_profile_matrix(O.msa_params(dna), lambda pcid, ml: (0, np.zeros((32, 32)), 1.0, 0))
    with pytest.raises(RuntimeError):
        O.msa(_seqs(codes, offs, lens), sim, w, steps, bad)


MATRICES_H = "/root/reference/CPU/clustalw-2.0.11/src/substitutionMatrix/matrices.h"


def _tables(text, pattern):
    """name -> list of ints of every `short name[]={...};` / `name[25*25] = {...};` block."""
    import re
    out = {}
    for m in re.finditer(pattern, text, re.S):
        body = re.sub(r"/\*.*?\*/", "", m.group(2), flags=re.S)
        out[m.group(1)] = [int(x) for x in re.findall(r"-?\d+", body)]
    return out


@pytest.mark.skipif(not os.path.exists(MATRICES_H), reason="reference sources not present")
def test_matrix_tables_equal_the_reference_text():
    """The Gonnet series and the two DNA tables that the library and the oracle
    compile (cuda-clustalw_amd/csrc/cw_matrices.inc) against the reference's
    matrices.h, read here as text: lower triangles in the reference's letter
    order (SubMatrix.cpp:179-180), re-indexed by residue code."""
    ref = _tables(open(MATRICES_H).read(), r"short\s+(\w+)\s*\[\]\s*=\s*\{(.*?)\};")
    inc_path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "cuda-clustalw_amd", "csrc", "cw_matrices.inc")
    inc = _tables(open(inc_path).read(), r"static const short\s+(\w+)\s*\[25 \* 25\]\s*=\s*\{(.*?)\};")
    codes = "ABCDEFGHIKLMNOPQRSTUVWXYZ"
    aa, dna = "ABCDEFGHIKLMNPQRSTVWXYZ", "ABCDGHKMNRSTUVWXY"
    pairs = [("gon80mt", "cw_gonnet80", aa), ("gon120mt", "cw_gonnet120", aa),
             ("gon160mt", "cw_gonnet160", aa), ("gon250mt", "cw_gonnet250", aa),
             ("gon350mt", "cw_gonnet350", aa), ("swgapdnamt", "cw_iub_dna", dna),
             ("clustalvdnamt", "cw_clustalv_dna", dna)]
    for rname, iname, order in pairs:
        tri = ref[rname]
        n = len(order)
        assert len(tri) == n * (n + 1) // 2, rname
        full = np.zeros((25, 25), np.int64)
        k = 0
        for i in range(n):
            for j in range(i + 1):
                a, b = codes.index(order[i]), codes.index(order[j])
                full[a, b] = full[b, a] = tri[k]
                k += 1
        got = np.array(inc[iname], np.int64).reshape(25, 25)
        np.testing.assert_array_equal(got, full, err_msg=rname)
