"""The profile matrix of each progressive step on the device path against the
oracle, row for row:

* UserParameters::distanceTree false (a guide tree without branch lengths,
  tree/Tree.cpp:561-565; the profile-to-profile path, multipleAlign/MSA.cpp:572)
  fixes the matrix of every step (gon250, substitutionMatrix/SubMatrix.cpp:565);
* a caller-supplied matrix series (cwhip_profile_matrix_fn, the route by which
  the C++ adapter hands over SubMatrix::getProfileAlignMatrix for blosum, pam,
  id and user matrices) replaces the built-in one, and its failure ends the
  alignment as MyersMillerProfileAlign.cpp:258-263 and MSA.cpp:147-152 do.
"""
import numpy as np
import pytest

import cwhip
import oracle as O
from cwhip import lib as cl
from cwhip import pipeline, seqgen
from fixtures import load
from test_msa_oracle import divergent_set

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cwhip.Context(0)
    yield c
    c.close()


def rooted_tree_without_lengths(names):
    """A rooted binary guide tree over the names, no branch lengths."""
    def build(lo, hi):
        if hi - lo == 1:
            return names[lo]
        mid = (lo + hi) // 2
        return "(%s,%s)" % (build(lo, mid), build(mid, hi))
    return build(0, len(names)) + ";\n"


def identity_or_gonnet(pcid, min_len):
    """A two-matrix series: an identity matrix (x100, mean off-diagonal 0, so
    the gap coefficient takes its matAvg <= 0 form) above 40% identity, the
    built-in gonnet choice below."""
    if pcid > 40.0:
        m = np.zeros((32, 32), np.int32)
        for c in range(25):
            m[c, c] = 100
        return 24, m, 0.75, 0
    return O.profile_matrix(False, pcid, min_len, distance_tree=1)


def fixture_seqs(setname):
    names, codes, offs, lens, dna, dnd = load(setname)
    return names, [codes[offs[i]:offs[i] + lens[i]].copy() for i in range(len(names))], dna, dnd


def compare(rows, order, r_rows, r_order):
    np.testing.assert_array_equal(order, r_order)
    for i in range(len(rows)):
        np.testing.assert_array_equal(rows[i], r_rows[i].astype(np.uint8))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_distance_tree_false_takes_gon250(ctx, seed):
    print("distance_tree 0, divergent set", seed, flush=True)
    seqs, sim, w, steps = divergent_set(seed)
    r_rows, r_order, _ = O.msa(seqs, sim, w, steps, O.msa_params(False, distance_tree=0))
    rows, order, _ = ctx.msa([s.astype(np.uint8) for s in seqs], sim, w, steps,
                             params=cwhip.msa_params(False, distance_tree=0))
    compare(rows, order, r_rows, r_order)


def test_user_tree_without_branch_lengths(ctx):
    print("rooted tree without lengths, 100_97", flush=True)
    names, seqs, dna, _ = fixture_seqs("100_97")
    tree = rooted_tree_without_lengths(names)
    w, sim, steps, dtree = cwhip.guide_from_tree(tree, names, with_distance_tree=True)
    ow, osim, osteps = O.guide_from_dnd(tree, names)
    assert dtree == 0
    np.testing.assert_array_equal(w, ow)
    np.testing.assert_array_equal(steps, osteps)
    np.testing.assert_array_equal(sim, osim)
    _, _, rows, order = pipeline.align(ctx, names, seqs, dna, guide_dnd=tree)
    r_rows, r_order, _ = O.msa([s.astype(np.int32) for s in seqs], osim, ow, osteps,
                               O.msa_params(dna, distance_tree=0))
    compare(rows, order, r_rows, r_order)


@pytest.mark.parametrize("setname", ["100_97", "100_498"])
def test_caller_matrix_series(ctx, setname):
    print("caller matrix series", setname, flush=True)
    names, seqs, dna, dnd = fixture_seqs(setname)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    calls = []

    def series(pcid, min_len):
        calls.append(pcid)
        return identity_or_gonnet(pcid, min_len)

    r_rows, r_order, info = O.msa([s.astype(np.int32) for s in seqs], sim, w, steps,
                                  O.set_profile_matrix(O.msa_params(dna), identity_or_gonnet))
    rows, order, stats = ctx.msa(seqs, sim, w, steps,
                                 params=cl.set_profile_matrix(cwhip.msa_params(dna), series))
    compare(rows, order, r_rows, r_order)
    assert len(calls) == stats["calls"] == len(info)


def test_caller_matrix_series_random_families(ctx):
    """Divergent sets, so both matrices of the series are taken."""
    seen = []

    def series(pcid, min_len):
        seen.append(pcid)
        return identity_or_gonnet(pcid, min_len)

    for seed in range(6):
        rng = np.random.default_rng(500 + seed)
        a = seqgen.family(seed * 11 + 3, int(rng.integers(3, 15)), int(rng.integers(40, 300)),
                          branch_mean=0.2, name_prefix="ma")
        b = seqgen.family(seed * 11 + 4, int(rng.integers(2, 10)), int(rng.integers(40, 300)),
                          branch_mean=1.5, name_prefix="mb")
        seqs, names = a.seqs + b.seqs, a.names + b.names
        n = len(seqs)
        print("caller matrix series random", seed, n, flush=True)
        codes = np.concatenate(seqs)
        lens = np.array([len(s) for s in seqs], np.int32)
        offs = np.zeros(n, np.int64)
        offs[1:] = np.cumsum(lens[:-1])
        pairs = O.range_pairs(n, 0, n, 0, n)
        _, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring())
        dnd = O.nj_dnd(O.distance_matrix(n, pairs, dist), names)
        w, sim, steps = O.guide_from_dnd(dnd, names)
        r_rows, r_order, _ = O.msa([s.astype(np.int32) for s in seqs], sim, w, steps,
                                   O.set_profile_matrix(O.msa_params(False), identity_or_gonnet))
        rows, order, _ = ctx.msa(seqs, sim, w, steps,
                                 params=cl.set_profile_matrix(cwhip.msa_params(False), series))
        compare(rows, order, r_rows, r_order)
    assert any(p > 40.0 for p in seen) and any(p <= 40.0 for p in seen)


def test_caller_matrix_failure_ends_the_alignment(ctx):
    print("caller matrix failure", flush=True)
    names, seqs, dna, dnd = fixture_seqs("100_97")
    w, sim, steps = O.guide_from_dnd(dnd, names)

    def failing(pcid, min_len):
        if pcid < 95.0:                          # a few of this set's steps
            return 0, np.zeros((32, 32), np.int32), 1.0, 0
        return O.profile_matrix(False, pcid, min_len)

    with pytest.raises(cwhip.CwhipError, match=r"\(-8\)"):
        ctx.msa(seqs, sim, w, steps, params=cl.set_profile_matrix(cwhip.msa_params(dna), failing))
    with pytest.raises(RuntimeError):
        O.msa([s.astype(np.int32) for s in seqs], sim, w, steps,
              O.set_profile_matrix(O.msa_params(dna), failing))
    # one step on its own: profileAlign returns -1 and aligns nothing
    def fail_always(pcid, min_len):
        return -1, np.zeros((32, 32), np.int32), 1.0, 0

    aligned = np.ones(len(names) + 1, np.int32)
    info, _, new = cl.profile_align(ctx, seqs, sim, w, steps[-1], aligned,
                                    cl.set_profile_matrix(cwhip.msa_params(dna), fail_always))
    assert info[0] == -1 and info[1] == 0 and all(x is None for x in new)
