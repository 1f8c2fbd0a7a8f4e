"""Profile-alignment parity on the GPU: the device Myers-Miller DP
(cwhip_profile_dp) against the oracle's progDiff, bit for bit (score and edit
script), and the whole progressive stage (cwhip_msa) against the reference's
.aln fixtures and against the oracle on divergent synthetic sets."""
import numpy as np
import pytest

import cwhip
import oracle as O
from cwhip import seqgen
from fixtures import load, parse_aln, row_text, text
from test_msa_oracle import random_profile_pairs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cwhip.Context(0)
    yield c
    c.close()


def _check_batch(ctx, probs, egp=0):
    got = ctx.profile_dp([(p1, l1, p2, l2) for p1, l1, p2, l2 in probs], end_gap_penalties=egp)
    for (p1, l1, p2, l2), (score, displ) in zip(probs, got):
        s_ref, d_ref, _ = O.prog_diff(p1, l1, p2, l2, egp)
        assert score == s_ref, (l1, l2)
        np.testing.assert_array_equal(displ, d_ref)


@pytest.mark.parametrize("setname", ["100_97", "100_498", "100_1002", "100_1523"])
def test_profile_dp_matches_oracle(ctx, setname):
    print("profile_dp", setname, flush=True)
    _check_batch(ctx, random_profile_pairs(setname, 12, 7))


def test_profile_dp_end_gap_penalties(ctx):
    print("profile_dp egp", flush=True)
    _check_batch(ctx, random_profile_pairs("100_498", 6, 9, egp=1), egp=1)


def test_profile_dp_small_and_degenerate(ctx):
    """Lengths 0..3 and very unequal lengths (M <= 1 leaves at the root,
    single-row sweeps, empty profiles)."""
    print("profile_dp small", flush=True)
    names, codes, offs, lens, dna, dnd = load("100_97")
    seqs = [codes[offs[i]:offs[i] + lens[i]].astype(np.int32) for i in range(len(lens))]
    n = len(seqs)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    probs = []
    for la, lb in [(1, 1), (1, 5), (5, 1), (2, 3), (3, 97), (97, 2), (0, 4), (4, 0), (1, 97)]:
        s2 = [s.copy() for s in seqs]
        s2[0] = s2[0][:la]
        s2[1] = s2[1][:lb]
        group = np.zeros(n + 1, np.int32)
        group[1], group[2] = 1, 2
        r = O.step_profiles(s2, sim, w, group, np.ones(n + 1, np.int32), O.msa_params(False))
        if r is not None:
            probs.append(r)
    _check_batch(ctx, probs)


SETS = ["100_97", "100_498", "1000_97", "100_1002", "100_1523", "1000_498"]


@pytest.mark.parametrize("setname", SETS)
def test_gpu_msa_reproduces_reference_aln(ctx, setname):
    print("msa", setname, flush=True)
    names, codes, offs, lens, dna, dnd = load(setname)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    seqs = [codes[offs[i]:offs[i] + lens[i]] for i in range(len(lens))]
    rows, order, stats = ctx.msa(seqs, sim, w, steps, dna=dna)
    ref_order, ref_rows = parse_aln(text(setname + ".aln"))
    assert [names[i - 1] for i in order] == ref_order
    for i, nm in enumerate(names):
        assert row_text(rows[i]) == ref_rows[nm], nm
    assert stats["calls"] == len(names) - 1


def _divergent(seed, n1, n2, L):
    a = seqgen.family(seed, n1, L, branch_mean=0.6, name_prefix="fa")
    b = seqgen.family(seed + 1, n2, int(L * 0.8), branch_mean=1.5, name_prefix="fb")
    return a.seqs + b.seqs, a.names + b.names


@pytest.mark.parametrize("seed,n1,n2,L,dna", [(11, 12, 8, 120, False), (12, 20, 10, 200, False),
                                               (13, 5, 25, 150, False), (14, 10, 10, 300, True)])
def test_gpu_msa_matches_oracle_divergent(ctx, seed, n1, n2, L, dna):
    """Sets with unrelated members: the delayed (divergent) pass of
    multiSeqAlign, steps with an emptied group, and every matrix of the
    gonnet series."""
    print("msa divergent", seed, flush=True)
    seqs, names = _divergent(seed, n1, n2, L)
    if dna:
        seqs = [np.array([(0, 2, 6, 18)[c % 4] for c in s], np.uint8) for s in seqs]
    n = len(seqs)
    codes = np.concatenate(seqs)
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(n, np.int64)
    offs[1:] = np.cumsum(lens[:-1])
    pairs = O.range_pairs(n, 0, n, 0, n)
    cnt, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring(dna))
    dnd = O.nj_dnd(O.distance_matrix(n, pairs, dist), names)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    r_rows, r_order, info = O.msa([s.astype(np.int32) for s in seqs], sim, w, steps,
                                  O.msa_params(dna))
    rows, order, stats = ctx.msa(seqs, sim, w, steps, dna=dna)
    np.testing.assert_array_equal(order, r_order)
    for i in range(n):
        np.testing.assert_array_equal(rows[i], r_rows[i].astype(np.uint8))


@pytest.mark.parametrize("setname", SETS)
def test_gpu_pipeline_reproduces_reference_dnd_and_aln(ctx, setname):
    """FASTA -> device pairwise -> NJ -> guide -> device progressive alignment
    -> CLUSTAL text, all on the library: the reference's .dnd and .aln files
    byte for byte."""
    from cwhip import pipeline
    print("pipeline", setname, flush=True)
    names, codes, offs, lens, dna, dnd = load(setname)
    seqs = [codes[offs[i]:offs[i] + lens[i]] for i in range(len(lens))]
    out_dnd, out_aln, rows, order = pipeline.align(ctx, names, seqs, dna)
    assert out_dnd == dnd
    assert out_aln == text(setname + ".aln")


@pytest.mark.parametrize("setname", ["100_97", "100_498", "100_1002"])
def test_gpu_profile_align_step_by_step(ctx, setname):
    """cwhip_profile_align called once per guide-tree step, as MSA.cpp:80-157
    calls profileAlign (the fixtures have no divergent sequences, so the first
    pass aligns everything): the final rows are the reference's."""
    from cwhip.lib import msa_params, profile_align
    print("profile_align steps", setname, flush=True)
    names, codes, offs, lens, dna, dnd = load(setname)
    n = len(names)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    seqs = [codes[offs[i]:offs[i] + lens[i]].copy() for i in range(n)]
    aligned = np.zeros(n + 1, np.int32)
    p = msa_params(dna)
    for s in steps:
        for i in range(1, n + 1):
            if s[i] != 0:
                aligned[i] = 1
        info, g, new = profile_align(ctx, seqs, sim, w, s, aligned, p)
        assert info[0] >= 0
        for i in range(n):
            if new[i] is not None:
                seqs[i] = new[i]
    ref_order, ref_rows = parse_aln(text(setname + ".aln"))
    for i, nm in enumerate(names):
        assert row_text(seqs[i]) == ref_rows[nm], nm


def test_gpu_msa_random_families(ctx):
    """40 seeded random sets: two families of 1-19 and 1-14 members, lengths
    15-350 and 0.3-1.3 times that, random divergence, a third of them DNA; the
    device alignment equals the oracle's row for row."""
    for seed in range(40):
        rng = np.random.default_rng(seed)
        dna = rng.random() < 0.3
        n1, n2 = int(rng.integers(1, 20)), int(rng.integers(1, 15))
        L = int(rng.integers(15, 350))
        a = seqgen.family(seed * 7 + 1, n1, L, branch_mean=float(rng.uniform(0.1, 1.0)), name_prefix="fa")
        b = seqgen.family(seed * 7 + 2, n2, max(5, int(L * rng.uniform(0.3, 1.3))),
                          branch_mean=float(rng.uniform(0.3, 2.0)), name_prefix="fb")
        seqs, names = a.seqs + b.seqs, a.names + b.names
        if dna:
            seqs = [np.array([(0, 2, 6, 18)[c % 4] for c in s], np.uint8) for s in seqs]
        n = len(seqs)
        print("msa random", seed, n, L, dna, flush=True)
        codes = np.concatenate(seqs)
        lens = np.array([len(s) for s in seqs], np.int32)
        offs = np.zeros(n, np.int64)
        offs[1:] = np.cumsum(lens[:-1])
        pairs = O.range_pairs(n, 0, n, 0, n)
        cnt, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring(dna))
        dnd = O.nj_dnd(O.distance_matrix(n, pairs, dist), names)
        w, sim, steps = O.guide_from_dnd(dnd, names)
        r_rows, r_order, info = O.msa([s.astype(np.int32) for s in seqs], sim, w, steps, O.msa_params(dna))
        rows, order, stats = ctx.msa(seqs, sim, w, steps, dna=dna)
        np.testing.assert_array_equal(order, r_order)
        for i in range(n):
            np.testing.assert_array_equal(rows[i], r_rows[i].astype(np.uint8))


def test_delayed_pass_is_one_profile_align_step(ctx):
    """MSA::multiSeqAlign's second pass (MSA.cpp:187-335) for one divergent
    sequence X: the first pass aligns the others, then X is aligned to all of
    them in one profileAlign step with the weights recomputed as tree weight x
    identity with X, normalised to INT_SCALE_FACTOR (:266-296).  So the final
    alignment equals one cwhip_profile_align step of X against the first-pass
    rows (the final rows without X and without the columns only X fills), with
    those weights, and X comes last in the output order."""
    fam = seqgen.family(4711, 9, 180, branch_mean=0.15, name_prefix="d")
    rng = np.random.default_rng(4711)
    x = rng.choice([0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 17, 18, 20, 21, 23], 170).astype(np.uint8)
    seqs, names = fam.seqs + [x], fam.names + ["xout"]
    n = len(seqs)
    codes = np.concatenate(seqs)
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(n, np.int64)
    offs[1:] = np.cumsum(lens[:-1])
    pairs = O.range_pairs(n, 0, n, 0, n)
    _, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring())
    dnd = O.nj_dnd(O.distance_matrix(n, pairs, dist), names)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    maxid = [max(int(sim[i, j]) for j in range(1, n + 1) if j != i) for i in range(1, n + 1)]
    assert maxid[-1] <= 30 and min(maxid[:-1]) > 30         # X alone is delayed
    rows, order, stats = ctx.msa(seqs, sim, w, steps)
    r_rows, r_order, _ = O.msa([s.astype(np.int32) for s in seqs], sim, w, steps, O.msa_params(False))
    np.testing.assert_array_equal(order, r_order)
    assert order[-1] == n
    S = np.array(rows[:-1])
    first_pass = S[:, (S < 30).any(axis=0)]
    wt = np.array(w, np.int64).copy()
    wt[:-1] = (w[:-1].astype(np.float64) * sim[1:n, n]).astype(np.int64)
    total = int(wt[:-1].sum())
    if total == 0:
        wt[:] = 1
        total = n
    wt[:-1] = np.maximum((wt[:-1] * 1000) // total, 1)
    group = np.zeros(n + 1, np.int32)
    group[1:n] = 1
    group[n] = 2
    from cwhip.lib import msa_params, profile_align
    info, g, new = profile_align(ctx, [r for r in first_pass] + [x], sim, wt.astype(np.int32), group,
                                 np.ones(n + 1, np.int32), msa_params(False))
    assert info[0] >= 0
    for i in range(n):
        np.testing.assert_array_equal(new[i], rows[i])


def test_msa_schedules_agree():
    """Every schedule of the profile alignments' Myers-Miller levels gives the
    oracle's alignment: the root's half sweeps as 2..5 pipelined stripes on
    separate workgroups or in the pair's workgroup, 1..6 levels swept over the
    chip, and batches left entirely to one workgroup per pair (the schedule of
    batches of more than 512 pairs).  Two families of ~600 residues make roots
    of 256+ rows from the first level on."""
    a = seqgen.family(901, 14, 640, branch_mean=0.3, name_prefix="sa")
    b = seqgen.family(902, 10, 560, branch_mean=0.5, name_prefix="sb")
    seqs, names = a.seqs + b.seqs, a.names + b.names
    n = len(seqs)
    codes = np.concatenate(seqs)
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(n, np.int64)
    offs[1:] = np.cumsum(lens[:-1])
    pairs = O.range_pairs(n, 0, n, 0, n)
    cnt, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring(False))
    dnd = O.nj_dnd(O.distance_matrix(n, pairs, dist), names)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    r_rows, r_order, info = O.msa([s.astype(np.int32) for s in seqs], sim, w, steps, O.msa_params(False))
    schedules = [(256, 5, 512), (64, 5, 512), (128, 6, 512), (0, 3, 512), (256, 1, 512), (256, 5, 0)]
    c = cwhip.Context(0)
    try:
        for wide_rows, wide_levels, wide_pairs in schedules:
            print("msa schedule", wide_rows, wide_levels, wide_pairs, flush=True)
            c.set_msa_tuning(wide_rows, wide_levels, wide_pairs)
            rows, order, stats = c.msa(seqs, sim, w, steps)
            np.testing.assert_array_equal(order, r_order)
            for i in range(n):
                np.testing.assert_array_equal(rows[i], r_rows[i].astype(np.uint8))
    finally:
        c.close()


def test_msa_stripe_wait_limit_reports_error():
    """A pipelined root stripe whose wait for the stripe above runs out gives up
    (every access stays in bounds; the others stop on the shared error flag),
    and cwhip_msa fails with a kernel error instead of hanging; the same context
    then aligns as the oracle does.  A one-sleep limit runs out before the
    stripe above has flushed its first 64 columns (~127 of its steps)."""
    a = seqgen.family(911, 6, 700, branch_mean=0.3, name_prefix="wa")
    b = seqgen.family(912, 5, 650, branch_mean=0.4, name_prefix="wb")
    seqs, names = a.seqs + b.seqs, a.names + b.names
    n = len(seqs)
    codes = np.concatenate(seqs)
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(n, np.int64)
    offs[1:] = np.cumsum(lens[:-1])
    pairs = O.range_pairs(n, 0, n, 0, n)
    cnt, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring(False))
    dnd = O.nj_dnd(O.distance_matrix(n, pairs, dist), names)
    w, sim, steps = O.guide_from_dnd(dnd, names)
    r_rows, r_order, info = O.msa([s.astype(np.int32) for s in seqs], sim, w, steps, O.msa_params(False))
    c = cwhip.Context(0)
    try:
        c.set_msa_tuning(wide_rows=256, wide_levels=5, wide_pairs=512, wait_limit=1)
        with pytest.raises(cwhip.CwhipError):
            c.msa(seqs, sim, w, steps)
        c.set_msa_tuning(wait_limit=0)
        rows, order, stats = c.msa(seqs, sim, w, steps)
        np.testing.assert_array_equal(order, r_order)
        for i in range(n):
            np.testing.assert_array_equal(rows[i], r_rows[i].astype(np.uint8))
    finally:
        c.close()
