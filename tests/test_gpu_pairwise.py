"""GPU parity of the pairwise kernels against the CPU restatement (the oracle).

Everything is compared bit for bit: identity counts, the float distance, and
the intermediate cells (maxScore, se1, se2, sb1, sb2) of every pair.
"""
import numpy as np
import pytest

import cwhip
from cwhip import seqgen
import oracle as O

pytestmark = pytest.mark.gpu

SC = O.Scoring()
FULL = [0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 17, 18, 20, 21, 23]


@pytest.fixture(scope="module")
def ctx():
    c = cwhip.Context(0)
    yield c
    c.close()


def pack(seqs):
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(len(seqs), np.int64)
    if len(seqs) > 1:
        offs[1:] = np.cumsum(lens[:-1])
    codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if lens.sum() else np.zeros(1, np.uint8)
    return codes.astype(np.uint8), offs, lens


def all_pairs(n):
    return O.range_pairs(n, 0, n, 0, n)


def run_and_compare(ctx, codes, offs, lens, pairs, scoring=SC, params=None,
                    kernels=(cwhip.KERNEL_PLAIN, cwhip.KERNEL_BATCHED)):
    if params is None:
        params = cwhip.pair_params(max_len=max(int(lens.max()), 1))
    cnt_o, dist_o, res = O.pairwise_pairs(codes, offs, lens, pairs, scoring, nthreads=16,
                                          want_results=True)
    info_o = np.array([[r.max_score, r.se1, r.se2, r.sb1, r.sb2] for r in res][:len(pairs)],
                      np.int32).reshape(-1, 5)
    mm_o = np.array([r.cells_mm for r in res][:len(pairs)], np.int64)
    ctx.load_sequences(codes, offs, lens)
    out = []
    for k in kernels:
        ctx.stage(params, pairs)
        ctx.run_staged(kernel=k)
        cnt = ctx.fetch_counts()
        info = ctx.fetch_info()
        np.testing.assert_array_equal(cnt, cnt_o, err_msg="identity counts, kernel %d" % k)
        nonempty = cnt_o >= 0
        np.testing.assert_array_equal(info[nonempty, :5], info_o[nonempty],
                                      err_msg="maxScore/se/sb, kernel %d" % k)
        assert (info[:, 5] == 0).all()
        np.testing.assert_array_equal(info[nonempty, 7].view(np.uint32).astype(np.int64), mm_o[nonempty])
        c2, d2 = ctx.pairwise_pairs(params, pairs, kernel=k)
        np.testing.assert_array_equal(c2, cnt_o)
        np.testing.assert_array_equal(d2.view(np.uint32), dist_o.view(np.uint32))
        out.append(cnt)
    return out


def test_small97_all_pairs(ctx):
    fam = seqgen.workload("small97")
    codes, offs, lens = fam.packed()
    run_and_compare(ctx, codes, offs, lens, all_pairs(len(lens)))


@pytest.mark.parametrize("rmax,wpg", [(1, 4), (2, 1), (3, 2), (5, 4), (8, 4)])
def test_stripe_shapes_give_identical_results(ctx, rmax, wpg):
    # every stripe decomposition (direct single-lane output, dummy first
    # registers, multi-stripe ping-pong) must reproduce the oracle
    rng = np.random.default_rng(100 + rmax)
    seqs = [rng.choice(FULL, int(L)).astype(np.uint8) for L in
            [1, 2, 63, 64, 65, 127, 128, 129, 300, 513, 700, 1025]]
    fam = seqgen.family(77, 6, 400)
    seqs += fam.seqs
    codes, offs, lens = pack(seqs)
    ctx.set_tuning(rmax=rmax, waves_per_group=wpg)
    try:
        run_and_compare(ctx, codes, offs, lens, all_pairs(len(seqs)))
    finally:
        ctx.set_tuning(rmax=8, waves_per_group=4)


def test_ties_homopolymers_and_small_alphabets(ctx):
    rng = np.random.default_rng(7)
    seqs = []
    for L in (1, 2, 3, 5, 17, 64, 65, 150, 333):
        seqs.append(np.full(L, 10, np.uint8))                       # LLLL...
    for L in (4, 9, 70, 200, 260):
        seqs.append(rng.choice([0, 6], L).astype(np.uint8))         # A/G only
        seqs.append(np.tile(np.array([0, 6], np.uint8), L)[:L])     # AGAG...
    for L in (30, 90):
        seqs.append(rng.choice([0, 17, 18], L).astype(np.uint8))
    codes, offs, lens = pack(seqs)
    run_and_compare(ctx, codes, offs, lens, all_pairs(len(seqs)))


def test_empty_gaps_unrelated_and_identical(ctx):
    # empty sequences (distance 1.0, FullPairwiseAlign.cpp:113-118), a gap-only
    # sequence and '-' codes inside sequences (excluded from len1/len2,
    # :100-108, :119-127), a one-residue sequence; the empty pairs fall in the
    # last bin of the batched queue, after pairs that run the full DP
    rng = np.random.default_rng(11)
    base = rng.choice(FULL, 400).astype(np.uint8)
    gapped = base.copy()
    gapped[rng.choice(400, 40, replace=False)] = 31                 # '-' codes in the input
    seqs = [np.zeros(0, np.uint8), base, base.copy(), gapped, rng.choice(FULL, 380).astype(np.uint8),
            np.array([21], np.uint8), np.zeros(0, np.uint8), base[::-1].copy(),
            np.full(50, 31, np.uint8), np.concatenate([base[:100], np.full(3, 31, np.uint8), base[100:]])]
    codes, offs, lens = pack(seqs)
    run_and_compare(ctx, codes, offs, lens, all_pairs(len(seqs)))


def test_long_sequences_multi_stripe(ctx):
    fam = seqgen.family(2024, 5, 4000)
    codes, offs, lens = fam.packed()
    assert lens.min() > 3500
    run_and_compare(ctx, codes, offs, lens, all_pairs(len(lens)))


def test_longest_sequences_fewer_waves_per_workgroup(ctx):
    # up to CWHIP_MAX_LEN residues: two sequences' LDS buffers per wave leave
    # room for two waves per workgroup (four below ~16000); one longer is refused
    fam = seqgen.family(31, 3, 24000)
    rng = np.random.default_rng(31)
    s0 = np.concatenate([fam.seqs[0], rng.choice(FULL, 40000).astype(np.uint8)])[:cwhip.MAX_LEN]
    seqs = [s0, fam.seqs[1], fam.seqs[2][:16500]]
    codes, offs, lens = pack(seqs)
    assert lens.max() == cwhip.MAX_LEN and lens.min() == 16500
    print("longest sequences", lens.tolist(), flush=True)
    run_and_compare(ctx, codes, offs, lens, all_pairs(len(seqs)))
    codes, offs, lens = pack([np.zeros(cwhip.MAX_LEN + 1, np.uint8), s0])
    with pytest.raises(cwhip.CwhipError):
        ctx.load_sequences(codes, offs, lens)


def test_large1523_sample(ctx):
    fam = seqgen.workload("large1523")
    codes, offs, lens = fam.packed()
    rng = np.random.default_rng(5)
    pairs = all_pairs(len(lens))[rng.choice(499500, 400, replace=False)]
    run_and_compare(ctx, codes, offs, lens, pairs)


def test_uneven_sample_with_longest(ctx):
    fam = seqgen.workload("uneven")
    codes, offs, lens = fam.packed()
    rng = np.random.default_rng(6)
    longest = np.argsort(lens)[-6:]
    shortest = np.argsort(lens)[:4]
    extra = np.array([(int(a), int(b)) for a in longest for b in np.concatenate([longest, shortest]) if a < b],
                     np.int32).reshape(-1, 2)
    pairs = np.concatenate([all_pairs(len(lens))[rng.choice(1999000, 300, replace=False)], extra])
    run_and_compare(ctx, codes, offs, lens, pairs)


@pytest.mark.parametrize("r", [(0, 30, 0, 30), (5, 12, 0, 30), (0, 30, 27, 30), (29, 30, 0, 30),
                               (3, 3, 0, 30), (10, 40, 20, 25)])
def test_range_semantics(ctx, r):
    fam = seqgen.family(99, 30, 120)
    codes, offs, lens = fam.packed()
    ctx.load_sequences(codes, offs, lens)
    params = cwhip.pair_params(max_len=int(lens.max()))
    pr, cnt, dist = ctx.pairwise_range(params, *r)
    want = O.range_pairs(30, *r)
    np.testing.assert_array_equal(pr, want)
    if len(want):
        co, do = O.pairwise_pairs(codes, offs, lens, want, SC)
        np.testing.assert_array_equal(cnt, co)
        np.testing.assert_array_equal(dist.view(np.uint32), do.view(np.uint32))


def test_range_errors(ctx):
    fam = seqgen.family(1, 4, 20)
    ctx.load_sequences(*fam.packed())
    params = cwhip.pair_params(max_len=40)
    for bad in [(-1, 3, 0, 3), (3, 2, 0, 3), (0, 3, -2, 3), (0, 3, 3, 2)]:
        with pytest.raises(cwhip.CwhipError):
            ctx.pairwise_range(params, *bad)


def test_dna_scoring(ctx):
    rng = np.random.default_rng(8)
    nuc = [0, 2, 6, 18]                                              # A C G T
    root = rng.choice(nuc, 300)
    seqs = []
    for k in range(12):
        s = root.copy()
        mut = rng.random(300) < 0.1 * (k + 1) / 4
        s[mut] = rng.choice(nuc, mut.sum())
        seqs.append(s[rng.integers(0, 20):].astype(np.uint8))
    codes, offs, lens = pack(seqs)
    dsc = O.Scoring(dna=True)
    params = cwhip.pair_params(dna=True, max_len=int(lens.max()))
    run_and_compare(ctx, codes, offs, lens, all_pairs(len(seqs)), scoring=dsc, params=params)


def test_full_large1523_batched_equals_plain(ctx):
    """All 499,500 pairs of the 1000 x ~1523 workload: both kernels agree bit for
    bit, every pair is consistent, and a strided sample matches the oracle."""
    fam = seqgen.workload("large1523")
    codes, offs, lens = fam.packed()
    pairs = all_pairs(len(lens))
    ctx.load_sequences(codes, offs, lens)
    params = cwhip.pair_params(max_len=int(lens.max()))
    ctx.stage(params, pairs)
    ctx.run_staged(kernel=cwhip.KERNEL_BATCHED)
    cb = ctx.fetch_counts()
    ib = ctx.fetch_info()
    ctx.run_staged(kernel=cwhip.KERNEL_PLAIN)
    cp = ctx.fetch_counts()
    np.testing.assert_array_equal(cb, cp)
    assert (ib[:, 5] == 0).all() and (cb >= 0).all()
    sample = np.arange(0, len(pairs), 1249)
    co, _ = O.pairwise_pairs(codes, offs, lens, pairs[sample], SC, nthreads=16)
    np.testing.assert_array_equal(cb[sample], co)
