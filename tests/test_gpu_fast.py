"""k-tuple distances on the device (cwhip_fast_pairwise_range, the
FastPairwiseAlign / -QUICKTREE path) against the oracle, bit for bit: the
reference's fixtures (protein and DNA), partial ranges, every parameter the
reference exposes, the partial-alignment cut-off and degenerate sequences."""
import numpy as np
import pytest

import cwhip
import oracle as O
from cwhip import seqgen
from fixtures import load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cwhip.Context(0)
    yield c
    c.close()


def _load(ctx, seqs):
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(len(seqs), np.int64)
    if len(seqs) > 1:
        offs[1:] = np.cumsum(lens[:-1])
    codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if len(seqs) else np.zeros(0, np.uint8)
    ctx.load_sequences(codes, offs, lens)


def _check(ctx, seqs, dna=False, ranges=None, max_aln=None, **kw):
    n = len(seqs)
    _load(ctx, seqs)
    r = ranges if ranges is not None else (0, n, 0, n)
    if max_aln is not None:
        kw["max_aln"] = max_aln
    p, d = ctx.fast_pairwise_range(*r, dna=dna, **kw)
    po, do = O.fast_pairwise([np.asarray(s, np.int32) for s in seqs], dna=dna, ranges=r,
                             max_aln=max_aln, **{k: v for k, v in kw.items() if k != "max_aln"})
    np.testing.assert_array_equal(p, po)
    np.testing.assert_array_equal(d.view(np.uint64), do.view(np.uint64))
    return d


def _fixture(setname, n=None):
    names, codes, offs, lens, dna, dnd = load(setname)
    n = n or len(lens)
    return [codes[offs[i]:offs[i] + lens[i]] for i in range(n)], dna


@pytest.mark.parametrize("setname,n", [("100_97", None), ("100_498", 40), ("1000_97", 150),
                                        ("100_1002", 30), ("100_1523", 24)])
def test_fast_matches_oracle_on_fixtures(ctx, setname, n):
    print("fast", setname, flush=True)
    seqs, dna = _fixture(setname, n)
    d = _check(ctx, seqs, dna)
    assert d.min() >= 0.0 and d.max() <= 1.0


def test_fast_partial_ranges(ctx):
    print("fast ranges", flush=True)
    seqs, dna = _fixture("100_97", 30)
    for r in [(0, 30, 0, 30), (5, 12, 0, 30), (0, 30, 20, 30), (3, 9, 7, 15), (10, 10, 0, 30),
              (0, 30, 29, 30)]:
        _check(ctx, seqs, dna, ranges=r)
    _load(ctx, seqs)
    for bad in [(-1, 5, 0, 5), (5, 4, 0, 5), (0, 5, -1, 5), (0, 5, 5, 4), (0, 31, 0, 30)]:
        with pytest.raises(cwhip.CwhipError):
            ctx.fast_pairwise_range(*bad)


@pytest.mark.parametrize("kw", [dict(ktup=2), dict(window=1), dict(window=12), dict(signif=1),
                                dict(signif=40), dict(wind_gap=0), dict(wind_gap=9), dict(percent=0)])
def test_fast_parameters(ctx, kw):
    print("fast params", kw, flush=True)
    seqs, dna = _fixture("100_498", 25)
    _check(ctx, seqs, dna, **kw)


@pytest.mark.parametrize("ktup", [1, 3, 4])
def test_fast_dna_ktup(ctx, ktup):
    print("fast dna ktup", ktup, flush=True)
    seqs, dna = _fixture("100_1002", 16)
    _check(ctx, seqs, True, ktup=ktup)


def test_fast_partial_alignment_cutoff(ctx):
    """A small maxAlnLength stops pairAlign at 2 x max_aln fragments
    ("(Partial alignment)"), keeping the best fragment so far."""
    print("fast partial", flush=True)
    seqs, dna = _fixture("100_498", 12)
    for m in (1, 5, 40, 300):
        _check(ctx, seqs, dna, max_aln=m)


def test_fast_hbm_lane_kernel(ctx):
    """Pairs whose arrays do not fit a wave's LDS (here: fragment numbers beyond
    16 bits with maxAlnLength 20000) take the one-lane-per-pair kernel with its
    arrays in HBM; same results."""
    print("fast lane kernel", flush=True)
    seqs, dna = _fixture("100_498", 10)
    _check(ctx, seqs, dna, max_aln=20000)
    seqs, dna = _fixture("100_1002", 6)
    _check(ctx, seqs, True, max_aln=20000)


def test_fast_degenerate(ctx):
    print("fast degenerate", flush=True)
    rng = np.random.default_rng(7)
    seqs = [rng.integers(0, 25, size=L).astype(np.uint8) for L in (0, 1, 2, 5, 40, 0, 3)]
    seqs.append(np.full(20, 30, np.uint8))                      # all gaps
    seqs.append(seqs[4].copy())                                 # identical pair
    d = _check(ctx, seqs, False)
    _check(ctx, seqs, False, ktup=2)
    _check(ctx, [np.array([0, 2, 6, 18, 19, 1, 0, 2], np.uint8)] * 3, True)


def test_fast_large_family_sample(ctx):
    """The 1000 x ~1523 workload's first 64 sequences (2,016 pairs of ~1500
    residues: long fragment lists, the partial-alignment limit in reach)."""
    print("fast large", flush=True)
    fam = seqgen.workload("large1523")
    _check(ctx, fam.seqs[:64], False)


def test_fast_two_pass_overflow(ctx):
    """Long low-complexity sequences make thousands of fragments per pair: with
    sequences this long the first pass runs with about two thirds of the
    fragment slots, these pairs run out and are run again with all of them,
    some up to the partial-alignment limit; same results."""
    print("fast two-pass", flush=True)
    rng = np.random.default_rng(11)
    seqs = [rng.choice(np.array([0, 4, 9], np.uint8), size=L) for L in (1400, 1350, 1420, 1380)]
    seqs.append(rng.integers(0, 20, size=1410).astype(np.uint8))
    d = _check(ctx, seqs, False)
    assert d.min() >= 0.0 and d.max() <= 1.0


def _random_case(seed):
    """A seeded family and parameter set: DNA or protein, 2-7 sequences of 0-400
    residues (about one case in seven: 1100-1500, the two-pass path), full or
    2-5 letter alphabets, related or unrelated sequences, a gap code now and
    then, every parameter the reference exposes, now and then a tiny
    maxAlnLength."""
    rng = np.random.default_rng(seed)
    dna = bool(rng.integers(0, 2))
    nseq = int(rng.integers(2, 8))
    longmode = rng.random() < 0.15
    alpha = int(rng.integers(2, 6)) if rng.random() < 0.4 else (4 if dna else 20)
    pool = (np.array([0, 2, 6, 18, 19][:max(2, min(alpha, 5))], np.uint8) if dna
            else rng.choice(25, size=alpha, replace=False).astype(np.uint8))
    size = (lambda: int(rng.integers(1100, 1500))) if longmode else (lambda: int(rng.integers(0, 400)))
    base = rng.choice(pool, size=size())
    seqs = []
    for _ in range(nseq):
        L = size()
        x = base[:L].copy() if (len(base) and rng.random() < 0.6) else rng.choice(pool, size=L)
        if len(x) < L:
            x = np.concatenate([x, rng.choice(pool, size=L - len(x))])
        m = rng.random(len(x)) < rng.random() * 0.5
        x[m] = rng.choice(pool, size=int(m.sum()))
        if rng.random() < 0.1 and len(x):
            x[rng.integers(0, len(x))] = 30
        seqs.append(x.astype(np.uint8))
    kw = dict(ktup=int(rng.integers(1, (4 if dna else 2) + 1)), window=int(rng.integers(0, 13)),
              wind_gap=int(rng.integers(0, 10)), signif=int(rng.integers(1, 41)),
              percent=int(rng.integers(0, 2)))
    max_aln = int(rng.integers(1, 60)) if rng.random() < 0.15 else None
    return seqs, dna, kw, max_aln


def test_fast_random_cases(ctx):
    """300 seeded random families and parameter sets, bit for bit."""
    print("fast random", flush=True)
    for seed in range(300):
        seqs, dna, kw, max_aln = _random_case(seed)
        _check(ctx, seqs, dna, max_aln=max_aln, **kw)


def test_fast_block_chains(ctx):
    """The closed-form block layouts of tests/test_fast_oracle.py (one block,
    two chained blocks, the same diagonal, crossing blocks, top diagonals and
    their window) on the device, against the oracle, under the parameters that
    change their scores."""
    from test_fast_oracle import B1, B2, F1, F2, _two_blocks
    single = (np.concatenate([np.full(20, F1), B1, np.full(15, F1)]).astype(np.uint8),
              np.concatenate([np.full(5, F2), B1, np.full(30, F2)]).astype(np.uint8))
    cross = (np.concatenate([np.full(10, F1), B1, np.full(4, F1), B2, np.full(10, F1)]).astype(np.uint8),
             np.concatenate([np.full(10, F2), B2, np.full(4, F2), B1, np.full(10, F2)]).astype(np.uint8))
    layouts = [single, cross] + [tuple(x.astype(np.uint8) for x in _two_blocks(4, g)) for g in (4, 7, 16)]
    for s1, s2 in layouts:
        for kw in (dict(), dict(wind_gap=1), dict(wind_gap=6), dict(signif=1), dict(signif=2),
                   dict(signif=2, window=2), dict(signif=3), dict(signif=2, window=12), dict(ktup=2),
                   dict(percent=0)):
            _check(ctx, [s1, s2], **kw)
