"""cwhip_pairwise_noalloc (the pairwise stage with its work space in a file-scope
device array, for callers that cannot allocate) against cwhip_pairwise_pairs
with the batched kernel, through the C ABI.

The test itself allocates the caller's device memory (inputs and outputs) with
the HIP runtime libcwhip.so is linked against; the entry point allocates none.
Two calls in a row, the second with fewer, shorter sequences and fewer pairs,
must each equal the existing entry point's identity counts bit for bit (and the
oracle's); a pair naming a sequence outside the set or longer than max_len is
refused by the kernels' own checks with count -2 and status bit 4.
"""
import ctypes as C

import numpy as np
import pytest

import cwhip
from cwhip import lib as cl
from cwhip import seqgen
import oracle as O

pytestmark = pytest.mark.gpu

FULL = [0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 17, 18, 20, 21, 23]
H2D, D2H = 1, 2


class Hip:
    """The few HIP runtime calls the caller side of the test needs."""

    def __init__(self):
        cl.load()                                   # libcwhip.so brings its HIP runtime in
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.L.hipSetDevice.argtypes = [C.c_int]
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, "%s returned %d" % (what, rc)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), max(arr.nbytes, 4)), "hipMalloc")
        self.bufs.append(p)
        if arr.nbytes:
            self.check(self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")
        return p

    def download(self, p, dtype, n):
        out = np.zeros(max(n, 1), dtype)
        if n:
            self.check(self.L.hipMemcpy(out.ctypes.data, p, out[:n].nbytes, D2H), "hipMemcpy D2H")
        return out[:n]

    def free_all(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def pack(seqs):
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(len(seqs), np.int64)
    if len(seqs) > 1:
        offs[1:] = np.cumsum(lens[:-1])
    codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if lens.sum() else np.zeros(1, np.uint8)
    return codes.astype(np.uint8), offs, lens


def run_noalloc(hip, params, codes, offs, lens, pairs, max_len=None):
    """Inputs to device memory, one cwhip_pairwise_noalloc call on the null
    stream, a device synchronisation, the counts and the status back."""
    L = cl.load()
    max_len = int(lens.max()) if max_len is None else max_len
    d_mat = hip.upload(np.ctypeslib.as_array(params.matrix).astype(np.int32))
    d_go = hip.upload(params._keep[:max_len + 1].astype(np.int32))
    dp = cl.PairDevParams(d_mat, d_go, params.gap_extend, params.gap_pos1, params.gap_pos2, max_len)
    d_codes, d_offs, d_lens = hip.upload(codes), hip.upload(offs), hip.upload(lens)
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    d_pairs = hip.upload(pairs)
    d_cnt = hip.upload(np.full(len(pairs), 12345, np.int32))
    d_st = hip.upload(np.full(1, 777, np.int32))
    rc = L.cwhip_pairwise_noalloc(C.byref(dp), d_codes, d_offs, d_lens, len(lens), d_pairs,
                                  len(pairs), d_cnt, d_st, None)
    assert rc == 0, "cwhip_pairwise_noalloc returned %d" % rc
    hip.check(hip.L.hipDeviceSynchronize(), "hipDeviceSynchronize")
    cnt = hip.download(d_cnt, np.int32, len(pairs))
    st = int(hip.download(d_st, np.int32, 1)[0])
    hip.free_all()
    return cnt, st


def direct(ctx, params, codes, offs, lens, pairs):
    ctx.load_sequences(codes, offs, lens)
    cnt, _ = ctx.pairwise_pairs(params, pairs, kernel=cwhip.KERNEL_BATCHED)
    return cnt


def first_input():
    # a seeded family plus the edge cases of the pairwise stage: empty
    # sequences, a gap-only sequence, '-' codes inside a sequence, one residue
    rng = np.random.default_rng(31)
    fam = seqgen.family(4242, 40, 700)
    seqs = list(fam.seqs)
    base = rng.choice(FULL, 300).astype(np.uint8)
    gapped = base.copy()
    gapped[rng.choice(300, 30, replace=False)] = 31
    seqs += [np.zeros(0, np.uint8), base, gapped, np.array([21], np.uint8), np.full(20, 31, np.uint8),
             rng.choice(FULL, 1300).astype(np.uint8)]
    return pack(seqs)


def second_input():
    fam = seqgen.family(99, 12, 150)
    return pack(fam.seqs)


def test_noalloc_matches_batched_entry_point_twice():
    hip = Hip()
    ctx = cwhip.Context(0)
    try:
        # first call: 46 sequences up to ~1300 residues, all 1,035 pairs
        codes, offs, lens = first_input()
        pairs = O.range_pairs(len(lens), 0, len(lens), 0, len(lens))
        params = cwhip.pair_params(max_len=int(lens.max()))
        ref = direct(ctx, params, codes, offs, lens, pairs)
        cnt_o, _ = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring(), nthreads=16)
        np.testing.assert_array_equal(ref, cnt_o)
        got, st = run_noalloc(hip, params, codes, offs, lens, pairs)
        assert st == 0
        np.testing.assert_array_equal(got, ref)
        # second call, smaller: 12 sequences of ~150 residues, 40 of their pairs
        codes2, offs2, lens2 = second_input()
        pairs2 = O.range_pairs(len(lens2), 0, len(lens2), 0, len(lens2))[:40]
        params2 = cwhip.pair_params(max_len=int(lens2.max()))
        ref2 = direct(ctx, params2, codes2, offs2, lens2, pairs2)
        cnt_o2, _ = O.pairwise_pairs(codes2, offs2, lens2, pairs2, O.Scoring(), nthreads=16)
        np.testing.assert_array_equal(ref2, cnt_o2)
        got2, st2 = run_noalloc(hip, params2, codes2, offs2, lens2, pairs2)
        assert st2 == 0
        np.testing.assert_array_equal(got2, ref2)
    finally:
        ctx.close()
        hip.free_all()


def test_noalloc_refuses_pairs_outside_what_it_was_given():
    # the kernels' own checks: a sequence index >= nseq, and a sequence longer
    # than max_len; the other pairs still equal the oracle
    hip = Hip()
    codes, offs, lens = second_input()
    n = len(lens)
    good = O.range_pairs(n, 0, n, 0, 6)
    pairs = np.concatenate([good, np.array([[0, n], [n + 5, 1]], np.int32)])
    params = cwhip.pair_params(max_len=int(lens.max()))
    got, st = run_noalloc(hip, params, codes, offs, lens, pairs)
    cnt_o, _ = O.pairwise_pairs(codes, offs, lens, good, O.Scoring(), nthreads=16)
    np.testing.assert_array_equal(got[:len(good)], cnt_o)
    assert (got[len(good):] == -2).all() and st == 4
    short = int(np.sort(lens)[n // 2])            # about half the sequences are longer
    got, st = run_noalloc(hip, params, codes, offs, lens, good, max_len=short)
    too_long = (lens[good[:, 0]] > short) | (lens[good[:, 1]] > short)
    assert too_long.any() and (~too_long).any()
    assert (got[too_long] == -2).all() and st == 4
    np.testing.assert_array_equal(got[~too_long], cnt_o[~too_long])


def test_noalloc_argument_limits():
    # rejected on the host before any device work
    L = cl.load()
    dp = cl.PairDevParams(None, None, 1, 30, 31, 100)
    one = C.c_void_p(16)          # never dereferenced: every call below is refused first
    assert L.cwhip_pairwise_noalloc(None, one, one, one, 1, one, 1, one, one, None) == -1
    assert L.cwhip_pairwise_noalloc(C.byref(dp), one, one, one, 1, one, 1, one, one, None) == -1
    dp = cl.PairDevParams(one, one, 1, 30, 31, 100)
    assert L.cwhip_pairwise_noalloc(C.byref(dp), one, one, one, 1, one, cl.NOALLOC_MAX_PAIRS + 1,
                                    one, one, None) == -6
    dp.max_len = cl.MAX_LEN + 1   # one past CWHIP_MAX_LEN (its value is checked against the header)
    assert L.cwhip_pairwise_noalloc(C.byref(dp), one, one, one, 1, one, 1, one, one, None) == -4
