"""The batched pairwise path: length-binned device work queue, hipGraph replay
of the batched stages, per-class shapes and the autotuner.

None of these may change a result: every run is compared bit for bit with the
plain kernel and, on a sample of pairs, with the oracle.
"""
import numpy as np
import pytest

import cwhip
from cwhip import seqgen
import oracle as O

pytestmark = pytest.mark.gpu

SC = O.Scoring()


def uneven_subset(nseq, seed=3):
    """nseq sequences of the 2000-sequence lognormal workload (lengths 50..3000)."""
    fam = seqgen.workload("uneven")
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(len(fam.seqs), nseq, replace=False))
    seqs = [fam.seqs[i] for i in pick]
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(len(seqs), np.int64)
    offs[1:] = np.cumsum(lens[:-1])
    return np.concatenate(seqs).astype(np.uint8), offs, lens


@pytest.fixture(scope="module")
def job():
    codes, offs, lens = uneven_subset(160)
    pairs = O.range_pairs(len(lens), 0, len(lens), 0, len(lens))
    params = cwhip.pair_params(max_len=int(lens.max()))
    # the oracle on every 11th pair (about 1,150 pairs of lengths 50..3000)
    sample = np.arange(0, len(pairs), 11)
    cnt_o, dist_o = O.pairwise_pairs(codes, offs, lens, pairs[sample], SC, nthreads=16)
    return codes, offs, lens, pairs, params, sample, cnt_o, dist_o


@pytest.fixture(scope="module")
def plain(job):
    codes, offs, lens, pairs, params, sample, cnt_o, dist_o = job
    ctx = cwhip.Context(0)
    ctx.load_sequences(codes, offs, lens)
    cnt, dist = ctx.pairwise_pairs(params, pairs, kernel=cwhip.KERNEL_PLAIN)
    ctx.close()
    np.testing.assert_array_equal(cnt[sample], cnt_o)
    np.testing.assert_array_equal(dist[sample].view(np.uint32), dist_o.view(np.uint32))
    return cnt, dist


def fresh_ctx(codes, offs, lens, params):
    """A context whose result buffer first holds a different job's counts, so a
    pair the queue dropped would show up as a stale value."""
    ctx = cwhip.Context(0)
    ctx.load_sequences(codes, offs, lens)
    n = len(lens)
    decoy = np.array([(i, (i + 1) % n) for i in range(n)], np.int32)
    ctx.pairwise_pairs(params, np.concatenate([decoy] * 60), kernel=cwhip.KERNEL_PLAIN)
    return ctx


def test_binned_queue_matches_plain_and_oracle(job, plain):
    codes, offs, lens, pairs, params, sample, cnt_o, dist_o = job
    ctx = fresh_ctx(codes, offs, lens, params)
    try:
        cnt, dist = ctx.pairwise_pairs(params, pairs, kernel=cwhip.KERNEL_BATCHED)
        np.testing.assert_array_equal(cnt, plain[0])
        np.testing.assert_array_equal(dist.view(np.uint32), plain[1].view(np.uint32))
        np.testing.assert_array_equal(cnt[sample], cnt_o)
    finally:
        ctx.close()


@pytest.mark.parametrize("nlong", [0, 1, 3])
def test_long_and_short_parts(nlong):
    """Pairs with a sequence longer than the length at which the batched kernel
    still runs its most waves per CU are queued first and run in a launch of
    their own; the rest run after them with smaller sequence buffers.  Sets with
    no, one and three 2,400-2,900-residue sequences among short ones: every
    count equals the plain kernel's and a sample the oracle's."""
    rng = np.random.default_rng(40 + nlong)
    fam = seqgen.workload("uneven")
    lens_all = np.array([len(s) for s in fam.seqs])
    long_ix = np.flatnonzero((lens_all > 2400) & (lens_all < 2900))[:nlong]
    short_ix = rng.choice(np.flatnonzero(lens_all < 700), 60, replace=False)
    seqs = [fam.seqs[i] for i in np.concatenate([long_ix, short_ix])]
    assert len(seqs) == 60 + nlong
    lens = np.array([len(x) for x in seqs], np.int32)
    offs = np.zeros(len(seqs), np.int64)
    offs[1:] = np.cumsum(lens[:-1])
    codes = np.concatenate(seqs).astype(np.uint8)
    pairs = O.range_pairs(len(lens), 0, len(lens), 0, len(lens))
    params = cwhip.pair_params(max_len=int(lens.max()))
    ctx = fresh_ctx(codes, offs, lens, params)
    try:
        cnt_p, dist_p = ctx.pairwise_pairs(params, pairs, kernel=cwhip.KERNEL_PLAIN)
        for _ in range(2):                       # eager, then captured
            cnt, dist = ctx.pairwise_pairs(params, pairs, kernel=cwhip.KERNEL_BATCHED)
            np.testing.assert_array_equal(cnt, cnt_p)
            np.testing.assert_array_equal(dist.view(np.uint32), dist_p.view(np.uint32))
    finally:
        ctx.close()
    sample = np.arange(0, len(pairs), 7)
    cnt_o, _ = O.pairwise_pairs(codes, offs, lens, pairs[sample], SC, nthreads=16)
    np.testing.assert_array_equal(cnt_p[sample], cnt_o)


def test_graph_replays_match_eager(job, plain):
    # the captured stages are replayed several times with other device work and
    # host synchronisation between the replays; every replay must give the
    # same counts as the plain kernel and as the stages launched one by one
    codes, offs, lens, pairs, params, sample, cnt_o, dist_o = job
    ctx = fresh_ctx(codes, offs, lens, params)
    try:
        ctx.stage(params, pairs)
        ctx.set_graph(False)
        ctx.run_staged(kernel=cwhip.KERNEL_BATCHED)
        np.testing.assert_array_equal(ctx.fetch_counts(), plain[0])
        ctx.set_graph(True)
        for rep in range(4):
            ctx.run_staged(kernel=cwhip.KERNEL_BATCHED)
            np.testing.assert_array_equal(ctx.fetch_counts(), plain[0], err_msg="replay %d" % rep)
            ctx.run_staged(kernel=cwhip.KERNEL_PLAIN)          # eager work between replays
        ms, kms = ctx.run_staged(kernel=cwhip.KERNEL_BATCHED, reps=3)
        assert ms > 0
        np.testing.assert_array_equal(ctx.fetch_counts(), plain[0])
        # a different staged job runs directly, then is captured as a new graph;
        # the old graph must not run for it
        sub = pairs[::5]
        ctx.stage(params, sub)
        for rep in range(3):
            ctx.run_staged(kernel=cwhip.KERNEL_BATCHED)
            np.testing.assert_array_equal(ctx.fetch_counts(), plain[0][::5], err_msg="sub %d" % rep)
    finally:
        ctx.close()


def test_graph_capture_beside_other_contexts(job, plain):
    """Contexts used from several host threads at once: while one thread
    captures its batched stages, the others launch, read back counts and
    diagnostics and synchronise; no call may fail and every result must equal
    the plain kernel's."""
    import threading
    codes, offs, lens, pairs, params, sample, cnt_o, dist_o = job
    subs = [np.arange(k, len(pairs), 6) for k in range(6)]
    errors = []

    def work(k):
        try:
            ctx = cwhip.Context(0)
            try:
                ctx.load_sequences(codes, offs, lens)
                ctx.stage(params, pairs[subs[k]])
                for rep in range(4):
                    ctx.run_staged(kernel=cwhip.KERNEL_BATCHED)
                    np.testing.assert_array_equal(ctx.fetch_counts(), plain[0][subs[k]],
                                                  err_msg="thread %d rep %d" % (k, rep))
                    ctx.fetch_info()
            finally:
                ctx.close()
        except Exception as e:                      # reported by the main thread
            errors.append("thread %d: %r" % (k, e))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(subs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(240)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors


@pytest.mark.parametrize("rmax,colcap", [(1, 32), (2, 64), (3, 256), (5, 128), (8, 32), (8, 256)])
def test_tuning_tables_give_identical_results(job, plain, rmax, colcap):
    # colcap < 256 sends Myers-Miller nodes of 33..256 columns to the row path,
    # small rmax splits long sweeps into many stripes
    codes, offs, lens, pairs, params, sample, cnt_o, dist_o = job
    ctx = cwhip.Context(0)
    try:
        ctx.load_sequences(codes, offs, lens)
        ctx.set_tune_table([rmax, 8, rmax, 2, rmax, 4], [colcap, 256, 64, colcap, colcap, 128])
        t = ctx.tune_table()
        assert list(t.rmax) == [rmax, 8, rmax, 2, rmax, 4]
        ctx.stage(params, pairs)
        ctx.run_staged(kernel=cwhip.KERNEL_BATCHED)
        np.testing.assert_array_equal(ctx.fetch_counts(), plain[0])
        info = ctx.fetch_info()
        assert (info[:, 5] == 0).all()
    finally:
        ctx.close()


def test_bad_tuning_tables_are_refused():
    ctx = cwhip.Context(0)
    try:
        for rm, cc in [(0, 64), (9, 64), (4, 16), (4, 512)]:
            with pytest.raises(cwhip.CwhipError):
                ctx.set_tune_table(rm, cc)
        before = ctx.tune_table().as_dict()
        with pytest.raises(cwhip.CwhipError):
            ctx.set_tune_table([4] * 5 + [9], 64)
        assert ctx.tune_table().as_dict() == before
    finally:
        ctx.close()


def test_autotune_keeps_results_and_restages(job, plain):
    codes, offs, lens, pairs, params, sample, cnt_o, dist_o = job
    ctx = cwhip.Context(0)
    try:
        ctx.load_sequences(codes, offs, lens)
        ctx.stage(params, pairs)
        table, ms = ctx.autotune(params, pairs)
        assert ms > 0
        assert all(1 <= r <= 8 for r in table.rmax)
        assert all(32 <= c <= 256 for c in table.colcap)
        assert ctx.tune_table().as_dict() == table.as_dict()
        # the job staged before the call is staged again
        ctx.run_staged(kernel=cwhip.KERNEL_BATCHED)
        np.testing.assert_array_equal(ctx.fetch_counts(), plain[0])
        # contexts created afterwards on the same device start from the tuned table
        ctx2 = cwhip.Context(0)
        try:
            assert ctx2.tune_table().as_dict() == table.as_dict()
        finally:
            ctx2.close()
    finally:
        ctx.close()
