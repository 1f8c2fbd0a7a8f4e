"""The serving loop: jobs of different sizes run concurrently (continuous
admission into device slots) and every result is bit-identical to the same
job run alone, and to the plain one-call path."""
import numpy as np
import pytest

import cwhip
from cwhip import pipeline, seqgen

pytestmark = pytest.mark.gpu


def _requests(count, seed=3):
    out = []
    for i in range(count):
        kind, fam = seqgen.serving_request(seed, i, n_range=(10, 90), len_cap=600)
        out.append((kind, fam))
    return out


def _collect(srv, job, kind):
    d = srv.distances(job)
    if kind == cwhip.JOB_ALIGNMENT:
        return d, srv.text(job, 0), srv.text(job, 1)
    return d, None, None


def test_serving_results_same_alone_and_in_the_mix():
    reqs = _requests(10)
    assert any(k == cwhip.JOB_ALIGNMENT for k, _ in reqs)
    srv = cwhip.Server(0, slots=4, hbm_budget=16 << 30)
    try:
        alone = []
        for kind, fam in reqs:
            print("alone", kind, len(fam.seqs), flush=True)
            j = srv.submit(kind, fam.seqs, fam.names)
            assert srv.wait(j, 300)
            alone.append(_collect(srv, j, kind))
            srv.release(j)
        print("mix", flush=True)
        jobs = [srv.submit(kind, fam.seqs, fam.names) for kind, fam in reqs]
        for (kind, fam), j, ref in zip(reqs, jobs, alone):
            assert srv.wait(j, 300)
            d, dnd, aln = _collect(srv, j, kind)
            np.testing.assert_array_equal(d.view(np.uint32), ref[0].view(np.uint32))
            assert dnd == ref[1] and aln == ref[2]
            t = srv.times(j)
            assert t[0] <= t[1] <= t[2]
        st = srv.stats()
        assert st["max_running"] >= 2          # jobs overlapped
        assert st["peak_bytes"] <= st["budget"]
    finally:
        srv.close()
    # the served results equal the direct path (plain kernel, pipeline)
    ctx = cwhip.Context(0)
    for (kind, fam), ref in zip(reqs, alone):
        codes, offs, lens = fam.packed()
        ctx.load_sequences(codes, offs, lens)
        params = cwhip.pair_params(max_len=int(lens.max()))
        pairs, cnt, dist = ctx.pairwise_range(params, 0, len(lens), 0, len(lens),
                                              kernel=cwhip.KERNEL_PLAIN)
        np.testing.assert_array_equal(dist.view(np.uint32), ref[0].view(np.uint32))
        if kind == cwhip.JOB_ALIGNMENT:
            dnd, aln, rows, order = pipeline.align(ctx, fam.names, fam.seqs, False)
            assert dnd == ref[1] and aln == ref[2]
    ctx.close()


def test_serving_admission_by_footprint():
    """A budget that holds one job at a time serialises the jobs; a job larger
    than the whole budget is refused at submission."""
    reqs = _requests(4, seed=9)
    srv = cwhip.Server(0, slots=4, hbm_budget=3 << 30)
    try:
        jobs = [srv.submit(kind, fam.seqs, fam.names) for kind, fam in reqs]
        for j in jobs:
            assert srv.wait(j, 300)
        assert srv.stats()["peak_bytes"] <= 3 << 30
    finally:
        srv.close()
    tiny = cwhip.Server(0, slots=2, hbm_budget=1 << 20)
    try:
        with pytest.raises(cwhip.CwhipError):
            tiny.submit(0, reqs[0][1].seqs, reqs[0][1].names)
    finally:
        tiny.close()


def test_serving_all_slots_busy():
    """The largest server (16 slots) with more jobs submitted at once than it has
    slots: every job completes and equals the same job run on the direct path."""
    reqs = []
    for i in range(40):
        kind, fam = seqgen.serving_request(17, i, n_range=(10, 60), len_cap=400)
        reqs.append((kind, fam))
    srv = cwhip.Server(0, slots=16, hbm_budget=64 << 30)
    try:
        jobs = [srv.submit(kind, fam.seqs, fam.names) for kind, fam in reqs]
        got = []
        for (kind, fam), j in zip(reqs, jobs):
            assert srv.wait(j, 300)
            got.append(_collect(srv, j, kind))
        assert srv.stats()["max_running"] >= 4
    finally:
        srv.close()
    ctx = cwhip.Context(0)
    try:
        for (kind, fam), ref in zip(reqs, got):
            codes, offs, lens = fam.packed()
            ctx.load_sequences(codes, offs, lens)
            params = cwhip.pair_params(max_len=int(lens.max()))
            pairs, cnt, dist = ctx.pairwise_range(params, 0, len(lens), 0, len(lens),
                                                  kernel=cwhip.KERNEL_BATCHED)
            np.testing.assert_array_equal(dist.view(np.uint32), ref[0].view(np.uint32))
            if kind == cwhip.JOB_ALIGNMENT:
                dnd, aln, rows, order = pipeline.align(ctx, fam.names, fam.seqs, False)
                assert dnd == ref[1] and aln == ref[2]
    finally:
        ctx.close()


def test_serving_real_device_bytes_within_budget_and_in_order():
    """A budget of 512 MiB, narrower than what four slots keep from earlier jobs:
    the device bytes the slots' contexts actually hold never exceed it (idle
    slots' buffers are released when the next job needs the room, and a running
    job cannot grow past its reservation); jobs start in arrival order, so none
    is overtaken; every result equals the same job on the direct path."""
    reqs = _requests(12, seed=21)
    srv = cwhip.Server(0, slots=4, hbm_budget=512 << 20)
    try:
        jobs = [srv.submit(kind, fam.seqs, fam.names) for kind, fam in reqs]
        got = []
        for (kind, fam), j in zip(reqs, jobs):
            assert srv.wait(j, 300)
            got.append(_collect(srv, j, kind))
        starts = [srv.times(j)[1] for j in jobs]
        assert all(a <= b for a, b in zip(starts, starts[1:]))
        st = srv.stats()
        assert 0 < st["peak_bytes"] <= 512 << 20
    finally:
        srv.close()
    ctx = cwhip.Context(0)
    try:
        for (kind, fam), ref in zip(reqs, got):
            codes, offs, lens = fam.packed()
            ctx.load_sequences(codes, offs, lens)
            params = cwhip.pair_params(max_len=int(lens.max()))
            pairs, cnt, dist = ctx.pairwise_range(params, 0, len(lens), 0, len(lens))
            np.testing.assert_array_equal(dist.view(np.uint32), ref[0].view(np.uint32))
            if kind == cwhip.JOB_ALIGNMENT:
                dnd, aln, rows, order = pipeline.align(ctx, fam.names, fam.seqs, False)
                assert dnd == ref[1] and aln == ref[2]
    finally:
        ctx.close()
