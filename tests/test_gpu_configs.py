"""The two workloads whose full size the other GPU tests do not reach, checked
against the oracle:

* the progressive stage on 1000 sequences of ~1002 residues (prog1002): the
  device's multiSeqAlign along the NJ guide tree of the device distances, row
  for row against the oracle's multiSeqAlign on the same guide, plus a sample
  of the distances;
* serving at the stated request sizes (N ~ U[10, 500], lengths lognormal(ln
  300, 0.7) clipped to [30, 2000]): five requests of the seeded stream, among
  them a 487-sequence distance job and alignment jobs of 38, 254 and 426
  sequences, submitted together; every served DistMatrix equals the oracle's
  distances, and every served .dnd / .aln equals the oracle's NJ tree and
  multiSeqAlign written out.
"""
import numpy as np
import pytest

import cwhip
import oracle as O
from cwhip import pipeline, seqgen

pytestmark = pytest.mark.gpu


def test_prog1002_progressive_stage_equals_oracle():
    fam = seqgen.workload("prog1002")
    codes, offs, lens = fam.packed()
    n = len(lens)
    print("prog1002: device distances", n, flush=True)
    ctx = cwhip.Context(0)
    try:
        ctx.load_sequences(codes, offs, lens)
        params = cwhip.pair_params(max_len=int(lens.max()))
        pairs, cnt, dist = ctx.pairwise_range(params, 0, n, 0, n)
        sample = np.arange(0, len(pairs), len(pairs) // 300)
        cnt_o, dist_o = O.pairwise_pairs(codes, offs, lens, pairs[sample], O.Scoring(), nthreads=16)
        np.testing.assert_array_equal(cnt[sample], cnt_o)
        np.testing.assert_array_equal(dist[sample].view(np.uint32), dist_o.view(np.uint32))
        D = pipeline.distance_matrix(n, pairs, dist)
        dnd = cwhip.nj_tree(D, fam.names)
        assert dnd == O.nj_dnd(D, fam.names)
        w, sim, steps = cwhip.guide_from_tree(dnd, fam.names)
        print("prog1002: device multiSeqAlign", flush=True)
        rows, order, stats = ctx.msa(fam.seqs, sim, w, steps)
        assert stats["calls"] == n - 1
    finally:
        ctx.close()
    print("prog1002: oracle multiSeqAlign", flush=True)
    r_rows, r_order, _ = O.msa([s.astype(np.int32) for s in fam.seqs], sim, w, steps, O.msa_params(False))
    np.testing.assert_array_equal(order, r_order)
    for i in range(n):
        np.testing.assert_array_equal(rows[i], r_rows[i].astype(np.uint8))


# (seed, index) of the serving stream: distances N=487 (max length 2000),
# distances N=57, alignments N=38, N=254 (max length 2000) and N=426
SERVING_REQUESTS = [(4, 5), (4, 1), (4, 2), (4, 6), (4, 14)]


@pytest.fixture(scope="module")
def served():
    reqs = [seqgen.serving_request(s, i) for s, i in SERVING_REQUESTS]
    kinds = [k for k, _ in reqs]
    assert kinds == [cwhip.JOB_DISTANCES, cwhip.JOB_DISTANCES, cwhip.JOB_ALIGNMENT,
                     cwhip.JOB_ALIGNMENT, cwhip.JOB_ALIGNMENT]
    srv = cwhip.Server(0, slots=4, hbm_budget=32 << 30)
    out = []
    try:
        print("serving: submit", [len(f.seqs) for _, f in reqs], flush=True)
        jobs = [srv.submit(kind, fam.seqs, fam.names) for kind, fam in reqs]
        for (kind, fam), j in zip(reqs, jobs):
            assert srv.wait(j, 300)
            d = srv.distances(j)
            txt = (srv.text(j, 0), srv.text(j, 1)) if kind == cwhip.JOB_ALIGNMENT else (None, None)
            out.append((kind, fam, d) + txt)
        assert srv.stats()["max_running"] >= 2
    finally:
        srv.close()
    return out


def _oracle_distances(fam):
    codes, offs, lens = fam.packed()
    n = len(lens)
    pairs = O.range_pairs(n, 0, n, 0, n)
    _, dist = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring(), nthreads=16)
    return pairs, dist


@pytest.mark.parametrize("k", [0, 1])
def test_served_distance_jobs_equal_oracle(served, k):
    kind, fam, d, _, _ = served[k]
    print("serving: oracle distances", len(fam.seqs), flush=True)
    _, dist = _oracle_distances(fam)
    np.testing.assert_array_equal(d.view(np.uint32), dist.view(np.uint32))


@pytest.mark.parametrize("k", [2, 3, 4])
def test_served_alignment_jobs_equal_oracle(served, k):
    kind, fam, d, dnd, aln = served[k]
    n = len(fam.seqs)
    print("serving: oracle alignment", n, flush=True)
    pairs, dist = _oracle_distances(fam)
    np.testing.assert_array_equal(d.view(np.uint32), dist.view(np.uint32))
    o_dnd = O.nj_dnd(O.distance_matrix(n, pairs, dist), fam.names)
    assert dnd == o_dnd
    w, sim, steps = O.guide_from_dnd(o_dnd, fam.names)
    rows, order, _ = O.msa([s.astype(np.int32) for s in fam.seqs], sim, w, steps, O.msa_params(False))
    assert aln == cwhip.clustal_text(fam.names, [r.astype(np.uint8) for r in rows], order)
