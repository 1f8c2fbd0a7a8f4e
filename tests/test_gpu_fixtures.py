"""GPU distances reproduce the reference's own guide trees (all six reference
data sets, 100/1000 sequences, protein and DNA), and match the oracle."""
import numpy as np
import pytest

import cwhip
import oracle as O
from fixtures import load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cwhip.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("setname", ["100_97", "100_498", "1000_97", "1000_498", "100_1002", "100_1523"])
@pytest.mark.parametrize("kernel", [cwhip.KERNEL_PLAIN, cwhip.KERNEL_BATCHED])
def test_gpu_reproduces_reference_guide_tree(ctx, setname, kernel):
    names, codes, offs, lens, dna, dnd = load(setname)
    n = len(names)
    ctx.load_sequences(codes, offs, lens)
    params = cwhip.pair_params(dna=dna, max_len=int(lens.max()))
    pairs, cnt, dist = ctx.pairwise_range(params, 0, n, 0, n, kernel=kernel)
    assert O.nj_dnd(O.distance_matrix(n, pairs, dist), names) == dnd
    if n <= 100:
        co, do = O.pairwise_pairs(codes, offs, lens, pairs, O.Scoring(dna=dna), nthreads=16)
        np.testing.assert_array_equal(cnt, co)
        np.testing.assert_array_equal(dist.view(np.uint32), do.view(np.uint32))
