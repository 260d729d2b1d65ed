"""The entry point that allocates no device memory
(sdb_sample_pair_dist_noalloc) against the allocating one
(sdb_sample_pair_dist), through the C ABI: equal outputs on two calls in a
row, the second smaller, and a clear refusal beyond the static buffer."""
import numpy as np
import pytest

import sdbscan_amd as S
from oracle import oracle as O
from sdbscan_amd import datagen

pytestmark = pytest.mark.gpu


def test_noalloc_pair_sample_matches_the_allocating_entry_point(ctx):
    X, _ = datagen.mixture(6000, 784, sigma=0.3, seed=311)
    p = ctx.set_params(6000, 784, n_proj=1024, topK=10, topM=100, precision=S.BF16)
    ctx.load(X)
    ctx.build_index(O.hd3_bits(3 * p.fhtDim, 42))
    print("no-alloc pair sample, cosine 6000 x 784 BF16", flush=True)
    for n_samples, seed in ((S.NOALLOC_MAX_SAMPLES, 7), (5000, 7), (777, 99)):
        want = ctx.sample_pair_dist(n_samples, seed=seed)
        got = ctx.sample_pair_dist_noalloc(n_samples, seed=seed)
        for w, g in zip(want, got):
            assert np.array_equal(w, g), (n_samples, seed)
        print("samples", n_samples, "seed", seed, "equal", flush=True)
    # the second call's pairs are the first ones of the first call (each
    # sample depends only on the seed and its own index)
    a_big = ctx.sample_pair_dist_noalloc(S.NOALLOC_MAX_SAMPLES, seed=7)
    a_small = ctx.sample_pair_dist_noalloc(5000, seed=7)
    for w, g in zip(a_big, a_small):
        assert np.array_equal(w[:5000], g)
    with pytest.raises(S.SdbError) as e:
        ctx.sample_pair_dist_noalloc(S.NOALLOC_MAX_SAMPLES + 1, seed=7)
    assert e.value.code == 1 and "static scratch" in str(e.value)
