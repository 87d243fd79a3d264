"""dem_sort_particles_noalloc (the sort with the library's fixed per-GPU
scratch, for callers that cannot allocate) against dem_sort_particles (the
same sort with context-allocated scratch) and numpy's stable argsort, on two
calls in a row, the second one smaller: bit-exact keys and permutation."""
import numpy as np
import pytest

from demso_amd import _lib, api
from demso_amd import scene as sc
from helpers import dev, host

pytestmark = pytest.mark.gpu
NOALLOC_MAX = 4194304


def sort_both(ctx, keys):
    n = keys.shape[0]
    idx = np.arange(n, dtype=np.uint32)
    h1, i1 = dev(keys), dev(idx)
    h2, i2 = dev(keys), dev(idx)
    api.sortParticles_noalloc(ctx, h1, i1, n)
    api.sortParticles(ctx, h2, i2, n)
    ctx.synchronize()
    return (host(h1, np.uint32), host(i1, np.uint32)), (host(h2, np.uint32), host(i2, np.uint32))


def test_noalloc_sort_matches_allocating_sort():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    assert _lib.load().dem_sort_particles_noalloc is not None
    ctx = api.Context(0)
    try:
        rng = np.random.default_rng(5)
        # call 1: the maximum count, 32-bit keys (no parameters set: 4 passes,
        # the result ends in the caller's arrays) with heavy ties
        n1 = NOALLOC_MAX
        k1 = (rng.integers(0, 1 << 20, n1, dtype=np.uint64) * 4093).astype(np.uint32)
        (a_h, a_i), (b_h, b_i) = sort_both(ctx, k1)
        order = np.argsort(k1, kind="stable").astype(np.uint32)
        assert np.array_equal(a_h, b_h) and np.array_equal(a_i, b_i)
        assert np.array_equal(a_i, order) and np.array_equal(a_h, k1[order])
        # call 2: smaller, keys of a scene's cell range (18 bits: 3 passes, the
        # result copied back from the scratch), scratch holding call 1's data
        s = sc.settling(side=256)
        ctx.load_scene(s)
        n2 = 70001
        k2 = rng.integers(0, s.total_cells, n2, dtype=np.uint32)
        (a_h, a_i), (b_h, b_i) = sort_both(ctx, k2)
        order = np.argsort(k2, kind="stable").astype(np.uint32)
        assert np.array_equal(a_h, b_h) and np.array_equal(a_i, b_i)
        assert np.array_equal(a_i, order) and np.array_equal(a_h, k2[order])
        # more keys than the scratch holds: refused before any launch
        big = NOALLOC_MAX + 1
        h, i = dev(np.zeros(big, np.uint32)), dev(np.zeros(big, np.uint32))
        with pytest.raises(RuntimeError, match="at most"):
            api.sortParticles_noalloc(ctx, h, i, big)
        # the context still works after the refusal
        (a_h, a_i), _ = sort_both(ctx, k2[:1000])
        assert np.array_equal(a_i, np.argsort(k2[:1000], kind="stable").astype(np.uint32))
    finally:
        ctx.close()
