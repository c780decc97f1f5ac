"""sbp_mdl_noalloc (the allocation-free variant of sbp_mdl, for callers that cannot allocate device memory) against
sbp_mdl through the C-ABI: the same description length and log posterior, bit for bit, on two calls in a row, the
second on a smaller blockmodel (the fixed work area keeps the first call's partials), and after it the first
blockmodel again.  A nonparametric blockmodel is refused."""
import numpy as np
import pytest

from sbp_amd import native
from sbp_amd.graph import dcsbm

pytestmark = pytest.mark.gpu


def bm_of(ctx, g, B, seed):
    a = np.random.default_rng(seed).integers(0, B, g.N).astype(np.int32)
    a[:B] = np.arange(B)
    return native.DeviceBlockmodel.build(ctx.upload(g), B, a)


def test_mdl_noalloc_matches_mdl_on_consecutive_calls():
    ctx = native.Context(0, seed=1)
    big = bm_of(ctx, dcsbm(60000, 40, gamma=2.1, kmin=1, kmax=500, mean_degree=16, p_in=0.6, seed=2), 3000, 2)
    small = bm_of(ctx, dcsbm(500, 6, seed=3), 20, 3)
    print(f"[noalloc] big nnz={big.info()}, small nnz={small.info()}", flush=True)
    for bm in (big, small, big):
        want = bm.mdl()
        got = bm.mdl_noalloc()
        assert got == want, (got, want)
    ctx.close()


def test_mdl_noalloc_leaves_device_footprint_unchanged():
    """The context's device bytes (buffer pool, scratch slots) do not change across the call; the ordinary
    sbp_mdl grows its scratch on first use, which shows the measure would see an allocation."""
    ctx = native.Context(0, seed=1)
    bm = bm_of(ctx, dcsbm(20000, 20, gamma=2.1, kmin=1, kmax=300, mean_degree=16, p_in=0.6, seed=5), 800, 5)
    before = ctx.memory()
    got = bm.mdl_noalloc()
    assert ctx.memory() == before
    assert bm.mdl() == got
    assert ctx.memory()[0] > before[0]
    ctx.close()


def test_mdl_noalloc_refuses_nonparametric():
    ctx = native.Context(0, seed=1)
    g = dcsbm(500, 6, seed=3)
    a = np.random.default_rng(0).integers(0, 10, g.N).astype(np.int32)
    a[:10] = np.arange(10)
    bm = native.DeviceBlockmodel.build(ctx.upload(g), 10, a, nonparametric=True)
    with pytest.raises(native.SbpError, match="parametric"):
        bm.mdl_noalloc()
    ctx.close()
