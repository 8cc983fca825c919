"""GPU: pe_pull_sum_noalloc (no device allocation; scratch in a static per-GPU array)
against pe_pull_sum, through the C-ABI.  The merge-path pull sum is deterministic, so
the two agree bit for bit: two calls in a row, the second on a smaller graph, then a
graph at the array's limit (RMAT-27 with 1024-item tiles: FP32 fits, FP64 does not)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pull_both(ctx, g, dtype, weighted, seed):
    from plato_amd import DenseState
    from plato_amd._lib import check
    rng = np.random.default_rng(seed)
    x = DenseState(ctx, g.V, dtype)
    x.upload((rng.random(g.V) - 0.25).astype(dtype))  # raw sums take any sign
    y_ref = DenseState(ctx, g.V, dtype)
    y_na = DenseState(ctx, g.V, dtype)
    y_ref.fill(-7.0)
    y_na.fill(3.0)
    check(g.lib.pe_pull_sum(g.h, x.h, y_ref.h, int(weighted)), ctx.h)
    check(g.lib.pe_pull_sum_noalloc(g.h, x.h, y_na.h, int(weighted)), ctx.h)
    a, b = y_ref.numpy(), y_na.numpy()
    for v in (x, y_ref, y_na):
        v.close()
    return a, b


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_pull_sum_equals_pull_sum(ctx, dtype):
    from plato_amd import Graph
    for scale, weighted in ((16, False), (12, True)):  # the second call on a smaller graph
        print(f"case: pull_sum_noalloc rmat s{scale} weighted={weighted} {np.dtype(dtype).name}", flush=True)
        g = Graph.rmat(ctx, scale, 16, seed=scale, weighted=weighted, relabel=scale == 16)
        a, b = _pull_both(ctx, g, dtype, weighted, scale)
        assert np.array_equal(a, b)
        g.close()


def test_noalloc_pull_sum_size_limit(ctx):
    from plato_amd import Graph
    from plato_amd._lib import PlatoHipError
    print("case: pull_sum_noalloc rmat s27 tile 1024", flush=True)
    g = Graph.rmat(ctx, 27, 16, seed=1)
    g.set_tile(1024)
    a, b = _pull_both(ctx, g, np.float32, False, 27)  # ~36 MB of scratch: fits
    assert np.array_equal(a, b)
    with pytest.raises(PlatoHipError, match="static array"):  # ~71 MB in FP64: rejected
        _pull_both(ctx, g, np.float64, False, 28)
    g.close()
