"""acc_gpu_top_k_heap_noalloc (selection scratch in a static device array,
no allocation) against acc_gpu_top_k_heap, through the C ABI: two calls in a
row, the second with a smaller input, and a third that takes the other
selection path; the array keeps each call's state, so every call must start
from its own reset."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def scores(n, seed):
    rng = np.random.default_rng(seed)
    y = rng.integers(-1000, 1000, n).astype(np.float32)  # many ties: the lower row must win
    y[::7] = -0.0
    return y


@pytest.mark.parametrize("calls", [[(3_000_000, 100), (50_000, 100), (200_000, 700)],
                                   [(1_000_000, 960), (70_000, 960), (5_000, 7)]])
def test_noalloc_top_k_heap_matches_allocating_call(acc, gpu, calls):
    for n, k in calls:
        print(f"[noalloc] n={n} k={k}", flush=True)
        y = scores(n, n + k)
        yd = acc.DeviceBuffer.from_numpy(y)
        a = acc.DeviceBuffer((2 * k,), np.int32)
        b = acc.DeviceBuffer((2 * k,), np.int32)
        acc.top_k_heap(yd, n, k, a)
        acc.top_k_heap_noalloc(yd, n, k, b)
        acc.synchronize()
        np.testing.assert_array_equal(b.numpy(), a.numpy(), err_msg=f"n={n} k={k}")
        oi, ov = O.topk(y, k)
        got = b.numpy().view(np.int32).reshape(-1, 2)
        m = min(n, k)
        np.testing.assert_array_equal(got[:m, 1], oi[:m])
        np.testing.assert_array_equal(got[:m, 0].view(np.float32), ov[:m].astype(np.float32))


def test_noalloc_rejects_bad_sizes(acc, gpu):
    y = acc.DeviceBuffer((16,), np.float32)
    out = acc.DeviceBuffer((2 * 961,), np.int32)
    with pytest.raises(acc.AccelesError, match="k in"):
        acc.top_k_heap_noalloc(y, 16, 961, out)
    with pytest.raises(acc.AccelesError, match="k in"):
        acc.top_k_heap_noalloc(y, -1, 10, out)
