"""GPU: hasty_dot_noalloc (the no-allocation form of hasty_dot, for callers
that cannot allocate) through the C-ABI against hasty_dot: bit for bit on
two calls in a row, the second on a smaller input (the static partial-sum
scratch keeps the first call's contents), and against an FP64 numpy dot."""
import ctypes

import numpy as np
import pytest
import torch

from tests.gpu_util import crandn, cuda, host

pytestmark = pytest.mark.gpu


def _dot(fn, a, b):
    from hastycucompute_amd._lib import check
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    check(fn(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), int(a.numel()),
             ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return host(out)


def test_dot_noalloc_matches_dot():
    from hastycucompute_amd._lib import lib
    rng = np.random.default_rng(40)
    for n in (3_000_000, 1000, 0):          # large, then smaller: stale partials must not leak in
        a_np, b_np = crandn(rng, n), crandn(rng, n)
        a, b = cuda(a_np), cuda(b_np)
        ref = _dot(lib().hasty_dot, a, b)
        got = _dot(lib().hasty_dot_noalloc, a, b)
        print(f"[parity] dot_noalloc n={n}: {got} vs hasty_dot {ref}", flush=True)
        assert np.array_equal(got, ref), (n, got, ref)
        exact = np.vdot(a_np.astype(np.complex128), b_np.astype(np.complex128))
        scale = max(1.0, np.linalg.norm(a_np) * np.linalg.norm(b_np))
        assert abs(complex(got[0], got[1]) - exact) <= 1e-12 * np.sqrt(max(n, 1)) * scale


def test_dot_noalloc_rejects_bad_arguments():
    from hastycucompute_amd._lib import lib
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    st = lib().hasty_dot_noalloc(None, None, -1, ctypes.c_void_p(out.data_ptr()), None)
    assert st != 0 and b"invalid" in lib().hasty_last_error()
