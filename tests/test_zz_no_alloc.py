"""pt_reduce_sum_f64_noalloc (no workspace, no allocation) against
pt_reduce_sum_f64 (caller workspace), through the C ABI.

The two share their kernels and reduction order, so the sums agree bit for
bit; two calls in a row (the second smaller) check that the library's shared
partial-sum array needs no clearing between calls.
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _sum_with_workspace(lib, x):
    from pantea_amd import _lib

    n = int(x.numel())
    ws = torch.empty(max(1, lib.pt_reduce_workspace_bytes(n)), dtype=torch.uint8, device=x.device)
    out = torch.full((1,), float("nan"), dtype=torch.float64, device=x.device)
    _lib.check(lib.pt_reduce_sum_f64(n, x.data_ptr(), out.data_ptr(), ws.data_ptr(), ws.numel(),
                                     _lib.stream_ptr()), "pt_reduce_sum_f64")
    return out


def _sum_noalloc(lib, x):
    from pantea_amd import _lib

    out = torch.full((1,), float("nan"), dtype=torch.float64, device=x.device)
    _lib.check(lib.pt_reduce_sum_f64_noalloc(int(x.numel()), x.data_ptr(), out.data_ptr(), _lib.stream_ptr()),
               "pt_reduce_sum_f64_noalloc")
    return out


def test_noalloc_sum_equals_workspace_sum_on_consecutive_calls():
    from pantea_amd import _lib

    assert torch.cuda.is_available()
    lib = _lib.load()
    rng = np.random.default_rng(7)
    for n in (3_000_017, 98_304, 1_000, 1):   # each call smaller than the one before
        host = rng.normal(size=n) * 10.0 ** rng.integers(-3, 4, size=n)
        x = torch.as_tensor(host, dtype=torch.float64, device="cuda")
        a = _sum_with_workspace(lib, x)
        b = _sum_noalloc(lib, x)
        torch.cuda.synchronize()
        va, vb = float(a.item()), float(b.item())
        print(f"n={n}: workspace {va!r} no-alloc {vb!r}", flush=True)
        assert va == vb   # same kernels, same order: bitwise equal
        assert abs(vb - math.fsum(host)) <= 1e-13 * float(np.sum(np.abs(host)))


def test_noalloc_sum_of_nothing_is_zero():
    from pantea_amd import _lib

    lib = _lib.load()
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    _lib.check(lib.pt_reduce_sum_f64_noalloc(0, x.data_ptr(), out.data_ptr(), _lib.stream_ptr()), "noalloc")
    torch.cuda.synchronize()
    assert float(out.item()) == 0.0


def test_noalloc_rejects_sizes_beyond_its_maximum():
    """Rejected on the host before any launch: the pointers are never read."""
    from pantea_amd import _lib

    lib = _lib.load()
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = lib.pt_reduce_sum_f64_noalloc((1 << 26) + 1, x.data_ptr(), out.data_ptr(), _lib.stream_ptr())
    assert rc == _lib.PT_ERR_ARG
    assert "maximum" in _lib.last_error()
    assert lib.pt_reduce_sum_f64_noalloc(-1, x.data_ptr(), out.data_ptr(), _lib.stream_ptr()) == _lib.PT_ERR_ARG
    # the library still works after the rejected calls
    torch.cuda.synchronize()
    assert float(_sum_noalloc(lib, torch.ones(10, dtype=torch.float64, device="cuda")).item()) == 10.0
