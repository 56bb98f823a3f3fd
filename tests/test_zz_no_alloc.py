"""avism_indexx_noalloc, INDEXX on a static scratch array with no device
allocation, against avism_indexx on the same keys (which the oracle and the
reference's INDEXX pin): two calls in a row, the second smaller, with many
tied keys (NR's tie order decides cube and void ids), and its size limit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc():
    from avism_amd._lib import lib
    return lib().avism_indexx_noalloc  # the C-ABI symbol, signature bound in _lib


def _keys(rng, n):
    k = rng.integers(0, max(n // 50, 2), n).astype(np.float32)  # ties
    k[::3] = rng.standard_normal(k[::3].size).astype(np.float32)
    return k


def test_noalloc_indexx_matches_indexx():
    import torch
    from avism_amd._lib import Context, lib
    c = Context(0)
    f = _noalloc()
    rng = np.random.default_rng(3)
    for n in (400_000, 70_001, 9):
        keys = _keys(rng, n)
        want = c.indexx(keys)
        dk = torch.from_numpy(keys).cuda()
        di = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        print(f"[indexx_noalloc n={n}]", flush=True)
        rc = f(c._h, C.c_void_p(dk.data_ptr()), n, C.c_void_p(di.data_ptr()))
        assert rc == 0, lib().avism_last_error(c._h)
        assert np.array_equal(di.cpu().numpy(), want)
    # a second context on the same device shares the array
    c2 = Context(0)
    keys = _keys(rng, 1000)
    dk = torch.from_numpy(keys).cuda()
    di = torch.zeros(1000, dtype=torch.int32, device="cuda")
    assert f(c2._h, C.c_void_p(dk.data_ptr()), 1000, C.c_void_p(di.data_ptr())) == 0
    assert np.array_equal(di.cpu().numpy(), c.indexx(keys))


def test_noalloc_indexx_rejects_beyond_its_scratch():
    import torch
    from avism_amd._lib import Context, lib
    c = Context(0)
    f = _noalloc()
    n = (1 << 19) + 1
    dk = torch.zeros(n, device="cuda")
    di = torch.zeros(n, dtype=torch.int32, device="cuda")
    rc = f(c._h, C.c_void_p(dk.data_ptr()), n, C.c_void_p(di.data_ptr()))
    assert rc == -4  # AVISM_ERR_LIMIT
    assert b"524288" in lib().avism_last_error(c._h)
