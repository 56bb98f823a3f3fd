"""INDEXX on the GPU reproduces the reference's permutation, ties included."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def _ctx():
    from avism_amd._lib import Context
    return Context(0)


def test_indexx_reference_vectors(golden_dir):
    z = np.load(golden_dir / "indexx.npz")
    c = _ctx()
    for k in z.files:
        if not k.startswith("keys_"):
            continue
        nm = k[5:]
        print(f"[indexx fixture {nm}]", flush=True)
        assert np.array_equal(c.indexx(z[k]), z[f"indx_{nm}"]), nm


@pytest.mark.parametrize("n,levels", [(1, 1), (2, 2), (7, 3), (8, 3), (9, 2), (1000, 5), (100_000, 50),
                                      (1_000_000, 1000), (1_000_000, 0)])
def test_indexx_vs_oracle(n, levels):
    rng = np.random.default_rng(n + levels)
    if levels == 0:
        keys = rng.standard_normal(n).astype(np.float32)
    else:
        keys = rng.integers(0, levels, size=n).astype(np.float32)
    print(f"[indexx n={n} levels={levels}]", flush=True)
    c = _ctx()
    assert np.array_equal(c.indexx(keys), O.indexx(keys))
