"""The CPU reference of the proposal network (oracle/cnn_ref.py) is pinned by
finite differences of its own loss: every parameter block's analytic gradient
must match a central difference, and a planted mistake (the ReLU mask dropped
from the backward) must fail the same check."""
from __future__ import annotations

import numpy as np
import pytest

import cnn_ref as R


def _loss(params, x, lab):
    lo, _ = R.forward(params, x, round_acts=False)
    m = lo.max(1, keepdims=True)
    e = np.exp(lo - m)
    pr = e / e.sum(1, keepdims=True)
    return float(-np.log(pr[np.arange(len(lab)), lab]).mean())


def _setup():
    rng = np.random.default_rng(0)
    p = {k: v.astype(np.float64) for k, v in R.init_params(1).items()}
    for k in ("b1", "b2", "b3", "bh"):
        p[k] = rng.standard_normal(p[k].shape) * 0.1
    x = rng.standard_normal((3, 3, 9, 9, 9))
    lab = np.array([1, 3, 7])
    return rng, p, x, lab


def _fd_check(grad_fn):
    rng, p, x, lab = _setup()
    _, g, _ = grad_fn(p, x, lab)
    worst = 0.0
    for name in R.ORDER:
        for _ in range(3):
            idx = tuple(int(rng.integers(0, s)) for s in p[name].shape)
            if name == "w1" and (idx[0] % R.C0P) >= R.C0:
                continue  # padding channel: zero input, zero gradient
            eps = 1e-6
            q = {k: v.copy() for k, v in p.items()}
            q[name][idx] += eps
            lp = _loss(q, x, lab)
            q[name][idx] -= 2 * eps
            lm = _loss(q, x, lab)
            fd = (lp - lm) / (2 * eps)
            worst = max(worst, abs(fd - g[name][idx]) / max(1e-3, abs(fd)))
    return worst


def test_gradients_match_finite_differences():
    assert _fd_check(lambda p, x, l: R.loss_and_grads(p, x, l, round_acts=False)) < 1e-4


def test_planted_mistake_is_caught(monkeypatch):
    # backward without the ReLU mask: (a > 0) replaced by 1
    src_forward = R.forward

    def no_mask_grads(p, x, lab):
        logits, cache = src_forward(p, x, round_acts=False)
        for li in range(3):
            cache[f"a{li}"] = np.abs(cache[f"a{li}"]) + 1.0  # every unit "active"
        monkeypatch.setattr(R, "forward", lambda *a, **k: (logits, cache))
        try:
            return R.loss_and_grads(p, x, lab, round_acts=False)
        finally:
            monkeypatch.setattr(R, "forward", src_forward)
    assert _fd_check(no_mask_grads) > 1e-2


def test_padding_channel_is_inert():
    _, p, x, lab = _setup()
    lo1, _ = R.forward(p, x)
    p2 = {k: v.copy() for k, v in p.items()}
    p2["w1"].reshape(27, R.C0P, R.C1)[:, R.C0:, :] = 5.0
    lo2, _ = R.forward(p2, x)
    assert np.array_equal(lo1, lo2)


def test_flat_layout_roundtrip():
    p = R.init_params(3)
    v = R.flatten(p)
    assert v.size == R.NPARAM == 19064
    q = R.unflatten(v)
    assert all(np.array_equal(p[k], q[k]) for k in R.ORDER)
