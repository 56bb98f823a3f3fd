"""FP64 CPU reference of the half-width proposal network.

Test infrastructure only: tests, the smoke check and the benchmark's CPU
baseline use it as the checker; the library never calls it.

The network has no counterpart in the reference code: it proposes the cube
half-width that VOIDFIND's growth loop (voidfinding.f90:500-654) would reach
from a candidate centre, and the exact criteria re-verify every proposal.  This
module is the checker for the HIP forward/backward kernels
(avism_amd/csrc/cnn.hip):

    input   3 x 9^3 patch (delta, div/div_rms, |grad delta|/GRAD_THRE), bf16
    conv1   3 -> 16, 3^3 valid, bias, ReLU   -> 16 x 7^3
    conv2  16 -> 16, 3^3 valid, bias, ReLU   -> 16 x 5^3
    conv3  16 -> 16, 3^3 valid, bias, ReLU   -> 16 x 3^3
    head   432 -> NCLS logits (classes r = 0..NCLS-1, the last one meaning >=)

Weights are stored as W[k][n] with k = off*CIN + c: the 3^3 offset
off = (dz*3+dy)*3+dx first, the input channel c fastest (CIN = 4 for conv1,
whose 4th input channel is zero padding; 16 otherwise) and n the output
channel; the head's k = pos*16 + c over the 3^3 map (x fastest).
Activations between layers are rounded to bf16 exactly as the GPU stores them,
so the remaining difference is FP32 vs FP64 accumulation.
"""
from __future__ import annotations

import numpy as np

NCLS = 8
P = 9
C0, C1 = 3, 16
C0P = 4  # conv1 input channels padded to 4
SHAPES = {"w1": (C0P * 27, C1), "b1": (C1,), "w2": (C1 * 27, C1), "b2": (C1,), "w3": (C1 * 27, C1),
          "b3": (C1,), "wh": (C1 * 27, NCLS), "bh": (NCLS,)}
ORDER = ["w1", "b1", "w2", "b2", "w3", "b3", "wh", "bh"]
NPARAM = sum(int(np.prod(SHAPES[k])) for k in ORDER)


def flatten(params) -> np.ndarray:
    return np.concatenate([np.asarray(params[k], np.float32).reshape(-1) for k in ORDER])


def unflatten(vec: np.ndarray):
    out, o = {}, 0
    for k in ORDER:
        n = int(np.prod(SHAPES[k]))
        out[k] = np.asarray(vec[o:o + n], np.float32).reshape(SHAPES[k])
        o += n
    return out


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (round to nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed: int = 0):
    rng = np.random.default_rng(seed)
    p = {}
    for k, shp in SHAPES.items():
        if k.startswith("w"):
            fan_in = shp[0] if k != "w1" else C0 * 27
            p[k] = (rng.standard_normal(shp) * np.sqrt(2.0 / fan_in)).astype(np.float32)
            if k == "w1":
                p[k].reshape(27, C0P, C1)[:, C0:, :] = 0.0
        else:
            p[k] = np.zeros(shp, np.float32)
    return p


def _im2col(x: np.ndarray) -> np.ndarray:
    """x: [B, C, S, S, S] (index order c, z, y, x) -> [B, S-2, S-2, S-2, 27*C] (offset major)."""
    B, C, S = x.shape[0], x.shape[1], x.shape[2]
    O = S - 2
    cols = np.empty((B, O, O, O, 27, C), np.float64)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                cols[..., (dz * 3 + dy) * 3 + dx, :] = np.moveaxis(x[:, :, dz:dz + O, dy:dy + O, dx:dx + O], 1, -1)
    return cols.reshape(B, O, O, O, 27 * C)


def _col2im(dcols: np.ndarray, C: int, S: int) -> np.ndarray:
    B, O = dcols.shape[0], S - 2
    d = dcols.reshape(B, O, O, O, 27, C)
    dx_ = np.zeros((B, C, S, S, S), np.float64)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                dx_[:, :, dz:dz + O, dy:dy + O, dx:dx + O] += np.moveaxis(d[..., (dz * 3 + dy) * 3 + dx, :], -1, 1)
    return dx_


def forward(params, x0, round_acts=True, acts=None):
    """x0: [B, 3, 9, 9, 9] (c, z, y, x).  Returns logits and a cache.

    round_acts=True rounds inputs, weights and activations to bf16 as the GPU
    does; False is the plain FP64 network (for finite-difference checks).
    acts: optional given activations of the three layers ([B, 16, O, O, O]),
    used instead of recomputed ones (a backward check given the same forward)."""
    rnd = bf16 if round_acts else (lambda a: np.asarray(a, np.float64))
    W = {k: rnd(v) if k.startswith("w") else np.asarray(v, np.float64) for k, v in params.items()}
    x = rnd(x0)
    x = np.concatenate([x, np.zeros((x.shape[0], C0P - x.shape[1]) + x.shape[2:])], axis=1)
    cache = {"x0": x}
    for li, (w, b) in enumerate((("w1", "b1"), ("w2", "b2"), ("w3", "b3"))):
        cols = _im2col(x)
        z = cols @ W[w] + W[b]
        a = np.maximum(z, 0.0)
        a = np.moveaxis(a, -1, 1)  # [B, C1, O, O, O]
        a = rnd(a) if acts is None else np.asarray(acts[li], np.float64)
        cache[f"cols{li}"] = cols
        cache[f"a{li}"] = a
        x = a
    feat = np.moveaxis(x, 1, -1).reshape(x.shape[0], 27 * C1)  # position major, channel fastest
    logits = feat @ W["wh"] + W["bh"]
    cache["feat"] = feat
    cache["W"] = W
    return logits, cache


def loss_and_grads(params, x0, labels, round_acts=True, acts=None):
    """Mean softmax cross-entropy and FP64 gradients w.r.t. every parameter."""
    logits, cache = forward(params, x0, round_acts, acts)
    B = logits.shape[0]
    m = logits.max(1, keepdims=True)
    e = np.exp(logits - m)
    p = e / e.sum(1, keepdims=True)
    loss = float(-np.log(p[np.arange(B), labels]).mean())
    dlog = p.copy()
    dlog[np.arange(B), labels] -= 1.0
    dlog /= B
    W = cache["W"]
    g = {"wh": cache["feat"].T @ dlog, "bh": dlog.sum(0)}
    dfeat = dlog @ W["wh"].T
    da = np.moveaxis(dfeat.reshape(B, 3, 3, 3, C1), -1, 1)
    sizes = [9, 7, 5]
    chans = [C0P, C1, C1]
    for li in (2, 1, 0):
        a = cache[f"a{li}"]
        dz = da * (a > 0)
        dzc = np.moveaxis(dz, 1, -1)  # [B, O, O, O, C1]
        cols = cache[f"cols{li}"]
        wname, bname = f"w{li + 1}", f"b{li + 1}"
        g[wname] = cols.reshape(-1, cols.shape[-1]).T @ dzc.reshape(-1, C1)
        g[bname] = dzc.reshape(-1, C1).sum(0)
        if li > 0:
            dcols = dzc @ W[wname].T
            da = _col2im(dcols, chans[li], sizes[li])
    return loss, g, logits


def predict(params, x0):
    logits, _ = forward(params, x0)
    return logits.argmax(1)
