"""Learned half-width proposals for the cube growth (no counterpart in the reference).

A 3-D CNN (conv 3->16->16->16, 3^3 valid + ReLU, linear head to 8 classes) sees a
9^3 patch of (delta, div/div_rms, |grad delta|/GRAD_THRE) around a candidate
centre and predicts the half-width the plain growth reaches (class 7 = 7 or
more).  Growth mode GROW_LEARNED runs the network on every candidate and then
verifies each proposal with the exact criteria (voidfinding.f90:500-654), so
the catalogue is the plain method's whatever the network predicts.

Training data are (patch, half-width) pairs taken from plain-method runs; the
forward, backward and Adam steps are the library's HIP kernels
(avism_cnn_forward / _backward / _adam), driven from here.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from ._lib import CNN_NCLASS, CNN_NPARAM, GROW_PLAIN, Context

WEIGHTS = Path(__file__).resolve().parent / "weights" / "halfwidth_cnn_v1.npy"

# parameter blocks in the flat vector: (name, rows, cols)
LAYOUT = [("w1", 108, 16), ("b1", 1, 16), ("w2", 432, 16), ("b2", 1, 16), ("w3", 432, 16), ("b3", 1, 16),
          ("wh", 432, CNN_NCLASS), ("bh", 1, CNN_NCLASS)]
assert sum(r * c for _, r, c in LAYOUT) == CNN_NPARAM


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (fan-in = 27 * real input channels), zero biases; the
    padded 4th input channel of conv1 has zero weights."""
    rng = np.random.default_rng(seed)
    out = []
    for name, r, c in LAYOUT:
        if name.startswith("b"):
            out.append(np.zeros(r * c, np.float32))
            continue
        fan_in = 27 * 3 if name == "w1" else r
        w = rng.standard_normal((r, c)) * np.sqrt(2.0 / fan_in)
        if name == "w1":
            w.reshape(27, 4, c)[:, 3, :] = 0.0
        out.append(w.astype(np.float32).reshape(-1))
    return np.concatenate(out)


def labels_from_halfwidth(r: np.ndarray) -> np.ndarray:
    return np.clip(r, 0, CNN_NCLASS - 1).astype(np.int32)


@dataclass
class Samples:
    patches: object  # torch int16 [n, 3, 9, 9, 9] (bf16 bits) on the GPU
    labels: object   # torch int32 [n]


def collect(ctx: Context, max_samples: int, seed: int = 0) -> Samples:
    """(patch, half-width) pairs of the current fields' candidates, from a
    plain-method growth run; candidates skipped at the edge zone are left out."""
    import torch
    ctx.find_voids(GROW_PLAIN)
    _, r = ctx.candidates()
    ranks = np.nonzero(r >= 0)[0]
    if ranks.size > max_samples:
        rng = np.random.default_rng(seed)
        ranks = np.sort(rng.choice(ranks, max_samples, replace=False))
    dev = torch.device("cuda", torch.cuda.current_device())
    tr = torch.from_numpy(ranks.astype(np.int64)).to(dev)
    patches = ctx.cnn_patches(tr)
    labels = torch.from_numpy(labels_from_halfwidth(r[ranks])).to(dev)
    return Samples(patches, labels)


@dataclass
class TrainLog:
    loss: list = field(default_factory=list)


def train(ctx: Context, data: Samples, params: np.ndarray | None = None, epochs: int = 4, batch: int = 512,
          lr: float = 2e-3, seed: int = 0, log: TrainLog | None = None) -> np.ndarray:
    """Adam on the mean softmax cross-entropy, all steps on the GPU kernels."""
    import torch
    dev = data.patches.device
    p = torch.from_numpy(init_params(seed) if params is None else params.astype(np.float32)).to(dev)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    n = int(data.labels.numel())
    g = torch.Generator(device="cpu").manual_seed(seed)
    step = 0
    for _ in range(epochs):
        perm = torch.randperm(n, generator=g).to(dev)
        for s in range(0, n - batch + 1 if n >= batch else 1, batch):
            idx = perm[s:s + batch]
            x = data.patches.index_select(0, idx).contiguous()
            y = data.labels.index_select(0, idx).contiguous()
            logits, cache = ctx.cnn_forward(p, x)
            grads, loss = ctx.cnn_backward(p, x, cache, logits, y)
            step += 1
            ctx.cnn_adam(p, grads, m, v, lr, step)
            if log is not None:
                log.loss.append(float(loss.item()))
    return p.cpu().numpy()


def predict(ctx: Context, params: np.ndarray, patches, chunk: int = 65536):
    import torch
    p = torch.from_numpy(np.ascontiguousarray(params, np.float32)).to(patches.device)
    out = []
    for s in range(0, int(patches.shape[0]), chunk):
        logits, _ = ctx.cnn_forward(p, patches[s:s + chunk].contiguous(), want_cache=False)
        out.append(logits.argmax(1))
    return torch.cat(out) if out else torch.zeros(0, dtype=torch.int64, device=patches.device)


class Proposer:
    """Holds trained parameters and installs them in a context for GROW_LEARNED."""

    def __init__(self, params: np.ndarray):
        self.params = np.ascontiguousarray(params, np.float32)
        if self.params.size != CNN_NPARAM:
            raise ValueError("wrong parameter count")

    @classmethod
    def load_default(cls, device=None):
        if not WEIGHTS.exists():
            return None
        return cls(np.load(WEIGHTS, allow_pickle=False))

    def install(self, ctx: Context):
        ctx.cnn_set_params(self.params)
