"""Train the half-width proposal network on plain-method runs.

    python -m avism_amd.train_proposer --n 256 --box 500 --seeds 40-47 \
        --out avism_amd/weights/halfwidth_cnn_v1.npy

For every seed: Zel'dovich particles (avism_amd.synth) -> fields -> plain
growth; (patch, half-width) pairs of a sample of the candidates are kept on
the GPU.  The network is then trained with the library's HIP kernels (bf16
MFMA forward and backward, fp32 Adam).  Prints one progress line per seed and
per epoch, and the held-out accuracy on the last seed.
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np


def parse_seeds(s: str):
    if "-" in s:
        a, b = s.split("-")
        return list(range(int(a), int(b) + 1))
    return [int(x) for x in s.split(",")]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--box", type=float, default=500.0)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--seeds", default="40-47")
    ap.add_argument("--per-seed", type=int, default=100_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    import torch
    from . import learned, synth
    from ._lib import Context

    dev = torch.device("cuda:0")
    seeds = parse_seeds(args.seeds)
    ctx = Context(0)
    ctx.set_grid(args.n, args.box, 1)
    ctx.set_cosmology(0.678, 0.31, 0.0)
    ctx.set_void_params(-0.6, 10.0, 0.25, 0.0, 1.0, 1.0, 30_000_000)
    chunks = []
    for s in seeds:
        t0 = time.perf_counter()
        p = synth.zeldovich(args.n, args.box, seed=s, device=dev)
        ctx.load_particles(p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m)
        del p
        ctx.particles_to_grid(k=args.k)
        d = learned.collect(ctx, args.per_seed, seed=s)
        hist = np.bincount(d.labels.cpu().numpy(), minlength=8)
        print(f"seed {s}: {int(d.labels.numel())} samples, classes {hist.tolist()}, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        chunks.append(d)
    held = chunks[-1]
    train = learned.Samples(torch.cat([c.patches for c in chunks[:-1]]), torch.cat([c.labels for c in chunks[:-1]]))
    params = None
    for ep in range(args.epochs):
        t0 = time.perf_counter()
        log = learned.TrainLog()
        params = learned.train(ctx, train, params=params, epochs=1, batch=args.batch,
                               lr=args.lr * (0.5 ** ep), seed=ep, log=log)
        pred = learned.predict(ctx, params, held.patches).cpu().numpy()
        acc = float(np.mean(pred == held.labels.cpu().numpy()))
        print(f"epoch {ep}: loss {np.mean(log.loss[-20:]):.4f}, held-out accuracy {acc:.4f}, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
    y = held.labels.cpu().numpy()
    pred = learned.predict(ctx, params, held.patches).cpu().numpy()
    report = {"seeds": seeds, "n": args.n, "box": args.box, "train_samples": int(train.labels.numel()),
              "held_out_seed": seeds[-1], "held_out_accuracy": float(np.mean(pred == y)),
              "best_constant_accuracy": float(np.max(np.bincount(y, minlength=8)) / max(y.size, 1)),
              "within_one_class": float(np.mean(np.abs(pred - y) <= 1))}
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    np.save(out, params.astype(np.float32), allow_pickle=False)
    out.with_suffix(".json").write_text(json.dumps(report, indent=1) + "\n")
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
