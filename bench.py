#!/usr/bin/env python3
"""Benchmark of the AVISM grid-and-void path on one MI355X.

Workload: 512^3 Zel'dovich particles (P(k) ~ k^-1.5 exp(-k^2/0.25), rms
displacement 3 Mpc/h, f = 0.5) in a periodic 1000 Mpc/h box, SPH with k = 32
onto a 512^3 grid (extended to 640^3 by the periodic shell), thresholds of
config/voids.dat.  One step = particles resident in HBM -> fields -> marking,
growth, acceptance -> merge -> post-processing -> catalogue and MARCA in HBM.

Both growth modes run on the same particles: the plain method and the
learned-proposal method (3D CNN half-width proposals, exactly re-verified);
their catalogues must be identical.  `value` is the learned-proposal path's
cells/s (N^3 inner cells per second of wall time); the plain path's figure and
the ratio are reported beside it.

With --levels L (default 1) a step runs L void-in-void levels on the same
particles (N^3, then (2N)^3, ... cells; avism_next_level between them) and
the line carries a per-level record; `value` is then N^3 base cells per
second of the whole multi-level step.

Prints one JSON line (see the keys below).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512, help="particles and cells per axis")
    ap.add_argument("--box", type=float, default=1000.0)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed-mass", action="store_true", help="skip the mixed-mass field-stage timing")
    ap.add_argument("--cpu-n", type=int, default=64, help="side of the CPU-baseline sample")
    ap.add_argument("--mode", choices=["both", "plain", "learned"], default="both")
    ap.add_argument("--levels", type=int, default=1, help="void-in-void levels per step (N, 2N, ...)")
    return ap.parse_args()


def stage_table(tim):
    out = {}
    for name, ms in tim:
        out[name] = out.get(name, 0.0) + ms
    return out


def run_step(ctx, parts, k, growth_mode, proposer=None, levels=1, grid=None, per_level=None):
    from avism_amd._lib import AVISM_DEVICE
    if levels > 1:
        ctx.set_grid(*grid)  # back to the base level
    ctx.load_particles(parts.x, parts.y, parts.z, parts.vx, parts.vy, parts.vz, parts.m, where=AVISM_DEVICE)
    for lev in range(levels):
        t0 = time.perf_counter()
        if lev:
            ctx.next_level()
        ctx.particles_to_grid(k=k)
        nv = ctx.find_voids(growth_mode)
        ctx.merge_and_post()
        if per_level is not None:  # the library calls return with their work complete
            st = ctx.stats()
            rec = per_level.setdefault(lev, dict(n=ctx.grid().n, ms=0.0, stages={}))
            rec["ms"] += (time.perf_counter() - t0) * 1e3
            rec.update(ncand=int(st.ncand), ncubes=int(nv), nvoids=int(st.nvoidt))
            for name, ms in ctx.timings():
                rec["stages"][name] = rec["stages"].get(name, 0.0) + ms
    return nv


def cpu_baseline(n_side: int, box_per_cell: float, k: int):
    """The reference's own routines (oracle/_ref: the reference's Fortran
    sources built with amdflang -O2 -fopenmp, driven by the oracle harness) on
    a bounded sample of the same workload: n_side^3 particles and cells, same
    cell size, all host cores.  Falls back to the oracle (C port, 1 thread)
    when the reference build is absent."""
    import tempfile
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    from avism_amd import synth
    L0 = n_side * box_per_cell
    p = synth.zeldovich(n_side, L0, seed=104)
    arr = p.numpy()
    if O.ref_available(1):
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        os.environ["OMP_NUM_THREADS"] = str(cores)
        with tempfile.TemporaryDirectory() as d:
            t0 = time.perf_counter()
            r = O.run_reference(d, 1, n_side, L0, 1, O.VoidParams(nvoid_max=4_000_000), parts=arr, k=k)
            dt = time.perf_counter() - t0
        return {"value": n_side ** 3 / dt, "unit": "cells/s", "cores": cores, "kind": "reference",
                "sample": f"reference routines (amdflang -O2 -fopenmp, {cores} threads) full level on "
                          f"{n_side}^3 Zel'dovich particles / {n_side}^3 cells, cell size {box_per_cell:.4g} "
                          f"Mpc/h, k={k}: {dt:.2f} s wall incl. a few MB of file I/O; "
                          f"{r['nvoid']} cubes, {r['nvoidt']} voids"}
    t0 = time.perf_counter()
    f = O.fields_from_particles(*arr, n_side, L0, 1, k=k)
    rete = O.cosmo(0.678, 0.31, 0.0)[1]
    O.void_stage(f["grid"], f["u1co"], f["div"], O.VoidParams(nvoid_max=4_000_000), meandens=f["meandens"],
                 rete=rete)
    dt = time.perf_counter() - t0
    return {"value": n_side ** 3 / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"oracle (C, -O2, 1 thread) full level on {n_side}^3 Zel'dovich particles / "
                      f"{n_side}^3 cells, cell size {box_per_cell:.4g} Mpc/h, k={k}: {dt:.2f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): this library drives a single "
              "device, and a sharded multi-GPU run is not implemented.", file=sys.stderr)
        sys.exit(2)
    import torch
    from avism_amd import synth
    from avism_amd._lib import Context
    from avism_amd import learned

    dev = torch.device("cuda:0")
    n = args.n
    t_gen = time.perf_counter()
    parts = synth.zeldovich(n, args.box, seed=args.seed, device=dev)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen
    ctx = Context(0)
    ctx.set_grid(n, args.box, 1)
    ctx.set_cosmology(0.678, 0.31, 0.0)
    ctx.set_void_params(-0.6, 10.0, 0.25, 0.0, 1.0, 1.0, 30_000_000)
    gi = ctx.grid()
    proposer = learned.Proposer.load_default(dev) if args.mode != "plain" else None
    if proposer is not None:
        proposer.install(ctx)

    results = {}
    catalogues = {}
    levels = {}
    GL = 2  # GROW_LEARNED
    modes = {"both": [0, GL], "plain": [0], "learned": [GL]}[args.mode]
    for mode in modes:
        if mode == GL and proposer is None:
            continue
        grid = (n, args.box, 1)
        for _ in range(args.warmup):
            run_step(ctx, parts, args.k, mode, proposer, args.levels, grid)
        torch.cuda.synchronize()
        ctx.set_timing(True)
        per_level = {} if args.levels > 1 else None
        t0 = time.perf_counter()
        for _ in range(args.steps):
            nv = run_step(ctx, parts, args.k, mode, proposer, args.levels, grid, per_level)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        tim = stage_table(ctx.timings())
        if per_level is not None:
            levels[mode] = [dict(n=v["n"], ms=round(v["ms"] / args.steps, 3), ncand=v["ncand"], ncubes=v["ncubes"],
                                 nvoids=v["nvoids"],
                                 stages_ms={k2: round(x / args.steps, 3) for k2, x in v["stages"].items()})
                            for _, v in sorted(per_level.items())]
            tim = stage_table([(k2, x * args.steps) for k2, x in levels[mode][0]["stages_ms"].items()])
        ctx.set_timing(False)
        st = ctx.stats()
        cat = ctx.catalogue(nv)
        catalogues[mode] = (ctx.marca(), cat["uvoid"].copy(), cat["indice"].copy())
        results[mode] = dict(ms=dt * 1e3, stages={k: v / args.steps for k, v in tim.items()},
                             ncubes=int(st.ncubes), nvoids=int(st.nvoidt), ncand=int(st.ncand))
    # The workload's particles all have the same mass (as N-body snapshots do),
    # so PPART takes k*m as the exact mass sum.  For the record, the field stage
    # once more with masses spread over [0.5, 1.5] x m (the neighbour-sum path).
    mixed = None
    if not args.no_mixed_mass and args.levels == 1:
        from avism_amd._lib import AVISM_DEVICE
        gen = torch.Generator(device=dev).manual_seed(7)
        m2 = parts.m * (0.5 + torch.rand(parts.m.shape, device=dev, generator=gen))
        torch.cuda.synchronize()
        ctx.load_particles(parts.x, parts.y, parts.z, parts.vx, parts.vy, parts.vz, m2, where=AVISM_DEVICE)
        ctx.set_timing(True)
        ctx.particles_to_grid(k=args.k)
        tim = stage_table(ctx.timings())
        ctx.set_timing(False)
        mixed = {s2: round(tim[s2], 3) for s2 in ("ppart_dens", "vvel_sph_vw", "ddens_scatter") if s2 in tim}
        del m2
    identical = None
    if 0 in catalogues and GL in catalogues:
        import numpy as np
        identical = all(np.array_equal(a, b) for a, b in zip(catalogues[0], catalogues[GL]))

    main_mode = GL if GL in results else 0
    r = results[main_mode]
    cells = n ** 3
    value = cells / (r["ms"] / 1e3)

    # roofline of the dominant kernel under the algorithmic byte model
    # (per-unit figures: neighbour records gathered per query, cell records
    # streamed per cell)
    k = args.k
    byte_model = {
        # gather model: k neighbour records (xyz, m, rho, velocity, HPART) + 3 outputs
        "vvel_sph_vw": cells * (k * 32 + 12 + k * 8),
        # gather model: k neighbour records (xyz, m) + 1 output
        "ppart_dens": n ** 3 * (k * 16 + 4),
        # compulsory: particle records (x,y,z,m,h) and NORM out, cell sums filled and read
        "ddens_scatter": n ** 3 * 24 + cells * 16,
        "ddens_finalize": cells * 8,
        "mark_flags": gi.next ** 3 * 9 + cells * 8,
        "mean_div": cells * 24,
    }
    stages = r["stages"]
    dom = max((s for s in stages if s in byte_model), key=lambda s: stages[s])
    achieved = byte_model[dom] / (stages[dom] / 1e3) / 1e9
    traffic = None  # HBM bytes per launch of that kernel, from the committed counter profile
    tfile = ROOT / "profiles" / "hbm_traffic.json"
    if tfile.exists():
        t = json.loads(tfile.read_text()).get("kernels", {}).get(dom)
        if t and t.get("n") == n and t.get("k") == k:
            traffic = t["bytes"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_model": byte_model[dom], "kernel_ms": round(stages[dom], 3)}
    # every modelled stage's algorithmic bytes over the whole step's wall time
    e2e_bytes = sum(byte_model[s2] for s2 in stages if s2 in byte_model)
    roofline["end_to_end_achieved"] = round(e2e_bytes / (r["ms"] / 1e3) / 1e9, 1)
    out = {
        "metric": "cells/s end-to-end at 512^3: learned-proposal vs plain GPU, identical voids",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Zel'dovich particles generated on the GPU)",
        "config": {"workload": f"zeldovich {n}^3 particles, {n}^3 grid (ext {gi.next}^3), periodic "
                               f"{args.box:g} Mpc/h, SPH k={k}"
                               + (f", {args.levels} void-in-void levels ({n}^3 .. {n * 2 ** (args.levels - 1)}^3 cells)"
                                  if args.levels > 1 else ""),
                   "growth": "learned" if main_mode else "plain"},
        "plain_ms_per_step": results.get(0, {}).get("ms"),
        "learned_ms_per_step": results.get(GL, {}).get("ms"),
        "learned_over_plain": (results[0]["ms"] / results[GL]["ms"]) if (0 in results and GL in results) else None,
        "identical_voids": identical,
        "stages_ms": {m: {k2: round(v, 3) for k2, v in results[m]["stages"].items()} for m in results},
        "ncand": r["ncand"], "ncubes": r["ncubes"], "nvoids": r["nvoids"],
        "roofline": roofline,
        "field_stage_ms_mixed_masses": mixed,
        "levels": {("learned" if m else "plain"): v for m, v in levels.items()} or None,
        "input_gen_s": round(t_gen, 2),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.box / n, k)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
