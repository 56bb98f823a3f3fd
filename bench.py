#!/usr/bin/env python3
"""Benchmark of the AVISM grid-and-void path on one MI355X.

Workload: 512^3 Zel'dovich particles (P(k) ~ k^-1.5 exp(-k^2/0.25), rms
displacement 3 Mpc/h, f = 0.5) in a periodic 1000 Mpc/h box, SPH with k = 32
onto a 512^3 grid (extended to 640^3 by the periodic shell), thresholds of
config/voids.dat.

One step (the headline, `value`) runs from the particle SoA resident in HBM to
the catalogue and the MARCA / U1CO / DIVERCO maps resident in HBM: binning,
fields, marking, growth, acceptance, merge and post-processing; its stage
timings give the roofline.  The same step from the particle SoA in pinned host
RAM to the catalogue and maps in pinned host RAM (the PCIe transfers included:
upload, velocities overlapped with binning and PPART_DENS, U1CO and DIVERCO
downloaded while the void stage runs) is reported beside it (`host_ram_*`).

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
import hashlib
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# kernel sources whose HBM traffic profiles/hbm_traffic.json records: a figure
# measured on other sources is not reported
TRAFFIC_SOURCES = ["avism_amd/csrc/fields.hip", "avism_amd/csrc/knn_search.h", "avism_amd/csrc/voids.hip"]


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
    ap.add_argument("--no-host", action="store_true", help="time the HBM-resident step only")
    ap.add_argument("--cpu-n", type=int, default=128, help="side of the CPU-baseline sample")
    ap.add_argument("--mode", choices=["both", "plain", "learned"], default="both")
    ap.add_argument("--levels", type=int, default=1, help="void-in-void levels per step (N, 2N, ...)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-patch batch record")
    ap.add_argument("--refine", type=int, default=0,
                    help="refined void-in-void levels in windows around the parent voids (avism_refine_*)")
    ap.add_argument("--margin", type=int, default=8, help="window margin of the refined levels, cells")
    ap.add_argument("--max-cells", type=int, default=1 << 30, help="cells per atlas of a refined level")
    ap.add_argument("--config", choices=["c4", "c5"], default="c4",
                    help="c5: 1024^3 particles in a periodic 2000 Mpc/h box, x2 and x4 refinement in the "
                         "parent voids (the refined-level run)")
    a = ap.parse_args()
    if a.config == "c5":
        a.n, a.box, a.seed = 1024, 2000.0, 5
        a.refine = a.refine or 2
    return a


def sources_hash() -> str:
    h = hashlib.sha256()
    for s in TRAFFIC_SOURCES:
        h.update((ROOT / s).read_bytes())
    return h.hexdigest()[:16]


def progress(msg):
    """A line on stderr when a phase of the benchmark has ended (stdout holds
    only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_table(tim):
    out = {}
    for name, ms in tim:
        out[name] = out.get(name, 0.0) + ms
    return out


class HostOut:
    """Pinned host buffers the host-RAM step downloads into, per grid size."""

    def __init__(self):
        self.bufs = {}

    def get(self, n):
        import torch
        if n not in self.bufs:
            n3 = n ** 3
            self.bufs[n] = dict(u1co=torch.empty(n3, dtype=torch.float32).pin_memory(),
                                div=torch.empty(n3, dtype=torch.float32).pin_memory(),
                                marca=torch.empty(n3, dtype=torch.int32).pin_memory())
        return self.bufs[n]


def run_step(ctx, parts, k, growth_mode, levels=1, grid=None, per_level=None, host_out=None):
    """One step.  host_out None: particles from HBM, results left in HBM.
    Otherwise particles from pinned host RAM (parts holds host tensors) and
    the catalogue and the maps downloaded into host_out's pinned buffers."""
    from avism_amd._lib import AVISM_DEVICE, AVISM_HOST
    if levels > 1:
        ctx.set_grid(*grid)  # back to the base level
    ctx.load_particles(parts.x, parts.y, parts.z, parts.vx, parts.vy, parts.vz, parts.m,
                       where=AVISM_HOST if host_out is not None else AVISM_DEVICE)
    cat = None
    for lev in range(levels):
        t0 = time.perf_counter()
        if lev:
            ctx.next_level()
        ctx.particles_to_grid(k=k)
        n = ctx.grid().n
        if host_out is not None:
            hb = host_out.get(n)
            ctx.fields_async(hb["u1co"], hb["div"])  # overlaps the void stage
        nv = ctx.find_voids(growth_mode)
        ctx.merge_and_post()
        if host_out is not None:
            cat = ctx.catalogue(nv)
            ctx.marca(inner_only=True, out=hb["marca"])
            ctx.sync()
        if per_level is not None:  # the library calls return with their work complete
            st = ctx.stats()
            rec = per_level.setdefault(lev, dict(n=n, ms=0.0, stages={}))
            rec["ms"] += (time.perf_counter() - t0) * 1e3
            rec.update(ncand=int(st.ncand), ncubes=int(nv), nvoids=int(st.nvoidt))
            for name, ms in ctx.timings():
                rec["stages"][name] = rec["stages"].get(name, 0.0) + ms
    return nv, cat


def cpu_baseline(n_side: int, box_per_cell: float, k: int):
    """The reference's own routines (oracle/_ref: its Fortran sources built
    with amdflang -O2 -fopenmp, particles.f90 included, driven by the oracle
    harness) on a bounded sample of the same workload: n_side^3 particles and
    cells, same cell size, all host cores.  Falls back to the oracle (C port,
    1 thread) when the reference build is absent."""
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
        omp_fields = O.ref_fields_openmp()
        with tempfile.TemporaryDirectory() as d:
            t0 = time.perf_counter()
            r = O.run_reference(d, 1, n_side, L0, 1, O.VoidParams(nvoid_max=4_000_000), parts=arr, k=k)
            dt = time.perf_counter() - t0
        how = ("every routine OpenMP" if omp_fields else
               "field routines (particles.f90) on 1 thread, the rest OpenMP")
        return {"value": n_side ** 3 / dt, "unit": "cells/s", "cores": cores, "kind": "reference",
                "sample": f"reference routines (amdflang -O2 -fopenmp, {how}, {cores} threads) full level on "
                          f"{n_side}^3 Zel'dovich particles / {n_side}^3 cells, periodic {L0:g} Mpc/h (same cell "
                          f"size), k={k}: {dt:.2f} s wall incl. file I/O of the harness; "
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


def batch_bench(dev, k, steps, warmup):
    """64 independent survey-like patches in one batch (Zel'dovich fields,
    particle counts log-uniform in 2e5..8e6, grids of 32^3..160^3 cells,
    non-periodic, 20% of each patch's cells masked): one step = pinned host
    particles and masks -> every patch's fields (one graph launch) -> every
    patch's catalogue and MARCA (void stages on 8 host threads).  Beside it, the
    same patches one at a time through the single-grid path."""
    import numpy as np
    import torch
    from avism_amd import synth
    from avism_amd._lib import Context
    t = time.perf_counter()
    # on the host: 64 lattice sides would need as many GPU FFT plans
    pats = synth.patch_batch(64, seed0=3000, device="cpu")
    names = ("x", "y", "z", "vx", "vy", "vz", "m")
    host = [torch.cat([getattr(p["particles"], f) for p in pats]).pin_memory() for f in names]
    mask = torch.from_numpy(np.concatenate([p["mask"] for p in pats])).pin_memory()
    gen_s = time.perf_counter() - t
    ns = [p["n"] for p in pats]
    nps = [p["particles"].n for p in pats]
    ctx = Context(0)
    ctx.set_cosmology(0.678, 0.31, 0.0)
    ctx.set_void_params(-0.6, 10.0, 0.25, 0.0, 1.0, 1.0, 30_000_000)
    ctx.batch_set(ns, [p["L0"] for p in pats], nps)

    def step():
        ctx.batch_load(*host, smask=mask)
        ctx.batch_fields(k=k)
        t1 = time.perf_counter()
        ctx.batch_voids(threads=8)
        return time.perf_counter() - t1

    t = time.perf_counter()
    step()  # tunes and captures
    first_s = time.perf_counter() - t
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ctx.set_timing(True)
    vs = 0.0
    t = time.perf_counter()
    for _ in range(steps):
        vs += step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) / steps * 1e3
    st = stage_table(ctx.timings())
    ctx.set_timing(False)
    marca_b = [ctx.batch_patch(i).marca(inner_only=True) for i in range(len(pats))]
    # the same patches one at a time through the single-grid path
    single = Context(0)
    single.set_cosmology(0.678, 0.31, 0.0)
    single.set_void_params(-0.6, 10.0, 0.25, 0.0, 1.0, 1.0, 30_000_000)
    off = np.concatenate([[0], np.cumsum(nps)])
    moff = np.concatenate([[0], np.cumsum([n ** 3 for n in ns])])

    def one(i):
        single.set_grid(ns[i], pats[i]["L0"], 0)
        single.set_mask(mask[moff[i]:moff[i + 1]].numpy())
        a, b = int(off[i]), int(off[i + 1])
        single.load_particles(*[h[a:b] for h in host])
        single.particles_to_grid(k=k)
        single.find_voids(0)
        single.merge_and_post()
        return single.marca(inner_only=True)

    one(0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    same = all(np.array_equal(one(i), marca_b[i]) for i in range(len(pats)))
    unbatched_ms = (time.perf_counter() - t) * 1e3
    info = ctx.batch_info()
    cells = int(sum(n ** 3 for n in ns))
    return {"patches": len(pats), "particles": int(sum(nps)), "cells": cells,
            "grids": {str(n): ns.count(n) for n in sorted(set(ns))},
            "ms_per_batch": round(ms, 2), "patches_per_s": round(len(pats) / (ms / 1e3), 1),
            "cells_per_s": cells / (ms / 1e3),
            "stages_ms": {k2: round(v / steps, 3) for k2, v in st.items()},
            "void_stages_wall_ms": round(vs / steps * 1e3, 2),
            "unbatched_ms": round(unbatched_ms, 2), "batched_over_unbatched": round(unbatched_ms / ms, 3),
            "identical_to_unbatched": same, "graph_launches": info["graph_launches"],
            "first_batch_s_incl_tuning": round(first_s, 2), "generation_s": round(gen_s, 1),
            "tuning": info["tuning"]}


def refine_walk(ctx, k, mode, depth, margin, max_cells, recs, lev=1):
    """Refined levels lev..depth in windows around the parent voids, depth
    first: plan the level from the current catalogue, then each atlas's fields,
    void stage and catalogue (downloaded), and its own sublevels."""
    import torch
    t = time.perf_counter()
    na = ctx.refine_plan(margin=margin, max_cells=max_cells)
    rec = recs.setdefault(lev, dict(atlases=0, windows=0, window_cells=0, atlas_cells=0, max_side=0, ncand=0,
                                    ncubes=0, nvoids=0, ms=0.0, min_free_gb=1e9, stages={}))
    rec["ms"] += (time.perf_counter() - t) * 1e3
    for a in range(na):
        t = time.perf_counter()
        ctx.refine_load(a, k)
        _, info = ctx.refine_windows()
        nv = ctx.find_voids(mode)
        ctx.merge_and_post()
        cat = ctx.catalogue(nv)
        st = cat["stats"]
        rec["ms"] += (time.perf_counter() - t) * 1e3
        rec["atlases"] += 1
        rec["windows"] += int(info["nwindows"])
        rec["window_cells"] += int(info["window_cells"])
        rec["atlas_cells"] += int(info["side"]) ** 3
        rec["max_side"] = max(rec["max_side"], int(info["side"]))
        rec["ncand"] += int(st.ncand)
        rec["ncubes"] += int(nv)
        rec["nvoids"] += int(st.nvoidt)
        rec["min_free_gb"] = min(rec["min_free_gb"], torch.cuda.mem_get_info()[0] / 2 ** 30)
        for name, ms in ctx.timings():
            rec["stages"][name] = rec["stages"].get(name, 0.0) + ms
        progress(f"level {lev + 1} atlas {a + 1}/{na}: {int(info['nwindows'])} windows, side {int(info['side'])}, "
                 f"{int(nv)} cubes, {int(st.nvoidt)} voids, {(time.perf_counter() - t) * 1e3:.0f} ms")
        if lev < depth:
            refine_walk(ctx, k, mode, depth, margin, max_cells, recs, lev + 1)
    ctx.refine_pop()
    return recs


def refine_bench(args):
    """The refined-level run: the base level on the whole box, then
    args.refine levels (x2, x4, ...) computed only in windows around the
    parent voids, atlas by atlas (memory bounded by the largest atlas).  One
    step = particles in HBM -> base fields, voids and catalogue -> every
    refined level's atlases and catalogues (catalogues downloaded)."""
    import torch
    from avism_amd import learned, synth
    from avism_amd._lib import AVISM_DEVICE, Context
    dev = torch.device("cuda:0")
    n, k = args.n, args.k
    t = time.perf_counter()
    parts = synth.zeldovich(n, args.box, seed=args.seed, device=dev)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t
    torch.cuda.empty_cache()
    progress(f"generated {n}^3 particles in {gen_s:.1f} s")
    ctx = Context(0)
    ctx.set_grid(n, args.box, 1)
    ctx.set_cosmology(0.678, 0.31, 0.0)
    ctx.set_void_params(-0.6, 10.0, 0.25, 0.0, 1.0, 1.0, 30_000_000)
    gi = ctx.grid()
    GL = 2
    mode = 0 if args.mode == "plain" else GL
    if mode == GL:
        learned.Proposer.load_default(dev).install(ctx)
    total_gb = torch.cuda.mem_get_info()[1] / 2 ** 30

    def step(recs):
        ctx.set_grid(n, args.box, 1)
        ctx.load_particles(parts.x, parts.y, parts.z, parts.vx, parts.vy, parts.vz, parts.m, where=AVISM_DEVICE)
        t0 = time.perf_counter()
        ctx.particles_to_grid(k=k)
        nv = ctx.find_voids(mode)
        ctx.merge_and_post()
        cat = ctx.catalogue(nv)
        st = cat["stats"]
        base = dict(n=n, ms=(time.perf_counter() - t0) * 1e3, ncand=int(st.ncand), ncubes=int(nv),
                    nvoids=int(st.nvoidt), free_gb=torch.cuda.mem_get_info()[0] / 2 ** 30,
                    stages={})
        for name, ms in ctx.timings():
            base["stages"][name] = base["stages"].get(name, 0.0) + ms
        progress(f"base level {n}^3: {base['ms']:.0f} ms, {int(nv)} cubes, {int(st.nvoidt)} voids")
        refine_walk(ctx, k, mode, args.refine, args.margin, args.max_cells, recs)
        return base

    for _ in range(args.warmup):
        step({})
        progress("warm-up step done")
    torch.cuda.synchronize()
    ctx.set_timing(True)
    ctx.timings()
    recs, bases = {}, []
    t = time.perf_counter()
    for _ in range(args.steps):
        bases.append(step(recs))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.steps
    ctx.set_timing(False)
    progress(f"refined-level step {dt:.2f} s")
    s_ = args.steps
    levels = [dict(n=n, ms=round(sum(b["ms"] for b in bases) / s_, 1), ncand=bases[-1]["ncand"],
                   ncubes=bases[-1]["ncubes"], nvoids=bases[-1]["nvoids"],
                   min_free_gb=round(min(b["free_gb"] for b in bases), 1),
                   stages_ms={kk: round(sum(b["stages"].get(kk, 0.0) for b in bases) / s_, 2)
                              for kk in bases[-1]["stages"]})]
    for lev in sorted(recs):
        r = recs[lev]
        levels.append(dict(n=n * 2 ** lev, ms=round(r["ms"] / s_, 1), atlases=r["atlases"] // s_,
                           windows=r["windows"] // s_, window_cells=r["window_cells"] // s_,
                           atlas_cells=r["atlas_cells"] // s_, max_atlas_side=r["max_side"],
                           whole_grid_cells=(n * 2 ** lev) ** 3,
                           ncand=r["ncand"] // s_, ncubes=r["ncubes"] // s_, nvoids=r["nvoids"] // s_,
                           min_free_gb=round(r["min_free_gb"], 1),
                           stages_ms={kk: round(v / s_, 2) for kk, v in r["stages"].items()}))
    peak_gb = total_gb - min(lv["min_free_gb"] for lv in levels)
    return {
        "metric": f"cells/s of a {n}^3-particle run with {args.refine} refined levels in the parent voids",
        "value": n ** 3 / dt, "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Zel'dovich particles generated on the GPU)",
        "config": {"workload": f"zeldovich {n}^3 particles, {n}^3 base grid (ext {gi.next}^3), periodic "
                               f"{args.box:g} Mpc/h, SPH k={k}; refined levels x2..x{2 ** args.refine} in windows "
                               f"(margin {args.margin} cells, atlases of at most {args.max_cells} cells)",
                   "growth": "learned" if mode else "plain",
                   "boundary": "particles in HBM -> every level's catalogue in host RAM"},
        "levels": levels, "device_memory_gb": round(total_gb, 1), "peak_device_use_gb": round(peak_gb, 1),
        "input_gen_s": round(gen_s, 2),
    }


def timed(ctx, fn, warmup, steps):
    import torch
    for _ in range(warmup):
        fn(None)
    torch.cuda.synchronize()
    ctx.timings()  # drop the warm-up's records
    ctx.set_timing(True)
    per_level = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn(per_level)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    tim = ctx.timings()
    ctx.set_timing(False)
    return dt, tim, per_level, out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): this library drives a single "
              "device, and a sharded multi-GPU run is not implemented.", file=sys.stderr)
        sys.exit(2)
    if args.refine > 0:
        print(json.dumps(refine_bench(args)))
        return
    import numpy as np
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
    parts_h = None
    if not args.no_host:  # the same particles in pinned host RAM
        f = lambda t: t.cpu().pin_memory()
        parts_h = synth.Particles(f(parts.x), f(parts.y), f(parts.z), f(parts.vx), f(parts.vy), f(parts.vz),
                                  f(parts.m), parts.L0, parts.periodic)
    ctx = Context(0)
    ctx.set_grid(n, args.box, 1)
    ctx.set_cosmology(0.678, 0.31, 0.0)
    ctx.set_void_params(-0.6, 10.0, 0.25, 0.0, 1.0, 1.0, 30_000_000)
    gi = ctx.grid()
    proposer = learned.Proposer.load_default(dev) if args.mode != "plain" else None
    if proposer is not None:
        proposer.install(ctx)

    GL = 2  # GROW_LEARNED
    modes = {"both": [0, GL], "plain": [0], "learned": [GL]}[args.mode]
    grid = (n, args.box, 1)
    host_out = HostOut()
    dres, hres, catalogues, levels = {}, {}, {}, {}
    for mode in modes:
        lv = args.levels
        # HBM-resident step: stage timings and roofline
        dt, tim, per_level, _ = timed(
            ctx, lambda pl: run_step(ctx, parts, args.k, mode, lv, grid, pl if lv > 1 else None),
            args.warmup, args.steps)
        st = ctx.stats()
        counters = ctx.counters()
        stages = {s: v / args.steps for s, v in stage_table(tim).items()}
        if lv > 1:
            levels[mode] = [dict(n=v["n"], ms=round(v["ms"] / args.steps, 3), ncand=v["ncand"],
                                 ncubes=v["ncubes"], nvoids=v["nvoids"],
                                 stages_ms={k2: round(x / args.steps, 3) for k2, x in v["stages"].items()})
                            for _, v in sorted(per_level.items())]
            stages = dict(levels[mode][0]["stages_ms"])
        dres[mode] = dict(ms=dt * 1e3, stages=stages, ncubes=int(st.ncubes), nvoids=int(st.nvoidt),
                          ncand=int(st.ncand), counters=counters)
        progress(f"growth mode {mode}: HBM-resident step {dt * 1e3:.1f} ms")
        if parts_h is not None:
            dt_h, tim_h, _, out = timed(
                ctx, lambda pl: run_step(ctx, parts_h, args.k, mode, lv, grid, None, host_out),
                args.warmup, args.steps)
            hres[mode] = dict(ms=dt_h * 1e3, stages={s: v / args.steps for s, v in stage_table(tim_h).items()})
            progress(f"growth mode {mode}: host-RAM step {dt_h * 1e3:.1f} ms")
            cat = out[1]
            hb = host_out.get(ctx.grid().n)
            catalogues[mode] = (hb["marca"].numpy().copy(), cat["uvoid"].copy(), cat["indice"].copy(),
                                hb["u1co"].numpy().copy(), hb["div"].numpy().copy())
        else:
            st_nv = ctx.stats().ncubes
            cat = ctx.catalogue(int(st_nv))
            catalogues[mode] = (ctx.marca(), cat["uvoid"].copy(), cat["indice"].copy())
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
        identical = all(np.array_equal(a, b) for a, b in zip(catalogues[0], catalogues[GL]))

    main_mode = GL if GL in dres else 0
    r = dres[main_mode]
    cells = n ** 3
    use_host = main_mode in hres
    ms_step = r["ms"]
    value = cells / (ms_step / 1e3)

    # Σ stages of the HBM-resident step against its wall time: the stage
    # timers cover the whole step (gaps are host work between launches)
    stages = r["stages"]
    stage_sum = sum(stages.values())
    coverage = stage_sum / r["ms"]
    if args.levels == 1 and not (0.95 <= coverage <= 1.02):
        print(f"warning: stage timings sum to {stage_sum:.1f} ms of a {r['ms']:.1f} ms step", file=sys.stderr)

    # roofline of the dominant kernel under the algorithmic byte model
    # (per-unit figures: neighbour records gathered per query, cell records
    # streamed per cell)
    k = args.k
    cnt = r["counters"]
    byte_model = {
        # gather model: k neighbour records (xyz, m, rho, velocity, HPART) + 3 outputs
        "vvel_sph_vw": cnt["vvel_queries"] * (k * 32 + 12 + k * 8),
        # gather model: k neighbour records (xyz, m) + 1 output
        "ppart_dens": cnt["ppart_queries"] * (k * 16 + 4),
        # gather model: 24 bytes per (particle, cell) pair within h + the cell sums
        "ddens_scatter": cnt["ddens_pairs"] * 24 + cells * 4,
        # compulsory: read the int64 cell sums, write U1
        "ddens_finalize": cells * 8 + cells * 4,
        # compulsory: read 3 velocity fields + U1, write U1CO and DIVERCO
        "mean_div": cells * 24,
        # the reference's bytes per extended cell: read U1CO, DIVERCO, MARCAP2
        # (4 B each) and SMASK (1 B), write the 8 flag bytes
        "mark_flags": gi.next ** 3 * 21,
    }
    dom = max((s for s in stages if s in byte_model), key=lambda s: stages[s])
    achieved = byte_model[dom] / (stages[dom] / 1e3) / 1e9
    traffic = None  # HBM bytes per launch of that kernel, from counters measured on these sources
    tfile = ROOT / "profiles" / "hbm_traffic.json"
    src_hash = sources_hash()
    if tfile.exists():
        tj = json.loads(tfile.read_text())
        t = tj.get("kernels", {}).get(dom)
        if t and t.get("n") == n and t.get("k") == k and tj.get("sources") == src_hash:
            traffic = t["bytes"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_model": byte_model[dom], "kernel_ms": round(stages[dom], 3), "sources": src_hash}
    e2e_bytes = sum(byte_model[s2] for s2 in stages if s2 in byte_model)
    roofline["end_to_end_achieved"] = round(e2e_bytes / (r["ms"] / 1e3) / 1e9, 1)
    roofline["per_kernel_frac"] = {s2: round(byte_model[s2] / (stages[s2] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                                   for s2 in byte_model if s2 in stages}
    knn = {}
    if "ppart_dens" in stages:
        knn["ppart_queries_per_s"] = cnt["ppart_queries"] / (stages["ppart_dens"] / 1e3)
    if "vvel_sph_vw" in stages:
        knn["vvel_queries_per_s"] = cnt["vvel_queries"] / (stages["vvel_sph_vw"] / 1e3)
    if "ppart_dens" in stages and "vvel_sph_vw" in stages:
        knn["all_queries_per_s"] = ((cnt["ppart_queries"] + cnt["vvel_queries"]) /
                                    ((stages["ppart_dens"] + stages["vvel_sph_vw"]) / 1e3))
    out = {
        "metric": "cells/s end-to-end at 512^3: learned-proposal vs plain GPU, identical voids",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Zel'dovich particles generated on the GPU; pinned host copies for host_ram_*)",
        "config": {"workload": f"zeldovich {n}^3 particles, {n}^3 grid (ext {gi.next}^3), periodic "
                               f"{args.box:g} Mpc/h, SPH k={k}"
                               + (f", {args.levels} void-in-void levels ({n}^3 .. {n * 2 ** (args.levels - 1)}^3 cells)"
                                  if args.levels > 1 else ""),
                   "growth": "learned" if main_mode else "plain",
                   "boundary": "particle SoA in HBM -> catalogue + MARCA/U1CO/DIVERCO maps in HBM"},
        "plain_ms_per_step": dres.get(0, {}).get("ms"),
        "learned_ms_per_step": dres.get(GL, {}).get("ms"),
        "learned_over_plain": None,
        "identical_voids": identical,
        # the same step from the particle SoA in pinned host RAM to the catalogue
        # and maps in pinned host RAM: PCIe transfers included
        "host_ram_ms_per_step": {("learned" if m else "plain"): round(v["ms"], 3) for m, v in hres.items()} or None,
        "host_ram_value": cells / (hres[main_mode]["ms"] / 1e3) if use_host else None,
        "stage_sum_over_step": round(coverage, 4),
        "stages_ms": {("learned" if m else "plain"): {k2: round(v, 3) for k2, v in dres[m]["stages"].items()}
                      for m in dres},
        "host_step_stages_ms": {("learned" if m else "plain"): {k2: round(v, 3) for k2, v in hres[m]["stages"].items()
                                                              if k2.startswith(("upload", "download"))}
                                for m in hres} or None,
        "ncand": r["ncand"], "ncubes": r["ncubes"], "nvoids": r["nvoids"],
        "counters": cnt,
        # what the proposals change in the growth: its time, the face cells it
        # examines and the candidates grown by one wave each (HBM-resident step)
        "growth": {("learned" if m else "plain"): {
            "grow_ms": round(v["stages"].get("grow_proposal" if m else "grow", 0.0), 3),
            "propose_ms": round(v["stages"].get("cnn_propose", 0.0), 3),
            "cells_examined": v["counters"]["growth_cells"],
            "wave_candidates": v["counters"]["growth_wave_candidates"]} for m, v in dres.items()},
        "knn": knn,
        "roofline": roofline,
        "field_stage_ms_mixed_masses": mixed,
        "levels": {("learned" if m else "plain"): v for m, v in levels.items()} or None,
        "input_gen_s": round(t_gen, 2),
    }
    if 0 in dres and GL in dres:
        out["learned_over_plain"] = dres[0]["ms"] / dres[GL]["ms"]
    if not args.no_batch and args.levels == 1:
        del parts, parts_h
        torch.cuda.empty_cache()
        out["batch64"] = batch_bench(dev, k, steps=max(1, min(args.steps, 3)), warmup=1)
        progress(f"batch of 64 patches: {out['batch64']['ms_per_batch']} ms")
    if not args.no_cpu_baseline:
        progress(f"CPU baseline on {args.cpu_n}^3")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.box / n, k)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
