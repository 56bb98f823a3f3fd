"""Writes the golden fixtures under tests/golden/ by running the reference's own
routines (compiled in place by ``make -C oracle ref`` from /root/reference/src and
driven by oracle/ref_harness.f90).

Run in the build container (needs oracle/_ref):  python tests/golden/make_golden.py

Each fixture holds the inputs and the reference's outputs for one level:
  grid_p{0,1}_n24 : grid input (reader.f90:523-552) -> divergence, marking,
                    VOIDFIND, merge, post-processing, catalogue
  part_p{0,1}     : particle input (reader.f90:410-465) -> kNN, PPART_DENS,
                    VVEL_INTERP_SPH_VW, FILL_H_ZEROS, DDENS_INTERP_SPH, mean
                    density, divergence and the same void stage
  tsc_p{0,1}      : particle input with FLAG_VEL_INTERP = FLAG_DENS_INTERP = 0
                    -> VEL_INTERP_TSC, DENS_INTERP_TSC and the rest as above
  tscv_p{0,1}     : FLAG_VEL_INTERP = 0, FLAG_DENS_INTERP = 1 -> VEL_INTERP_TSC,
                    H_DISTANCE, FILL_H_ZEROS, DDENS_INTERP_SPH and the rest

  lev_p{0,1}      : two void-in-void levels on one particle set: the first
                    level at N, then the same particles at 2N with the first
                    level's MARCAP and REQP (avism.f90:2123-2161) -> the
                    second level's fields, cubes with FATHER, merge and
                    catalogue.  The reference's driver frees the particles
                    inside its level loop (avism.f90:1613-1629), so each level
                    is one run of the reference routines; RMIN_SUB is set so
                    that some parents are below it (voidfinding.f90:267-274).

  mask_p{0,1}, mask_g{0,1}
                  : particle (grid) input with a survey mask (SMASK, 20% of the cells
                    empty, the periodic shell masked as in survey mode): VVEL and
                    H_DISTANCE skip masked cells, MARK_ALL and the flag stencil
                    read it (voidfinding.f90:33, 283-300), DDENS ignores it

  amr_{p0,p1}_ir{m1,0,1}
                  : MASCLET-like AMR gas (base grid and level-1 patches, some
                    overlapping) onto the uniform grid of level IR: SMOOTH_2 or
                    the base grid, then DIVER_UNIFORM (IR <= 0); DIVER_FINA_GAS
                    and VMESH (IR = 1) -> U1GCO and DIVERGCO

`python tests/golden/make_golden.py tsc` (or `lev`, `mask`, `amr`) writes only those fixtures.
"""
from __future__ import annotations

import os
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import pyoracle as O  # noqa: E402
from avism_amd import synth  # noqa: E402

OUT = Path(__file__).resolve().parent
PARAMS = O.VoidParams(dens_thre=-0.6, dens_thre2=10.0, grad_thre=0.25, div_thre=0.0, rmin=1.0,
                      rmin_sub=1.0, nvoid_max=200000)


def _save(name, inputs: dict, ref: dict, meta: dict):
    keep = {k: v for k, v in ref.items() if isinstance(v, np.ndarray)}
    scal = {k: np.asarray(v) for k, v in ref.items() if not isinstance(v, np.ndarray)}
    np.savez_compressed(OUT / f"{name}.npz", **{f"in_{k}": v for k, v in inputs.items()},
                        **{f"ref_{k}": v for k, v in keep.items()},
                        **{f"ref_{k}": v for k, v in scal.items()},
                        **{f"meta_{k}": np.asarray(v) for k, v in meta.items()})
    print(name, {k: (v.shape if isinstance(v, np.ndarray) else v) for k, v in ref.items()
                 if k in ("nvoid", "nvoid2", "nvoidt")})


def tsc_fixtures(td: Path):
    """The TSC interpolations (particles.f90:16-351) through the reference's own
    routines: a Zel'dovich box (periodic) and mock voids (non-periodic)."""
    mock = synth.mock_voids(n0=5000, L0=60.0, nvoids=4, rmin=6.0, rmax=12.0, seed=25)[0].numpy()
    # non-periodic: particles kept two cells away from the faces -- the
    # reference's TSC writes outside its arrays for particles in the outer cells
    # (particles.f90:104-113 without the periodic wrap)
    keep = np.all(np.abs(np.stack(mock[:3])) < 30.0 - 2 * 3.0, axis=0)
    cases = (
        (1, synth.zeldovich(16, 40.0, seed=24, rms_disp=2.0).numpy(), 20, 40.0, 24),
        (0, [a[keep].copy() for a in mock], 20, 60.0, 25),
    )
    for per, arr, n, L0, seed in cases:
        for name, vi, di in (("tsc", 0, 0), ("tscv", 0, 1)):
            # one thread: VEL_INTERP_TSC adds REAL*4 with OMP ATOMIC, so only the
            # one-thread order is reproducible (particles.f90:226-249)
            os.environ["OMP_NUM_THREADS"] = "1"
            ref = O.run_reference(td / f"{name}{per}", 1, n, L0, per, PARAMS, k=16, parts=arr,
                                  vel_interp=vi, dens_interp=di)
            del os.environ["OMP_NUM_THREADS"]
            _save(f"{name}_p{per}", dict(zip(("x", "y", "z", "vx", "vy", "vz", "m"), arr)), ref,
                  dict(n=n, L0=L0, periodic=per, mode=1, k=16, seed=seed, vel_interp=vi, dens_interp=di))


def lev_fixtures(td: Path):
    cases = (
        (1, synth.zeldovich(16, 40.0, seed=22, rms_disp=2.0).numpy(), 16, 40.0, 22, 5.0),
        (0, synth.mock_voids(n0=5000, L0=60.0, nvoids=4, rmin=6.0, rmax=12.0, seed=23)[0].numpy(), 20, 60.0,
         23, 6.0),
    )
    for per, arr, n, L0, seed, rmin_sub in cases:
        P = O.VoidParams(**{**PARAMS.__dict__, "rmin_sub": rmin_sub})
        r1 = O.run_reference(td / f"lev{per}_1", 1, n, L0, per, P, k=16, parts=arr)
        marcap, reqp = O.next_level_inputs(O.Grid(n, L0, per), r1["marca"], r1["uvoid"], r1["vf"][11])
        r2 = O.run_reference(td / f"lev{per}_2", 1, 2 * n, L0, per, P, k=16, parts=arr, marcap=marcap, reqp=reqp)
        inputs = dict(zip(("x", "y", "z", "vx", "vy", "vz", "m"), arr))
        inputs.update(marcap=marcap, reqp=reqp)
        # the velocity and mass grids are checked at one level by part_p*; here
        # u1co, div and hpart (whose VVEL ownership depends on the grid) are kept
        ref = {k: v for k, v in r2.items() if k not in ("u", "v", "w", "mass", "part_dens")}
        ref.update(l1_marca=r1["marca"], l1_uvoid=r1["uvoid"], l1_req=r1["vf"][11], l1_nvoid=r1["nvoid"],
                   l1_nvoidt=r1["nvoidt"])
        _save(f"lev_p{per}", inputs, ref, dict(n=n, n2=2 * n, L0=L0, periodic=per, mode=1, k=16, seed=seed,
                                               rmin_sub=rmin_sub))


def mask_fixtures(td: Path):
    cases = (
        (1, synth.zeldovich(16, 40.0, seed=26, rms_disp=2.0).numpy(), 20, 40.0, 26),
        (0, synth.mock_voids(n0=5000, L0=60.0, nvoids=4, rmin=6.0, rmax=12.0, seed=27)[0].numpy(), 20, 60.0, 27),
    )
    for per, arr, n, L0, seed in cases:
        sm = synth.survey_mask(n, 0.2, seed=seed)
        ref = O.run_reference(td / f"mask{per}", 1, n, L0, per, PARAMS, k=16, parts=arr, smask=sm)
        inputs = dict(zip(("x", "y", "z", "vx", "vy", "vz", "m"), arr))
        inputs.update(smask=sm)
        _save(f"mask_p{per}", inputs, ref, dict(n=n, L0=L0, periodic=per, mode=1, k=16, seed=seed))
    # grid input with a mask: more voids for the masked marking and growth
    for per, seed in ((1, 28), (0, 29)):
        n, L0 = 24, 48.0
        d, vx, vy, vz = (a.numpy() for a in synth.grid_fields(n, L0, seed))
        sm = synth.survey_mask(n, 0.2, seed=seed)
        ref = O.run_reference(td / f"maskg{per}", 2, n, L0, per, PARAMS, fields=(d, vx, vy, vz), smask=sm)
        _save(f"mask_g{per}", dict(delta=d, vx=vx, vy=vy, vz=vz, smask=sm), ref,
              dict(n=n, L0=L0, periodic=per, mode=2, seed=seed))


def kbig_fixtures(td: Path):
    """Neighbour counts above 64: the reference's default KNEIGHBOURS = 256
    (periodic Zel'dovich, 4096 particles) and k = 100 (mock voids)."""
    p = synth.zeldovich(16, 40.0, seed=24, rms_disp=2.0)
    arr = p.numpy()
    ref = O.run_reference(td / "k256", 1, 16, 40.0, 1, PARAMS, k=256, parts=arr)
    _save("part_k256_p1", dict(zip(("x", "y", "z", "vx", "vy", "vz", "m"), arr)), ref,
          dict(n=16, L0=40.0, periodic=1, mode=1, k=256, seed=24))
    p, _, _ = synth.mock_voids(n0=5000, L0=60.0, nvoids=4, rmin=6.0, rmax=12.0, seed=25)
    arr = p.numpy()
    ref = O.run_reference(td / "k100", 1, 20, 60.0, 0, PARAMS, k=100, parts=arr)
    _save("part_k100_p0", dict(zip(("x", "y", "z", "vx", "vy", "vz", "m"), arr)), ref,
          dict(n=20, L0=60.0, periodic=0, mode=1, k=100, seed=25))


def novel_fixtures(td: Path):
    """Particles without velocities (FLAG_VEL_AVAILABLE = 0): density by SPH
    (H_DISTANCE, FILL_H_ZEROS, DDENS) or TSC, DIVERCO from linear theory."""
    for name, per, seed, di in (("nov_p0", 0, 41, 1), ("nov_p1", 1, 42, 1), ("novt_p1", 1, 43, 0)):
        if per:
            arr = synth.zeldovich(16, 40.0, seed=seed, rms_disp=2.0).numpy()
            n, L0 = 20, 40.0
        else:
            p, _, _ = synth.mock_voids(n0=5000, L0=60.0, nvoids=4, rmin=6.0, rmax=12.0, seed=seed)
            arr = p.numpy()
            n, L0 = 20, 60.0
        ref = O.run_reference(td / name, 1, n, L0, per, PARAMS, k=16, vel_interp=-1, dens_interp=di, parts=arr)
        _save(name, dict(zip(("x", "y", "z", "vx", "vy", "vz", "m"), arr)), ref,
              dict(n=n, L0=L0, periodic=per, mode=1, k=16, seed=seed, vel_interp=-1, dens_interp=di))


def survey_fields_fixtures(td: Path):
    """Survey mode (FLAG_SURVEY = 1): galaxies without velocities, randoms
    uniform in the box at 4 per cell, a survey mask with 20% of cells empty;
    density by SPH (KNEIGHBOURS2 = 128 for the randoms), periodic and not."""
    for name, per, seed, di in (("svy_p0", 0, 51, 1), ("svy_p1", 1, 52, 1)):
        p, _, _ = synth.mock_voids(n0=6000, L0=60.0, nvoids=4, rmin=6.0, rmax=12.0, seed=seed)
        arr = p.numpy()
        n, L0 = 20, 60.0
        rng = np.random.default_rng(seed)
        nr = 4 * n ** 3
        rands = [rng.uniform(-L0 / 2 + 1e-3, L0 / 2 - 1e-3, nr).astype(np.float32) for _ in range(3)]
        rands.append(np.full(nr, 1e10, np.float32))
        sm = synth.survey_mask(n, 0.2, seed=seed)
        ref = O.run_reference(td / name, 1, n, L0, per, PARAMS, k=16, vel_interp=-1, dens_interp=di, parts=arr,
                              smask=sm, rands=rands)
        inputs = dict(zip(("x", "y", "z", "vx", "vy", "vz", "m"), arr))
        inputs.update(smask=sm, rx=rands[0], ry=rands[1], rz=rands[2], rm=rands[3])
        _save(name, inputs, ref, dict(n=n, L0=L0, periodic=per, mode=1, k=16, seed=seed, vel_interp=-1,
                                      dens_interp=di))


AMR_CASES = [("amr_p0_ir1", 24, 48.0, 0, 1, 31), ("amr_p1_ir1", 24, 48.0, 1, 1, 32),
             ("amr_p1_ir0", 24, 48.0, 1, 0, 33), ("amr_p0_irm1", 32, 64.0, 0, -1, 34)]


def amr_fixtures(td: Path):
    for name, nb, L0, per, ir, seed in AMR_CASES:
        base, patches = synth.amr_gas(nb, L0, seed)
        pint, preal, pdata = synth.amr_pack(patches)
        u, d = O.run_reference_amr(td / name, nb, L0, per, ir, base, pint, preal, pdata)
        _save(name, dict(delta=base[0], vx=base[1], vy=base[2], vz=base[3], pint=pint, preal=preal, pdata=pdata),
              dict(u1gco=u, divgco=d), dict(nb=nb, L0=L0, periodic=per, ir=ir, seed=seed))


def main():
    for per in (0, 1):
        if not O.ref_available(per):
            raise SystemExit("oracle/_ref not built: run `make -C oracle ref` first")
    with tempfile.TemporaryDirectory(dir=ROOT / "_work" if (ROOT / "_work").exists() else None) as td:
        td = Path(td)
        if sys.argv[1:] == ["tsc"]:
            tsc_fixtures(td)
            return
        if sys.argv[1:] == ["lev"]:
            lev_fixtures(td)
            return
        if sys.argv[1:] == ["mask"]:
            mask_fixtures(td)
            return
        if sys.argv[1:] == ["kbig"]:
            kbig_fixtures(td)
            return
        if sys.argv[1:] == ["novel"]:
            novel_fixtures(td)
            return
        if sys.argv[1:] == ["survey"]:
            survey_fields_fixtures(td)
            return
        if sys.argv[1:] == ["amr"]:
            amr_fixtures(td)
            return
        tsc_fixtures(td)
        lev_fixtures(td)
        mask_fixtures(td)
        kbig_fixtures(td)
        novel_fixtures(td)
        survey_fields_fixtures(td)
        amr_fixtures(td)
        for per, seed in ((0, 11), (1, 12)):
            n, L0 = 24, 48.0
            d, vx, vy, vz = (a.numpy() for a in synth.grid_fields(n, L0, seed))
            ref = O.run_reference(td / f"g{per}", 2, n, L0, per, PARAMS, fields=(d, vx, vy, vz))
            _save(f"grid_p{per}_n{n}", dict(delta=d, vx=vx, vy=vy, vz=vz), ref,
                  dict(n=n, L0=L0, periodic=per, mode=2, seed=seed))
        # particle mode: Zel'dovich (periodic) and mock voids (non-periodic)
        p = synth.zeldovich(16, 40.0, seed=22, rms_disp=2.0)
        arr = p.numpy()
        ref = O.run_reference(td / "p1", 1, 20, 40.0, 1, PARAMS, k=16, parts=arr)
        _save("part_p1", dict(zip(("x", "y", "z", "vx", "vy", "vz", "m"), arr)), ref,
              dict(n=20, L0=40.0, periodic=1, mode=1, k=16, seed=22))
        p, _, _ = synth.mock_voids(n0=5000, L0=60.0, nvoids=4, rmin=6.0, rmax=12.0, seed=23)
        arr = p.numpy()
        ref = O.run_reference(td / "p0", 1, 20, 60.0, 0, PARAMS, k=16, parts=arr)
        _save("part_p0", dict(zip(("x", "y", "z", "vx", "vy", "vz", "m"), arr)), ref,
              dict(n=20, L0=60.0, periodic=0, mode=1, k=16, seed=23))
        # INDEXX permutations with many ties (the tie order decides cube and void
        # ids), run by the reference routine itself
        rng = np.random.default_rng(23)
        sets = {
            "ties_small": rng.integers(0, 40, size=3000).astype(np.float32),
            "ties_few": rng.integers(0, 3, size=2000).astype(np.float32),
            "distinct": rng.standard_normal(5000).astype(np.float32),
            "sorted": np.arange(1000, dtype=np.float32),
            "reversed": np.arange(1000, 0, -1).astype(np.float32),
            "all_equal": np.ones(777, np.float32),
            "tiny": np.array([3.0, 1.0, 2.0, 1.0, 3.0], np.float32),
        }
        out = {}
        for name, keys in sets.items():
            out[f"keys_{name}"] = keys
            out[f"indx_{name}"] = O.run_reference_indexx(td / f"ix_{name}", keys)
        np.savez_compressed(OUT / "indexx.npz", **out)


if __name__ == "__main__":
    main()
