"""GPU parity of void-in-void levels (avism.f90:742-759, 957-968, 2123-2161):
the parent map and radii (MARCAP, REQP), growth confined to one parent with
small parents stopped (FLAG1_GRID), the merge that keeps parents apart
(FATHER), and the hand-off of avism_next_level.

Given identical fields the second level's catalogue is bit-exact in MARCA,
cube extents, FATHER, UVOID and ids (tolerances on GXC, EPS/IP, UMEAN/MTOT as
in test_gpu_voids).  Where the GPU computes the fields itself, its void stage
is compared with the oracle's on those same fields.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O
from test_gpu_voids import PARAMS, compare, gpu_void_outputs

pytestmark = pytest.mark.gpu


def _ctx(n, L0, per, rmin_sub):
    from avism_amd._lib import Context
    c = Context(0)
    c.set_grid(n, L0, per)
    c.set_cosmology(0.678, 0.31, 0.0)
    c.set_void_params(PARAMS.dens_thre, PARAMS.dens_thre2, PARAMS.grad_thre, PARAMS.div_thre, PARAMS.rmin,
                      rmin_sub, PARAMS.nvoid_max)
    return c


def _oracle_params(rmin_sub):
    return O.VoidParams(**{**PARAMS.__dict__, "rmin_sub": rmin_sub})


def _as_ref(res: O.LevelResult) -> dict:
    return dict(nvoid=res.nvoid, cubes_i=res.cubes_i, cubes_f=res.cubes_f, father=res.father, indice=res.indice,
                nvoid2=res.nvoid2, uvoid=res.uvoid, marca=res.marca, nvoidt=res.nvoidt, ncellv=res.ncellv,
                volt_clean=res.volt_clean, vf=res.vf)


@pytest.mark.parametrize("name", ["lev_p0", "lev_p1"])
def test_sublevel_on_reference_fields(golden_dir, name):
    """The second level of the reference's own two-level run: its fields,
    MARCAP and REQP fed to the GPU void stage."""
    g = dict(np.load(golden_dir / f"{name}.npz"))
    n2, L0, per = int(g["meta_n2"]), float(g["meta_L0"]), int(g["meta_periodic"])
    print(f"[level 2 on reference fields {name}] n={n2}", flush=True)
    c = _ctx(n2, L0, per, float(g["meta_rmin_sub"]))
    c.load_void_fields(g["ref_u1co"], g["ref_div"], float(g["ref_meandens"]))
    c.set_parent_map(g["in_marcap"])
    c.set_parent_radii(g["in_reqp"])
    out = gpu_void_outputs(c)
    compare(out, lambda k: g[f"ref_{k}"])
    assert np.array_equal(out["father"], g["ref_father"])


@pytest.mark.parametrize("name", ["lev_p0", "lev_p1"])
def test_level_driver_vs_oracle(golden_dir, name):
    """Particles -> level 1 -> avism_next_level -> level 2 on the GPU.  Each
    level's void stage equals the oracle's on the GPU's fields, the hand-off
    equals the oracle's from the GPU's level-1 catalogue, and the level-2
    fields stay within the field tolerances of the reference's."""
    g = dict(np.load(golden_dir / f"{name}.npz"))
    n, n2, L0, per, k = (int(g["meta_n"]), int(g["meta_n2"]), float(g["meta_L0"]), int(g["meta_periodic"]),
                         int(g["meta_k"]))
    rmin_sub = float(g["meta_rmin_sub"])
    P = _oracle_params(rmin_sub)
    rete = O.cosmo(0.678, 0.31, 0.0)[1]
    arr = [g[f"in_{c}"] for c in ("x", "y", "z", "vx", "vy", "vz", "m")]
    print(f"[level driver {name}] n={n} -> {n2}", flush=True)
    c = _ctx(n, L0, per, rmin_sub)
    c.load_particles(*arr)
    c.particles_to_grid(k=k)
    f1 = c.fields()
    out1 = gpu_void_outputs(c)
    g1 = O.Grid(n, L0, per)
    r1 = O.void_stage(g1, f1["u1co"], f1["div"], P, meandens=c.stats().meandens, rete=rete)
    compare(out1, lambda kk: _as_ref(r1)[kk])
    marcap, reqp = O.next_level_inputs(g1, out1["marca"], out1["uvoid"], out1["vf"][11])
    gi = c.next_level()
    assert gi.n == n2
    c.particles_to_grid(k=k)
    f2 = c.fields()
    rel = np.max(np.abs(f2["u1co"] - g["ref_u1co"]) / np.maximum(np.abs(g["ref_u1co"]), 1e-30))
    assert rel <= 1e-5
    rms = np.sqrt(np.mean((f2["div"] - g["ref_div"]) ** 2) / max(np.mean(g["ref_div"] ** 2), 1e-30))
    assert rms <= 1e-5
    out2 = gpu_void_outputs(c)
    g2 = O.Grid(n2, L0, per)
    r2 = O.void_stage(g2, f2["u1co"], f2["div"], P, marcap=marcap, reqp=reqp, meandens=c.stats().meandens,
                      rete=rete)
    print(f"[level driver {name}] level 2: cubes={out2['nvoid']} voids={out2['nvoidt']}", flush=True)
    assert out2["nvoid"] > 0
    compare(out2, lambda kk: _as_ref(r2)[kk])
    assert np.array_equal(out2["father"], r2.father)


@pytest.mark.parametrize("per,seed", [(1, 301), (0, 302)])
def test_grid_levels_vs_oracle(per, seed):
    """Grid input at 32^3, then a second level on 64^3 grid fields with the
    hand-off of avism_next_level; RMIN_SUB at the median parent radius."""
    from avism_amd import synth
    n, L0 = 32, 64.0
    d, vx, vy, vz = (a.numpy() for a in synth.grid_fields(n, L0, seed))
    f1 = O.fields_from_grid(d, vx, vy, vz, n, L0, per)
    r1 = O.void_stage(f1["grid"], f1["u1co"], f1["div"], PARAMS, meandens=f1["meandens"])
    main = r1.uvoid == -1
    rmin_sub = float(np.median(r1.vf[11][main]))
    P = _oracle_params(rmin_sub)
    r1 = O.void_stage(f1["grid"], f1["u1co"], f1["div"], P, meandens=f1["meandens"])
    marcap, reqp = O.next_level_inputs(f1["grid"], r1.marca, r1.uvoid, r1.vf[11])
    d2, vx2, vy2, vz2 = (a.numpy() for a in synth.grid_fields(2 * n, L0, seed + 1000))
    f2 = O.fields_from_grid(d2, vx2, vy2, vz2, 2 * n, L0, per)
    r2 = O.void_stage(f2["grid"], f2["u1co"], f2["div"], P, marcap=marcap, reqp=reqp, meandens=f2["meandens"])
    print(f"[grid levels per={per}] level 1 voids={r1.nvoidt} level 2 cubes={r2.nvoid} voids={r2.nvoidt}",
          flush=True)
    c = _ctx(n, L0, per, rmin_sub)
    c.load_grid_fields(d, vx, vy, vz)
    compare(gpu_void_outputs(c), lambda kk: _as_ref(r1)[kk])
    c.next_level()
    c.load_grid_fields(d2, vx2, vy2, vz2)
    out2 = gpu_void_outputs(c)
    compare(out2, lambda kk: _as_ref(r2)[kk])
    assert np.array_equal(out2["father"], r2.father)
    assert len(np.unique(r2.father)) >= 2


def test_parent_ids_above_radii_rejected():
    from avism_amd._lib import AvismError
    from avism_amd import synth
    n, L0 = 24, 48.0
    d, vx, vy, vz = (a.numpy() for a in synth.grid_fields(n, L0, 7))
    c = _ctx(n, L0, 1, 1.0)
    c.load_grid_fields(d, vx, vy, vz)
    c.set_parent_map(np.full(n ** 3, 5, np.int32))
    c.set_parent_radii(np.ones(4, np.float32))
    with pytest.raises(AvismError):
        c.find_voids(0)
    c.set_parent_radii(np.ones(5, np.float32))
    c.find_voids(0)


def test_next_level_needs_a_catalogue():
    from avism_amd._lib import AvismError
    c = _ctx(16, 32.0, 0, 1.0)
    with pytest.raises(AvismError):
        c.next_level()
