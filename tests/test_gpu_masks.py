"""GPU parity with a survey mask (SMASK, avism.f90:403-418, 866-904): VVEL and
H_DISTANCE skip masked cells (particles.f90:312, 856), MARK_ALL marks nothing
there and the flag stencil ends a cell at a masked neighbour
(voidfinding.f90:33, 283-300).  Bars as in test_gpu_fields / test_gpu_voids."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O
from test_gpu_fields import _rel, _rms_rel
from test_gpu_voids import PARAMS, compare, gpu_void_outputs

pytestmark = pytest.mark.gpu

REL = 1e-5


def _ctx(n, L0, per):
    from avism_amd._lib import Context
    c = Context(0)
    c.set_grid(n, L0, per)
    c.set_cosmology(0.678, 0.31, 0.0)
    c.set_void_params(PARAMS.dens_thre, PARAMS.dens_thre2, PARAMS.grad_thre, PARAMS.div_thre, PARAMS.rmin,
                      PARAMS.rmin_sub, PARAMS.nvoid_max)
    return c


def _as_ref(res):
    return dict(nvoid=res.nvoid, cubes_i=res.cubes_i, cubes_f=res.cubes_f, indice=res.indice, nvoid2=res.nvoid2,
                uvoid=res.uvoid, marca=res.marca, nvoidt=res.nvoidt, ncellv=res.ncellv, volt_clean=res.volt_clean,
                vf=res.vf)


@pytest.mark.parametrize("name", ["mask_p0", "mask_p1"])
def test_masked_fields_vs_reference(golden_dir, name):
    g = dict(np.load(golden_dir / f"{name}.npz"))
    n, L0, per, k = (int(g["meta_n"]), float(g["meta_L0"]), int(g["meta_periodic"]), int(g["meta_k"]))
    arr = [g[f"in_{c}"] for c in ("x", "y", "z", "vx", "vy", "vz", "m")]
    sm = g["in_smask"]
    print(f"[masked fields {name}] n={n} masked={int((sm == 0).sum())}", flush=True)
    c = _ctx(n, L0, per)
    c.set_mask(sm)
    c.load_particles(*arr)
    c.particles_to_grid(k=k)
    f = c.fields()
    _, hp = c.particle_fields(arr[0].size)
    assert np.array_equal(hp, g["ref_hpart"])
    for comp in ("u", "v", "w"):
        assert np.all(f[comp][sm == 0] == 0.0), comp
        assert _rms_rel(f[comp], g[f"ref_{comp}"]) <= REL, comp
    assert _rel(f["u1co"], g["ref_u1co"]) <= REL
    assert _rms_rel(f["div"], g["ref_div"]) <= REL


@pytest.mark.parametrize("name", ["mask_p0", "mask_p1", "mask_g0", "mask_g1"])
def test_masked_void_stage_vs_reference(golden_dir, name):
    g = dict(np.load(golden_dir / f"{name}.npz"))
    n, L0, per = int(g["meta_n"]), float(g["meta_L0"]), int(g["meta_periodic"])
    print(f"[masked voids {name}]", flush=True)
    c = _ctx(n, L0, per)
    c.set_mask(g["in_smask"])
    if int(g["meta_mode"]) == 2:
        c.load_grid_fields(g["in_delta"], g["in_vx"], g["in_vy"], g["in_vz"])
    else:
        c.load_void_fields(g["ref_u1co"], g["ref_div"], float(g["ref_meandens"]))
    compare(gpu_void_outputs(c), lambda k: g[f"ref_{k}"])


@pytest.mark.parametrize("n,per,seed", [(48, 1, 401), (48, 0, 402)])
def test_masked_grid_vs_oracle_and_learned(n, per, seed):
    """Masked grid input at 48^3 against the oracle; proposals wrong both ways
    give the same catalogue; a second level upsamples the mask."""
    from avism_amd import synth
    L0 = 2.0 * n
    d, vx, vy, vz = (a.numpy() for a in synth.grid_fields(n, L0, seed))
    sm = synth.survey_mask(n, 0.2, seed=seed)
    f = O.fields_from_grid(d, vx, vy, vz, n, L0, per)
    ref = O.void_stage(f["grid"], f["u1co"], f["div"], PARAMS, smask=sm, meandens=f["meandens"])
    print(f"[masked grid n={n} per={per}] cubes={ref.nvoid} voids={ref.nvoidt}", flush=True)
    c = _ctx(n, L0, per)
    c.set_mask(sm)
    c.load_grid_fields(d, vx, vy, vz)
    plain = gpu_void_outputs(c)
    compare(plain, lambda k: _as_ref(ref)[k])
    lin, r = c.candidates()
    rng = np.random.default_rng(seed)
    c.set_proposals(np.clip(np.maximum(r, 0) + rng.integers(-2, 3, size=r.size), 0, 100).astype(np.int8))
    learned = gpu_void_outputs(c, 1)
    for k in ("marca", "uvoid", "indice", "cubes_i"):
        assert np.array_equal(learned[k], plain[k]), k
    # second level: MARCAP, REQP and the mask at 2N
    c2 = _ctx(n, L0, per)
    c2.set_mask(sm)
    c2.load_grid_fields(d, vx, vy, vz)
    gpu_void_outputs(c2)
    marcap, reqp = O.next_level_inputs(f["grid"], ref.marca, ref.uvoid, ref.vf[11])
    c2.next_level()
    d2, vx2, vy2, vz2 = (a.numpy() for a in synth.grid_fields(2 * n, L0, seed + 1000))
    c2.load_grid_fields(d2, vx2, vy2, vz2)
    f2 = O.fields_from_grid(d2, vx2, vy2, vz2, 2 * n, L0, per)
    sm2 = np.ascontiguousarray(sm.reshape((n,) * 3).repeat(2, 0).repeat(2, 1).repeat(2, 2)).ravel()
    ref2 = O.void_stage(f2["grid"], f2["u1co"], f2["div"], PARAMS, marcap=marcap, reqp=reqp, smask=sm2,
                        meandens=f2["meandens"])
    out2 = gpu_void_outputs(c2)
    compare(out2, lambda k: _as_ref(ref2)[k])


def test_uneven_masked_patches():
    """Survey-like patches of uneven size through one context (grids 32^3 to
    64^3, Zel'dovich particles at 0.5 to 2 per cell, 20% masked,
    non-periodic): each patch's void stage on its GPU fields equals the
    oracle's, and masked cells carry no velocity."""
    from avism_amd import synth
    from avism_amd._lib import Context
    rng = np.random.default_rng(3000)
    ctx = Context(0)  # one context, reused across patches of different sizes
    for i in range(6):
        n = int(rng.choice([32, 40, 48, 64]))
        side = int(round(n * rng.uniform(0.8, 1.26)))
        L0 = 2.0 * n
        p = synth.zeldovich(side, L0, seed=3000 + i)
        arr = p.numpy()
        sm = synth.survey_mask(n, 0.2, seed=3000 + i)
        ctx.set_grid(n, L0, 0)
        ctx.set_cosmology(0.678, 0.31, 0.0)
        ctx.set_void_params(PARAMS.dens_thre, PARAMS.dens_thre2, PARAMS.grad_thre, PARAMS.div_thre, PARAMS.rmin,
                            PARAMS.rmin_sub, PARAMS.nvoid_max)
        ctx.set_mask(sm)
        ctx.load_particles(*arr)
        ctx.particles_to_grid(k=32)
        f = ctx.fields()
        out = gpu_void_outputs(ctx)
        ref = O.void_stage(O.Grid(n, L0, 0), f["u1co"], f["div"], PARAMS, smask=sm, meandens=ctx.stats().meandens,
                           rete=O.cosmo(0.678, 0.31, 0.0)[1])
        print(f"[patch {i}] n={n} np={side ** 3} cubes={out['nvoid']} voids={out['nvoidt']}", flush=True)
        compare(out, lambda k: _as_ref(ref)[k])
        assert np.all(f["u"][sm == 0] == 0.0)


@pytest.mark.parametrize("name", ["svy_p0", "svy_p1"])
def test_survey_vs_reference_fixture(golden_dir, name):
    """avism_survey_density against the reference's survey-mode run: galaxy
    HPART bit-exact, U1CO within 1e-5, DIVERCO within 1e-5 of its RMS, MEANDENS
    the background; the void stage on the GPU's fields equals the oracle's, and
    on the reference's fields the reference's catalogue."""
    g = dict(np.load(golden_dir / f"{name}.npz"))
    n, L0, per, k = int(g["meta_n"]), float(g["meta_L0"]), int(g["meta_periodic"]), int(g["meta_k"])
    sm = g["in_smask"]
    print(f"[survey {name}] n={n} randoms={g['in_rx'].size}", flush=True)
    c = _ctx(n, L0, per)
    c.set_mask(sm)
    c.load_particles(g["in_x"], g["in_y"], g["in_z"], None, None, None, g["in_m"])
    c.load_randoms(g["in_rx"], g["in_ry"], g["in_rz"], g["in_rm"])
    c.survey_density(k=k, dens_interp=int(g["meta_dens_interp"]))
    f = c.fields()
    _, hp = c.particle_fields(g["in_x"].size)
    assert np.array_equal(hp, g["ref_hpart"])
    assert _rel(f["u1co"], g["ref_u1co"]) <= REL
    assert _rms_rel(f["div"], g["ref_div"]) <= REL
    assert np.float32(c.stats().meandens) == np.float32(g["ref_meandens"])
    out = gpu_void_outputs(c)
    ref = O.void_stage(O.Grid(n, L0, per), f["u1co"], f["div"], PARAMS, smask=sm, meandens=c.stats().meandens,
                       rete=O.cosmo(0.678, 0.31, 0.0)[1])
    compare(out, lambda kk: _as_ref(ref)[kk])
    c2 = _ctx(n, L0, per)
    c2.set_mask(sm)
    c2.load_void_fields(g["ref_u1co"], g["ref_div"], float(g["ref_meandens"]))
    compare(gpu_void_outputs(c2), lambda kk: g[f"ref_{kk}"])
