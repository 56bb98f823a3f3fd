"""Batches of independent patches (avism_batch_*): many uneven, masked,
non-periodic Zel'dovich patches through one field stage of persistent
launches, autotuned tile shapes and a captured graph, then each patch's void
stage on its own context.

Every patch must equal a run of the single-grid path on its particles and
mask: fields bit for bit (the batched kernels run the single-grid kernel
bodies per work item), and each patch's catalogue equals the oracle's void
stage on those fields (MARCA, cubes, UVOID, ids bit-exact; the catalogue's
floating-point columns as in test_gpu_voids).  The graph is replayed on new
particles of the same layout, with eager work and synchronisations between
the replays, and every replay is checked.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O
from test_gpu_voids import PARAMS, compare, gpu_void_outputs

pytestmark = pytest.mark.gpu

FIELDS = ("u1co", "div", "u", "v", "w", "part_dens", "hpart")


def _ctx():
    from avism_amd._lib import Context
    c = Context(0)
    c.set_cosmology(0.678, 0.31, 0.0)
    c.set_void_params(PARAMS.dens_thre, PARAMS.dens_thre2, PARAMS.grad_thre, PARAMS.div_thre, PARAMS.rmin,
                      PARAMS.rmin_sub, PARAMS.nvoid_max)
    return c


def _patches(npatch, seed0):
    from avism_amd import synth
    return synth.patch_batch(npatch, seed0=seed0, np_range=(8e3, 3e5), grids=(16, 24, 32, 40, 48))


def _load(c, pats):
    arr = [np.concatenate([p["particles"].numpy()[i] for p in pats]) for i in range(7)]
    mask = np.concatenate([p["mask"] for p in pats])
    c.batch_load(*arr, smask=mask)


def _single(pat, k=32):
    """The single-grid path on one patch: fields, particle fields, context."""
    c = _ctx()
    c.set_grid(pat["n"], pat["L0"], 0)
    c.set_mask(pat["mask"])
    arr = pat["particles"].numpy()
    c.load_particles(*arr)
    c.particles_to_grid(k=k)
    f = c.fields()
    f["part_dens"], f["hpart"] = c.particle_fields(arr[0].size)
    return f, c


def _layout(c, pats):
    c.batch_set([p["n"] for p in pats], [p["L0"] for p in pats], [p["particles"].n for p in pats])


def test_batch_fields_and_catalogues_match_single_patches():
    pats = _patches(10, 3000)
    c = _ctx()
    _layout(c, pats)
    _load(c, pats)
    print(f"[batch] {len(pats)} patches, grids {[p['n'] for p in pats]}, "
          f"particles {[p['particles'].n for p in pats]}", flush=True)
    c.batch_fields(k=32)
    c.batch_voids(threads=4)
    for i, pat in enumerate(pats):
        fb = c.batch_fields_of(i)
        fs, cs = _single(pat)
        for q in FIELDS:
            assert np.array_equal(fb[q], fs[q]), (i, q)
        sub = c.batch_patch(i)
        out = gpu_void_outputs_from(sub)
        ref = O.void_stage(O.Grid(pat["n"], pat["L0"], 0), fb["u1co"], fb["div"], PARAMS, smask=pat["mask"],
                           meandens=sub.stats().meandens, rete=O.cosmo(0.678, 0.31, 0.0)[1])
        compare(out, lambda k, r=ref: _as_ref(r)[k])
        # the patch's catalogue equals the single-grid context's own
        single = gpu_void_outputs(cs)
        for k in ("marca", "uvoid", "indice", "cubes_i", "ncellv", "vf"):
            assert np.array_equal(out[k], single[k]), (i, k)
    info = c.batch_info()
    t = info["tuning"]
    print(f"  tuning: vvel {t['vvel_shape']}/{t['vvel_blocks']}, ddens {t['ddens_shape']}/{t['ddens_blocks']}, "
          f"ppart {t['ppart_blocks']}", flush=True)
    assert t["vvel_shape"] in ("8x4x4", "4x8x4", "4x4x8")
    assert t["ddens_shape"] in ("8x8x4+4", "8x8x8+3", "16x8x4+4")
    assert t["k"] == 32 and t["patches"] == len(pats)
    assert info["graph_launches"] == 1


def gpu_void_outputs_from(sub):
    """The catalogue of a patch context whose void stage has run."""
    nv = int(sub.stats().ncubes)
    ci, cf, fa = sub.cubes(nv)
    cat = sub.catalogue(nv)
    st = cat["stats"]
    return dict(nvoid=nv, cubes_i=ci, cubes_f=cf, father=fa, indice=cat["indice"], uvoid=cat["uvoid"],
                ncellv=cat["ncellv"], vf=cat["vf"], nvoid2=st.nvoid2, nvoidt=st.nvoidt, volt_clean=st.volt_clean,
                marca=sub.marca(), nperiodic=st.nperiodic)


def _as_ref(res):
    return dict(nvoid=res.nvoid, cubes_i=res.cubes_i, cubes_f=res.cubes_f, indice=res.indice, nvoid2=res.nvoid2,
                uvoid=res.uvoid, marca=res.marca, nvoidt=res.nvoidt, ncellv=res.ncellv, volt_clean=res.volt_clean,
                vf=res.vf)


def test_batch_graph_replays_on_new_particles():
    """Three batches of one layout: the first tunes and captures, the next two
    replay the graph on new particles; each equals the single-grid path."""
    pats0 = _patches(6, 3100)
    c = _ctx()
    _layout(c, pats0)
    for rep in range(3):
        # same layout (grid, box, particle count per patch), new particles
        pats = pats0 if rep == 0 else [dict(p, particles=_reseed(p, 7000 + 100 * rep + i)) for i, p in
                                       enumerate(pats0)]
        _load(c, pats)
        c.batch_fields(k=32)
        print(f"[batch replay {rep}]", flush=True)
        for i, pat in enumerate(pats):
            fb = c.batch_fields_of(i)
            fs, _ = _single(pat)
            for q in FIELDS:
                assert np.array_equal(fb[q], fs[q]), (rep, i, q)
        assert c.batch_info()["graph_launches"] == rep + 1
    c.batch_voids()
    assert c.batch_patch(0).stats().ncubes >= 0


def _reseed(pat, seed):
    from avism_amd import synth
    side = round(pat["particles"].n ** (1.0 / 3.0))
    p = synth.zeldovich(side, pat["L0"], seed=seed)
    return synth.Particles(p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m, pat["L0"], 0)


def test_batch_learned_growth_equals_plain():
    from avism_amd import learned
    from avism_amd._lib import GROW_LEARNED
    pats = _patches(5, 3200)
    c = _ctx()
    _layout(c, pats)
    _load(c, pats)
    c.batch_fields(k=32)
    c.batch_voids()
    plain = [gpu_void_outputs_from(c.batch_patch(i)) for i in range(len(pats))]
    learned.Proposer.load_default().install(c)
    c.batch_voids(growth_mode=GROW_LEARNED)
    for i in range(len(pats)):
        lr = gpu_void_outputs_from(c.batch_patch(i))
        for k in ("marca", "uvoid", "indice", "cubes_i", "ncellv", "vf"):
            assert np.array_equal(lr[k], plain[i][k]), (i, k)


def test_batch_rejects():
    from avism_amd._lib import AvismError
    pats = _patches(2, 3300)
    c = _ctx()
    with pytest.raises(AvismError) as e:
        c.batch_fields(k=32)
    assert e.value.code == -5  # no batch
    _layout(c, pats)
    _load(c, pats)
    with pytest.raises(AvismError) as e:
        c.batch_fields(k=100)
    assert e.value.code == -6  # AVISM_ERR_UNSUPPORTED
    with pytest.raises(AvismError) as e:
        c.batch_patch(0)
    assert e.value.code == -5  # the void stage has not run
    # a particle outside its patch's box
    arr = [np.concatenate([p["particles"].numpy()[i] for p in pats]) for i in range(7)]
    arr[0][5] = pats[0]["L0"]
    with pytest.raises(AvismError) as e:
        c.batch_load(*arr)
    assert e.value.code == -3 and "patch 0" in str(e.value)


def test_batch_64_patches_at_full_size():
    """Configuration 3 at its full size: 64 patches, 2e5..8e6 particles each,
    grids of 32^3..160^3 cells, 20% of each mask empty, one batch.  Every
    patch's fields equal the single-grid path's bit for bit and its catalogue
    is the single-grid context's; the three smallest patches' void stages
    equal the oracle's on the batch's fields."""
    from avism_amd import synth
    pats = synth.patch_batch(64, seed0=3000, device="cpu")
    ns = [p["n"] for p in pats]
    print(f"[batch 64] particles {sum(p['particles'].n for p in pats)}, cells {sum(n ** 3 for n in ns)}",
          flush=True)
    c = _ctx()
    _layout(c, pats)
    _load(c, pats)
    c.batch_fields(k=32)
    c.batch_voids(threads=8)
    small = set(np.argsort([n ** 3 + p["particles"].n for n, p in zip(ns, pats)], kind="stable")[:3].tolist())
    for i, pat in enumerate(pats):
        fb = c.batch_fields_of(i)
        fs, cs = _single(pat)
        for q in FIELDS:
            assert np.array_equal(fb[q], fs[q]), (i, q)
        out = gpu_void_outputs_from(c.batch_patch(i))
        single = gpu_void_outputs(cs)
        for k in ("marca", "uvoid", "indice", "cubes_i", "ncellv", "vf"):
            assert np.array_equal(out[k], single[k]), (i, k)
        if i in small:
            ref = O.void_stage(O.Grid(pat["n"], pat["L0"], 0), fb["u1co"], fb["div"], PARAMS, smask=pat["mask"],
                               meandens=c.batch_patch(i).stats().meandens, rete=O.cosmo(0.678, 0.31, 0.0)[1])
            compare(out, lambda k, r=ref: _as_ref(r)[k])
        del fb, fs, cs
    assert c.batch_info()["graph_launches"] == 1
