"""GPU parity of the field stage against the oracle and the reference fixtures.

Tolerances: neighbour sets, HPART and the DDENS normalisation are bit-exact
by construction (same FP32 operation order, same tie rule); the accumulated
fields (FP64 sums on both sides, in another order; DDENS as int64 fixed point)
are held to max relative error 1e-5 (velocity and divergence errors
normalised by the field RMS).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-5


def _ctx():
    from avism_amd._lib import Context
    return Context(0)


def _rel(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _rms_rel(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b)) / np.sqrt(np.mean(b * b)))


def _gpu_fields(arr, n, L0, per, k, vel_interp=1, dens_interp=1):
    c = _ctx()
    c.set_grid(n, L0, per)
    c.set_cosmology(0.678, 0.31, 0.0)
    c.load_particles(*arr)
    c.particles_to_grid(k=k, vel_interp=vel_interp, dens_interp=dens_interp)
    f = c.fields()
    pd, hp = c.particle_fields(arr[0].size)
    return c, f, pd, hp


@pytest.mark.parametrize("name", ["part_p0", "part_p1", "part_k100_p0", "part_k256_p1"])
def test_fields_vs_reference_fixture(golden_dir, name):
    g = dict(np.load(golden_dir / f"{name}.npz"))
    n, L0, per, k = (int(g["meta_n"]), float(g["meta_L0"]), int(g["meta_periodic"]), int(g["meta_k"]))
    arr = [g[f"in_{c}"] for c in ("x", "y", "z", "vx", "vy", "vz", "m")]
    print(f"[fields fixture {name}] n={n} np={arr[0].size} k={k}", flush=True)
    _, f, pd, hp = _gpu_fields(arr, n, L0, per, k)
    assert np.array_equal(hp, g["ref_hpart"])
    assert _rel(pd, g["ref_part_dens"]) <= REL
    for c in ("u", "v", "w"):
        assert _rms_rel(f[c], g[f"ref_{c}"]) <= REL, c
    assert _rel(f["u1co"], g["ref_u1co"]) <= REL
    assert _rms_rel(f["div"], g["ref_div"]) <= REL


@pytest.mark.parametrize("case", ["zel32", "mock32"])
def test_fields_vs_oracle(case):
    import pyoracle as O
    from avism_amd import synth
    if case == "zel32":
        n, L0, per = 32, 62.5, 1
        p = synth.zeldovich(32, L0, seed=31)
    else:
        n, L0, per = 32, 100.0, 0
        p, _, _ = synth.mock_voids(n0=40000, L0=L0, nvoids=5, rmin=8, rmax=15, seed=32)
    arr = p.numpy()
    print(f"[fields oracle {case}] n={n} np={arr[0].size}", flush=True)
    ref = O.fields_from_particles(*arr, n, L0, per, k=32)
    _, f, pd, hp = _gpu_fields(arr, n, L0, per, 32)
    assert np.array_equal(hp, ref["hpart"])
    assert _rel(pd, ref["part_dens"]) <= REL
    for c in ("u", "v", "w"):
        assert _rms_rel(f[c], ref[c]) <= REL, c
    assert _rel(f["u1co"], ref["u1co"]) <= REL
    assert _rms_rel(f["div"], ref["div"]) <= REL


@pytest.mark.parametrize("per", [0, 1])
def test_fields_mixed_masses(per):
    """Particles of different masses: PPART sums the k nearest masses one by
    one (with equal masses the library takes k * m, which is the same FP64
    value); checked against the oracle."""
    import pyoracle as O
    from avism_amd import synth
    n, L0 = 24, 48.0
    p = synth.zeldovich(n, L0, seed=37)
    arr = list(p.numpy())
    rng = np.random.default_rng(38)
    arr[6] = (arr[6] * rng.uniform(0.5, 1.5, arr[6].size)).astype(np.float32)
    if not per:  # keep the particles inside the box without wrapping
        for a in arr[:3]:
            np.clip(a, -L0 / 2, L0 / 2 - 1e-4, out=a)
    print(f"[fields mixed masses periodic={per}]", flush=True)
    ref = O.fields_from_particles(*arr, n, L0, per, k=32)
    _, f, pd, hp = _gpu_fields(arr, n, L0, per, 32)
    assert np.array_equal(hp, ref["hpart"])
    assert _rel(pd, ref["part_dens"]) <= REL
    for c in ("u", "v", "w"):
        assert _rms_rel(f[c], ref[c]) <= REL, c
    assert _rel(f["u1co"], ref["u1co"]) <= REL


@pytest.mark.parametrize("name", ["tsc_p0", "tsc_p1", "tscv_p0", "tscv_p1"])
def test_tsc_vs_reference_fixture(golden_dir, name):
    """VEL_INTERP_TSC / DENS_INTERP_TSC / H_DISTANCE against the reference's own
    output (one OpenMP thread).  The reference sums the TSC velocities in FP32
    under OMP ATOMIC, so its last bits depend on the schedule; here the sums are
    exact (fixed point): velocities within 1e-5 of the field RMS."""
    g = dict(np.load(golden_dir / f"{name}.npz"))
    n, L0, per, k = (int(g["meta_n"]), float(g["meta_L0"]), int(g["meta_periodic"]), int(g["meta_k"]))
    vi, di = int(g["meta_vel_interp"]), int(g["meta_dens_interp"])
    arr = [g[f"in_{c}"] for c in ("x", "y", "z", "vx", "vy", "vz", "m")]
    print(f"[tsc fixture {name}] n={n} np={arr[0].size} vel={vi} dens={di}", flush=True)
    _, f, pd, hp = _gpu_fields(arr, n, L0, per, k, vi, di)
    assert not np.any(pd)  # PART_DENS is not computed without SPH velocities
    if di == 1:
        assert np.array_equal(hp, g["ref_hpart"])
    else:
        assert not np.any(hp)
    for c in ("u", "v", "w"):
        assert _rms_rel(f[c], g[f"ref_{c}"]) <= REL, c
    assert _rel(f["u1co"], g["ref_u1co"]) <= REL
    assert _rms_rel(f["div"], g["ref_div"]) <= REL


@pytest.mark.parametrize("k,per,vi,di", [(33, 0, 1, 1), (65, 1, 1, 1), (128, 0, 1, 1), (256, 1, 1, 1),
                                         (256, 1, 0, 1), (200, 0, 0, 1)])
def test_fields_large_k_vs_oracle(k, per, vi, di):
    """Each neighbour-count class (k up to 64, 128, 256 selection slots; the
    reference's default KNEIGHBOURS is 256) on SPH velocities and on
    H_DISTANCE with TSC velocities, at 24^3 with 13824 particles."""
    import pyoracle as O
    from avism_amd import synth
    n, L0 = 24, 48.0
    p = synth.zeldovich(24, L0, seed=70 + k)
    arr = list(p.numpy())
    if not per:
        keep = np.all(np.abs(np.stack(arr[:3])) < L0 / 2 - 2 * L0 / n, axis=0)
        arr = [a[keep].copy() for a in arr]
    print(f"[fields k={k} per={per} vel={vi} dens={di}] np={arr[0].size}", flush=True)
    ref = O.fields_from_particles(*arr, n, L0, per, k=k, vel_interp=vi, dens_interp=di, tsc_exact=True)
    _, f, pd, hp = _gpu_fields(arr, n, L0, per, k, vi, di)
    assert np.array_equal(hp, ref["hpart"])
    if vi == 1:
        assert _rel(pd, ref["part_dens"]) <= REL
    for c in ("u", "v", "w"):
        assert _rms_rel(f[c], ref[c]) <= REL, c
    assert _rel(f["u1co"], ref["u1co"]) <= REL
    assert _rms_rel(f["div"], ref["div"]) <= REL


@pytest.mark.parametrize("name", ["nov_p0", "nov_p1", "novt_p1"])
def test_velocity_free_vs_reference_fixture(golden_dir, name):
    """avism_particles_to_density on particles without velocities (the
    reference's FLAG_VEL_AVAILABLE = 0 run): HPART bit-exact, velocities 0,
    U1CO within 1e-5, DIVERCO = -OMEGA0^0.6 RETE HTE (U1CO - 1) within 1e-5 of
    its RMS."""
    g = dict(np.load(golden_dir / f"{name}.npz"))
    n, L0, per, k = (int(g["meta_n"]), float(g["meta_L0"]), int(g["meta_periodic"]), int(g["meta_k"]))
    di = int(g["meta_dens_interp"])
    print(f"[velocity-free fields {name}] n={n} dens_interp={di}", flush=True)
    c = _ctx()
    c.set_grid(n, L0, per)
    c.set_cosmology(0.678, 0.31, 0.0)
    c.load_particles(g["in_x"], g["in_y"], g["in_z"], None, None, None, g["in_m"])
    c.particles_to_density(k=k, dens_interp=di)
    f = c.fields()
    _, hp = c.particle_fields(g["in_x"].size)
    assert np.array_equal(hp, g["ref_hpart"])
    for comp in ("u", "v", "w"):
        assert not np.any(f[comp]), comp
    assert _rel(f["u1co"], g["ref_u1co"]) <= REL
    assert _rms_rel(f["div"], g["ref_div"]) <= REL


@pytest.mark.parametrize("per,di", [(1, 1), (0, 1), (1, 0)])
def test_velocity_free_vs_oracle(per, di):
    """The velocity-free field stage at 32^3 against the oracle, then the void
    stage on the GPU's fields against the oracle's on the same fields."""
    import pyoracle as O
    from avism_amd import synth
    from test_gpu_voids import PARAMS, compare, gpu_void_outputs
    n, L0 = 32, 62.5
    arr = list(synth.zeldovich(32, L0, seed=90 + per).numpy())
    if not per:
        keep = np.all(np.abs(np.stack(arr[:3])) < L0 / 2 - 2 * L0 / n, axis=0)
        arr = [a[keep].copy() for a in arr]
    x, y, z, m = arr[0], arr[1], arr[2], arr[6]
    print(f"[velocity-free oracle per={per} dens={di}] np={x.size}", flush=True)
    ref = O.fields_from_particles(x, y, z, None, None, None, m, n, L0, per, k=32, vel_interp=-1, dens_interp=di)
    c = _ctx()
    c.set_grid(n, L0, per)
    c.set_cosmology(0.678, 0.31, 0.0)
    c.set_void_params(PARAMS.dens_thre, PARAMS.dens_thre2, PARAMS.grad_thre, PARAMS.div_thre, PARAMS.rmin,
                      PARAMS.rmin_sub, PARAMS.nvoid_max)
    c.load_particles(x, y, z, None, None, None, m)
    c.particles_to_density(k=32, dens_interp=di)
    f = c.fields()
    _, hp = c.particle_fields(x.size)
    assert np.array_equal(hp, ref["hpart"])
    assert _rel(f["u1co"], ref["u1co"]) <= REL
    assert _rms_rel(f["div"], ref["div"]) <= REL
    out = gpu_void_outputs(c)
    res = O.void_stage(O.Grid(n, L0, per), f["u1co"], f["div"], PARAMS, meandens=c.stats().meandens,
                       rete=O.cosmo(0.678, 0.31, 0.0)[1])
    assert res.nvoidt > 0
    compare(out, lambda k: dict(nvoid=res.nvoid, cubes_i=res.cubes_i, cubes_f=res.cubes_f, indice=res.indice,
                                nvoid2=res.nvoid2, uvoid=res.uvoid, marca=res.marca, nvoidt=res.nvoidt,
                                ncellv=res.ncellv, volt_clean=res.volt_clean, vf=res.vf)[k])


def test_k_above_256_rejected():
    from avism_amd._lib import AvismError
    from avism_amd import synth
    arr = synth.zeldovich(8, 16.0, seed=5).numpy()
    c = _ctx()
    c.set_grid(8, 16.0, 1)
    c.set_cosmology(0.678, 0.31, 0.0)
    c.load_particles(*arr)
    with pytest.raises(AvismError):
        c.particles_to_grid(k=257)


@pytest.mark.parametrize("per,vi,di", [(1, 0, 0), (0, 0, 0), (1, 0, 1), (1, 1, 0)])
def test_tsc_vs_oracle(per, vi, di):
    """The TSC paths at a larger size and every flag combination, against the oracle."""
    import pyoracle as O
    from avism_amd import synth
    n, L0 = 32, 62.5
    p = synth.zeldovich(32, L0, seed=61 + per)
    arr = list(p.numpy())
    if not per:  # the reference drops no deposit only away from the faces
        keep = np.all(np.abs(np.stack(arr[:3])) < L0 / 2 - 2 * L0 / n, axis=0)
        arr = [a[keep].copy() for a in arr]
    print(f"[tsc oracle per={per} vel={vi} dens={di}] np={arr[0].size}", flush=True)
    # the oracle's exact (FP64) TSC velocity sums: the reference's REAL*4 sums
    # round at 1e-4 of the RMS here, in cells reached only by weight tails
    ref = O.fields_from_particles(*arr, n, L0, per, k=32, vel_interp=vi, dens_interp=di, tsc_exact=True)
    _, f, pd, hp = _gpu_fields(arr, n, L0, per, 32, vi, di)
    if di == 1 or vi == 1:
        assert np.array_equal(hp, ref["hpart"])
    for c in ("u", "v", "w"):
        assert _rms_rel(f[c], ref[c]) <= REL, c
    assert _rel(f["u1co"], ref["u1co"]) <= REL
    assert _rms_rel(f["div"], ref["div"]) <= REL


def test_fields_deterministic():
    from avism_amd import synth
    p = synth.zeldovich(24, 48.0, seed=33)
    arr = p.numpy()
    _, f1, pd1, hp1 = _gpu_fields(arr, 24, 48.0, 1, 32)
    _, f2, pd2, hp2 = _gpu_fields(arr, 24, 48.0, 1, 32)
    for key in f1:
        assert np.array_equal(f1[key], f2[key]), key
    assert np.array_equal(pd1, pd2) and np.array_equal(hp1, hp2)


def test_grid_input_fields_exact(golden_dir):
    g = dict(np.load(golden_dir / "grid_p1_n24.npz"))
    n, L0, per = int(g["meta_n"]), float(g["meta_L0"]), int(g["meta_periodic"])
    c = _ctx()
    c.set_grid(n, L0, per)
    c.load_grid_fields(g["in_delta"], g["in_vx"], g["in_vy"], g["in_vz"])
    f = c.fields()
    assert np.array_equal(f["u1co"], g["ref_u1co"])
    assert np.array_equal(f["div"], g["ref_div"])


def test_particles_outside_box_rejected():
    from avism_amd._lib import AvismError
    c = _ctx()
    c.set_grid(8, 10.0, 0)
    x = np.array([0.0, 6.0, 1.0], np.float32)
    o = np.zeros(3, np.float32)
    with pytest.raises(AvismError) as e:
        c.load_particles(x, o, o, o, o, o, np.ones(3, np.float32))
    assert e.value.code == -3


@pytest.mark.parametrize("mixed", [0, 1])
@pytest.mark.parametrize("per", [0, 1])
def test_fields_lattice_ties(per, mixed):
    """Particles on the cell centres of a perfect lattice: every query has many
    neighbours at exactly equal distances, so the tie rule (lower input index
    wins) decides the neighbour sets; HPART must still match bit for bit.
    With unequal masses PART_DENS sums the masses of exactly those sets, so a
    wrong member among the ties shows in it."""
    import pyoracle as O
    n, L0 = 16, 32.0
    d = L0 / n
    g = (np.arange(n, dtype=np.float32) + 0.5) * d - L0 / 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    rng = np.random.default_rng(41)
    perm = rng.permutation(n ** 3)  # input order unrelated to the bin order
    arr = [a.reshape(-1)[perm].astype(np.float32) for a in (x, y, z)]
    arr += [rng.standard_normal(n ** 3).astype(np.float32) * 100 for _ in range(3)]
    arr += [np.full(n ** 3, 1e10, np.float32)]
    if mixed:
        arr[6] = (arr[6] * rng.uniform(0.5, 1.5, n ** 3)).astype(np.float32)
    print(f"[fields lattice periodic={per} mixed={mixed}]", flush=True)
    ref = O.fields_from_particles(*arr, n, L0, per, k=32)
    _, f, pd, hp = _gpu_fields(arr, n, L0, per, 32)
    assert np.array_equal(hp, ref["hpart"])
    assert _rel(pd, ref["part_dens"]) <= REL
    for c in ("u", "v", "w"):
        assert _rms_rel(f[c], ref[c]) <= REL, c
    assert _rel(f["u1co"], ref["u1co"]) <= REL


@pytest.mark.parametrize("per", [0, 1])
def test_fields_clustered(per):
    """Dense blobs next to empty space: the blob tiles overflow the LDS
    staging (HBM fallback path), blob cells take the DX ball path (d_k <= DX)
    and the empty regions need wide searches."""
    import pyoracle as O
    n, L0 = 40, 80.0
    rng = np.random.default_rng(55 + per)
    nbg, nblob = 30000, 40000
    bg = rng.uniform(-L0 / 2, L0 / 2, (nbg, 3))
    centres = rng.uniform(-L0 / 3, L0 / 3, (4, 3))
    blob = centres[rng.integers(0, 4, nblob)] + rng.standard_normal((nblob, 3)) * 1.2
    xyz = np.clip(np.concatenate([bg, blob]), -L0 / 2, L0 / 2 - 1e-4).astype(np.float32)
    npart = xyz.shape[0]
    arr = [xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()]
    arr += [rng.standard_normal(npart).astype(np.float32) * 200 for _ in range(3)]
    arr += [np.full(npart, 2e11, np.float32)]
    print(f"[fields clustered periodic={per}] np={npart}", flush=True)
    ref = O.fields_from_particles(*arr, n, L0, per, k=32)
    _, f, pd, hp = _gpu_fields(arr, n, L0, per, 32)
    assert np.array_equal(hp, ref["hpart"])
    assert _rel(pd, ref["part_dens"]) <= REL
    for c in ("u", "v", "w"):
        assert _rms_rel(f[c], ref[c]) <= REL, c
    assert _rel(f["u1co"], ref["u1co"]) <= REL


@pytest.mark.parametrize("per", [0, 1])
def test_fields_mostly_empty_box(per):
    """Particles only in one corner of the box: the bin starts are runs of
    thousands of empty bins (the long-run path of the binning), and the
    searches from the empty cells reach across the box."""
    import pyoracle as O
    n, L0 = 48, 96.0
    rng = np.random.default_rng(71 + per)
    npart = 6000
    xyz = rng.uniform(-L0 / 2, -L0 / 2 + L0 / 5, (npart, 3)).astype(np.float32)
    arr = [xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()]
    arr += [rng.standard_normal(npart).astype(np.float32) * 200 for _ in range(3)]
    arr += [np.full(npart, 2e11, np.float32)]
    print(f"[fields mostly empty periodic={per}] np={npart}", flush=True)
    ref = O.fields_from_particles(*arr, n, L0, per, k=32)
    _, f, pd, hp = _gpu_fields(arr, n, L0, per, 32)
    assert np.array_equal(hp, ref["hpart"])
    assert _rel(pd, ref["part_dens"]) <= REL
    for c in ("u", "v", "w"):
        assert _rms_rel(f[c], ref[c]) <= REL, c
    # the int64 fixed point of the cell masses resolves 2^-62 n^3 of the mean
    # density (2.4e-14 here, one rounding per deposit): cells reached only by
    # weight tails (U1CO down to 1e-20 in the oracle's FP64) are held to that
    # absolute resolution, the others to the relative bar
    u, r = f["u1co"].astype(np.float64), ref["u1co"].astype(np.float64)
    big = r > 1e-5
    assert _rel(u[big], r[big]) <= REL
    assert np.max(np.abs(u[~big] - r[~big]), initial=0.0) <= 1e-10


def test_device_inputs_from_another_stream():
    """Particles written on a torch side stream and handed over as device
    pointers without a synchronize: the library waits for them (the fields
    must equal those of a synchronized load)."""
    import torch
    from avism_amd import synth
    from avism_amd._lib import AVISM_DEVICE
    p = synth.zeldovich(24, 48.0, seed=35)
    ref_c = _ctx()
    ref_c.set_grid(24, 48.0, 1)
    ref_c.load_particles(*p.numpy())
    ref_c.particles_to_grid(k=32)
    ref = ref_c.fields()
    side = torch.cuda.Stream()
    host = [torch.from_numpy(a) for a in p.numpy()]
    with torch.cuda.stream(side):
        slow = torch.ones(1 << 26, device="cuda")
        for _ in range(20):
            slow = slow * 1.0000001  # keeps the side stream busy before the copies
        dev = [h.to("cuda", non_blocking=True) for h in host]
    c = _ctx()
    c.set_grid(24, 48.0, 1)
    c.load_particles(*dev, where=AVISM_DEVICE)
    c.particles_to_grid(k=32)
    f = c.fields()
    torch.cuda.synchronize()
    for key in ("u1co", "div", "u"):
        assert np.array_equal(f[key], ref[key]), key


@pytest.mark.parametrize("n,k,per,vi", [(20, 128, 1, 0), (20, 32, 1, 1), (18, 64, 1, 1), (18, 32, 0, 1),
                                        (13, 128, 0, 0), (13, 16, 1, 1)])
def test_hpart_grid_not_multiple_of_tiles(n, k, per, vi):
    """Grids that are not multiples of the kernels' tile sizes (8 x 4 x 4 cells
    for the cell queries, 8 x 8 x 4 for DDENS), with the periodic wrap across
    partial tiles: HPART bit-exact against the oracle and U1CO within 1e-5,
    for SPH velocities and for H_DISTANCE."""
    import pyoracle as O
    from avism_amd import synth
    L0 = 3.0 * n
    p = synth.zeldovich(n, L0, seed=400 + n + k)
    arr = list(p.numpy())
    if per:  # denser: 2 particles per cell from a second lattice
        q = synth.zeldovich(n, L0, seed=500 + n + k)
        arr = [np.concatenate([a, b]) for a, b in zip(arr, q.numpy())]
    else:
        keep = np.all(np.abs(np.stack(arr[:3])) < L0 / 2 - 2 * L0 / n, axis=0)
        arr = [a[keep].copy() for a in arr]
    print(f"[hpart n={n} k={k} per={per} vel={vi}] np={arr[0].size}", flush=True)
    ref = O.fields_from_particles(*arr, n, L0, per, k=k, vel_interp=vi, dens_interp=1, tsc_exact=True)
    _, f, _, hp = _gpu_fields(arr, n, L0, per, k, vi, 1)
    assert np.array_equal(hp, ref["hpart"])
    assert _rel(f["u1co"], ref["u1co"]) <= REL
