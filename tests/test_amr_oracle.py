"""MASCLET AMR input on the CPU: the oracle's gas levels (SMOOTH_2,
DIVER_UNIFORM, DIVER_FINA_GAS, VMESH) against the reference's routines, and
the snapshot reader/writer (avism_amd/masclet.py) against the reference's
READ_MASCLET.

* Fixtures amr_*.npz (tests/golden/make_golden.py, the reference's routines):
  the oracle reproduces U1GCO and DIVERGCO bit for bit.
* Where oracle/_ref is built, fresh runs of the reference on new seeds.
* Closed forms: uniform gas and a linear velocity field.  A refined cell
  covered by c patches holds c (1 + delta) / (1 + c), VMESH's count starting at 1
  (amr2uniform.f90:162, 213-219), unrefined cells the base grid's value, and
  the divergence of v = a x is a everywhere, halved in the refined cells.
"""
from __future__ import annotations

import tempfile
from pathlib import Path

import numpy as np
import pytest

import pyoracle as O
from avism_amd import masclet, synth

CASES = ["amr_p0_ir1", "amr_p1_ir1", "amr_p1_ir0", "amr_p0_irm1"]


@pytest.mark.parametrize("name", CASES)
def test_amr_oracle_vs_reference_fixture(golden_dir, name):
    g = dict(np.load(golden_dir / f"{name}.npz"))
    base = [g["in_delta"], g["in_vx"], g["in_vy"], g["in_vz"]]
    u, d = O.amr_gas(int(g["meta_nb"]), float(g["meta_L0"]), int(g["meta_periodic"]), int(g["meta_ir"]), base,
                     g["in_pint"], g["in_preal"], g["in_pdata"])
    assert np.array_equal(u, g["ref_u1gco"])
    assert np.array_equal(d, g["ref_divgco"])


@pytest.mark.skipif(not O.ref_available(0), reason="oracle/_ref not built")
@pytest.mark.parametrize("per,ir,seed", [(0, 1, 71), (1, -1, 72), (0, -2, 73)])
def test_amr_oracle_vs_reference_fresh(per, ir, seed):
    nb, L0 = 32, 64.0
    base, patches = synth.amr_gas(nb, L0, seed, npatch=20)
    pint, preal, pdata = synth.amr_pack(patches)
    u, d = O.amr_gas(nb, L0, per, ir, base, pint, preal, pdata)
    with tempfile.TemporaryDirectory() as td:
        ru, rd = O.run_reference_amr(td, nb, L0, per, ir, base, pint, preal, pdata)
    assert np.array_equal(u, ru) and np.array_equal(d, rd)


def test_amr_closed_forms():
    nb, L0, a = 16, 32.0, 1e-3
    g0 = O.Grid(nb, L0, 0)
    x0 = np.array([O.lib().orc_grid_rad(g0._p, i) for i in range(1, nb + 1)], np.float32)
    zero = np.zeros((nb, nb, nb), np.float32)
    vx = np.broadcast_to(a * x0[None, None, :], (nb, nb, nb)).astype(np.float32)
    base = [zero.ravel(), vx.ravel(), zero.ravel(), zero.ravel()]
    # two overlapping patches of 8^3 fine cells
    _, patches = synth.amr_gas(nb, L0, 5, npatch=2, nmin=8, nmax=8)
    patches[0]["x"], patches[1]["x"] = (4, 4, 4), (6, 5, 4)
    gl = O.Grid(2 * nb, L0, 0)
    rx1, dxl = O.lib().orc_grid_rad(gl._p, 1), gl.dx
    xf = np.array([O.lib().orc_grid_rad(gl._p, i) for i in range(1, 2 * nb + 1)], np.float32)
    for p in patches:
        p["r"] = tuple(float(np.float32(rx1 + np.float32(2 * v - 2) * dxl + np.float32(0.25) * dxl)) for v in p["x"])
        i0 = [2 * v - 2 for v in p["x"]]
        vxp = np.broadcast_to(a * xf[None, None, i0[0]:i0[0] + 8], (8, 8, 8)).astype(np.float32)
        z8 = np.zeros(512, np.float32)
        p["fields"] = [z8, vxp.ravel(), z8, z8]
    pint, preal, pdata = synth.amr_pack(patches)
    u, d = O.amr_gas(nb, L0, 0, 1, base, pint, preal, pdata)
    u = u.reshape(2 * nb, 2 * nb, 2 * nb)
    d = d.reshape(2 * nb, 2 * nb, 2 * nb)
    cover = np.zeros(u.shape, np.int32)
    for p in patches:
        i0 = [2 * v - 2 for v in p["x"]]
        cover[i0[2]:i0[2] + 8, i0[1]:i0[1] + 8, i0[0]:i0[0] + 8] += 1
    refined = cover > 0
    c = cover[refined].astype(np.float32)
    assert np.array_equal(u[refined], c / (np.float32(1.0) + c))
    assert np.all(u[~refined] == 1.0)
    np.testing.assert_allclose(d[~refined], a, rtol=1e-4)
    np.testing.assert_allclose(d[refined], a * cover[refined] / (1.0 + cover[refined]), rtol=1e-4)


@pytest.mark.skipif(not O.ref_available(0), reason="oracle/_ref not built")
def test_masclet_files_read_by_the_reference():
    nb, L0 = 16, 40.0
    base, patches = synth.amr_gas(nb, L0, 9, npatch=6, nmin=4, nmax=8)
    rng = np.random.default_rng(4)
    dm0 = tuple(rng.uniform(-1, 1, 400).astype(np.float32) for _ in range(6))
    lv = [tuple(rng.uniform(-1, 1, 50 + 10 * i).astype(np.float32) for _ in range(7)) for i in range(2)]
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        masclet.write(td / "simu_masclet", 12, nb, base, patches, dm0, lv, map_mass=0.5)
        mine = masclet.read(td / "simu_masclet", 12, nb)
        ref = O.run_reference_read_masclet(td, nb, 12, npalev=16, nlevels=4, npmax=10_000)
    assert ref["nl"] == mine["nl"] == 2
    assert np.array_equal(ref["gas"][0], mine["base"][0] + np.float32(1.0))
    for q in range(1, 4):
        assert np.array_equal(ref["gas"][q], mine["base"][q])
    assert len(ref["patches"]) == len(mine["patches"])
    for (ints, reals, f), p in zip(ref["patches"], mine["patches"]):
        assert list(ints) == list(p["n"]) + list(p["x"])
        assert np.array_equal(reals, np.asarray(p["r"], np.float32))
        assert np.array_equal(f[0], p["fields"][0] + np.float32(1.0))
        for q in range(1, 4):
            assert np.array_equal(f[q], p["fields"][q])
    for a, b in zip(ref["particles"], mine["particles"]):
        assert np.array_equal(a, b)
