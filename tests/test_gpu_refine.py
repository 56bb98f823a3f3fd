"""GPU parity of refined levels computed in windows around the parent voids
(avism_refine_*): the void-in-void hand-off of avism.f90:742-759, 957-968,
2123-2161 restricted to windows, since the level's whole grid is not stored.

* Fields: a window cell's velocities equal those of the same cell of the
  level's whole grid (avism_next_level + avism_particles_to_grid) bit for bit,
  DIVERCO inside the window too; the cell masses (U1CO x MEANDENS) agree to
  the rounding of the two divisions wherever HPART is the whole grid's, which
  a margin wider than twice the largest smoothing length guarantees for the
  parent's cells.  MEANDENS (total mass over the level's cells) agrees with
  the whole grid's FP64 mean to 1e-6.
* Several atlases give every window the same fields as one atlas.
* The void stage of an atlas equals the oracle's void stage on the atlas's
  fields, with MARCAP built here from the parent level's MARCA and the window
  table, bit for bit (as test_gpu_voids).  Two nested levels, periodic and
  not.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O
from test_gpu_voids import PARAMS, compare, gpu_void_outputs

pytestmark = pytest.mark.gpu

K = 32


def _ctx(n, L0, per, rmin_sub=1.0):
    from avism_amd._lib import Context
    c = Context(0)
    c.set_grid(n, L0, per)
    c.set_cosmology(0.678, 0.31, 0.0)
    c.set_void_params(PARAMS.dens_thre, PARAMS.dens_thre2, PARAMS.grad_thre, PARAMS.div_thre, PARAMS.rmin,
                      rmin_sub, PARAMS.nvoid_max)
    return c


def _params(rmin_sub=1.0):
    return O.VoidParams(**{**PARAMS.__dict__, "rmin_sub": rmin_sub})


def _as_ref(res: O.LevelResult) -> dict:
    return dict(nvoid=res.nvoid, cubes_i=res.cubes_i, cubes_f=res.cubes_f, father=res.father, indice=res.indice,
                nvoid2=res.nvoid2, uvoid=res.uvoid, marca=res.marca, nvoidt=res.nvoidt, ncellv=res.ncellv,
                volt_clean=res.volt_clean, vf=res.vf)


def _cube(a, side):
    return np.asarray(a).reshape(side, side, side)  # [z, y, x]


def _window_cells(arr_atlas, A, win):
    par, w, ox, oy, oz, ax, ay, az = (int(v) for v in win)
    return _cube(arr_atlas, A)[az:az + w, ay:ay + w, ax:ax + w]


def _level_cells(arr_level, n, win, periodic):
    par, w, ox, oy, oz = (int(v) for v in win[:5])
    l = np.arange(w)
    ix, iy, iz = ox + l, oy + l, oz + l
    if periodic:
        ix, iy, iz = ix % n, iy % n, iz % n
    return _cube(arr_level, n)[np.ix_(iz, iy, ix)]


def _atlas_marcap(tab, A, parent_marca, src_of):
    """MARCAP of an atlas from the parent level's MARCA: window cell l reads the
    parent cell floor((o + l) / 2); src_of(parent id) gives (grid, origin in the
    parent level's cells, origin in that grid, side, periodic wrap)."""
    mp = np.zeros((A, A, A), np.int32)
    for win in tab:
        par, w, ox, oy, oz, ax, ay, az = (int(v) for v in win)
        grid, so, sa, sw, wrap = src_of(par)
        l = np.arange(w)
        idx = []
        for o, s0 in zip((oz, oy, ox), (so[2], so[1], so[0])):
            g = (o + l) // 2 - s0
            if wrap:
                g = g % sw
            idx.append(g)
        inside = [(g >= 0) & (g < sw) for g in idx]
        gz, gy, gx = (np.clip(g, 0, sw - 1) for g in idx)
        vals = grid[np.ix_(sa[2] + gz, sa[1] + gy, sa[0] + gx)]
        ok = inside[0][:, None, None] & inside[1][None, :, None] & inside[2][None, None, :]
        mp[az:az + w, ay:ay + w, ax:ax + w] = np.where(ok & (vals == par), par, 0)
    return mp.ravel()


def _base_level(arr, n, L0, per, rmin_sub=1.0):
    c = _ctx(n, L0, per, rmin_sub)
    c.load_particles(*arr)
    c.particles_to_grid(k=K)
    out = gpu_void_outputs(c)
    return c, out


def _atlas_vs_oracle(c, tab, info, parent_marca, src_of, reqp, rmin_sub=1.0):
    """Void stage of the loaded atlas on the GPU and in the oracle; the GPU's outputs."""
    A = int(info["side"])
    f = c.fields()
    marcap = _atlas_marcap(tab, A, parent_marca, src_of)
    L0a = float(np.float32(A) * np.float32(info["dx"]))
    ga = O.Grid(A, L0a, 0)
    rete = O.cosmo(0.678, 0.31, 0.0)[1]
    ref = O.void_stage(ga, f["u1co"], f["div"], _params(rmin_sub), marcap=marcap, reqp=reqp,
                       meandens=c.stats().meandens, rete=rete)
    out = gpu_void_outputs(c)
    compare(out, lambda kk: _as_ref(ref)[kk])
    assert np.array_equal(out["father"], ref.father)
    # every cube's parent is a window's parent and its centre lies in that window
    pw = {int(w[0]): w for w in tab}
    for v in range(out["nvoid"]):
        w = pw[int(out["father"][v])]
        for axis in range(3):
            ic = int(out["cubes_i"][6 + axis][v]) - 1  # 0-based atlas cell
            assert w[5 + axis] <= ic < w[5 + axis] + w[1]
    return out


@pytest.fixture(scope="module")
def zel48():
    from avism_amd import synth
    p = synth.zeldovich(48, 96.0, seed=11)
    return p.numpy(), 48, 96.0, 1


@pytest.fixture(scope="module")
def whole_level2(zel48):
    """The level's whole grid (2N) for comparison: avism_next_level + fields."""
    arr, n, L0, per = zel48
    c, _ = _base_level(arr, n, L0, per)
    c.next_level()
    c.particles_to_grid(k=K)
    f = c.fields()
    mean = float(c.stats().meandens)
    return f, mean


def test_refine_fields_match_whole_level(zel48, whole_level2):
    arr, n, L0, per = zel48
    fw, mean_w = whole_level2
    n2 = 2 * n
    c, out1 = _base_level(arr, n, L0, per)
    assert out1["nvoidt"] > 0
    margin = 20
    assert c.refine_plan(margin=margin) == 1
    c.refine_load(0, K)
    tab, info = c.refine_windows()
    A = int(info["side"])
    print(f"[refine fields] windows={len(tab)} atlas={A}^3 level={info['n_level']}", flush=True)
    assert info["n_level"] == n2 and len(tab) > 0
    fa = c.fields()
    mean_a = float(info["meandens"])
    assert abs(mean_a / mean_w - 1.0) <= 1e-6
    worst = 0.0
    for win in tab:
        w = int(win[1])
        for key in ("u", "v", "w"):
            assert np.array_equal(_window_cells(fa[key], A, win), _level_cells(fw[key], n2, win, per)), key
        inner = (slice(1, w - 1),) * 3
        assert np.array_equal(_window_cells(fa["div"], A, win)[inner], _level_cells(fw["div"], n2, win, per)[inner])
        # cell masses: U1CO x MEANDENS, on the cells at least margin - 4 inside
        cs = (slice(margin - 4, w - margin + 4),) * 3
        a = _window_cells(fa["u1co"], A, win)[cs].astype(np.float64) * mean_a
        b = _level_cells(fw["u1co"], n2, win, per)[cs].astype(np.float64) * mean_w
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30 * mean_w)
        worst = max(worst, float(rel.max(initial=0.0)))
    print(f"[refine fields] worst cell-mass difference {worst:.2e}", flush=True)
    assert worst <= 3e-7


def test_refine_void_stage_vs_oracle_two_levels(zel48):
    """Level 2 in one atlas, then level 3 planned from that atlas."""
    arr, n, L0, per = zel48
    c, out1 = _base_level(arr, n, L0, per)
    g1 = O.Grid(n, L0, per)
    m1 = O.inner_of_ext(g1, out1["marca"]).reshape(n, n, n)
    reqp1 = np.where(out1["uvoid"] == -1, out1["vf"][11], 0).astype(np.float32)
    assert c.refine_plan(margin=8) == 1
    c.refine_load(0, K)
    tab2, info2 = c.refine_windows()
    src1 = lambda par: (m1, (0, 0, 0), (0, 0, 0), n, per)
    out2 = _atlas_vs_oracle(c, tab2, info2, m1, src1, reqp1)
    print(f"[refine level 2] windows={len(tab2)} cubes={out2['nvoid']} voids={out2['nvoidt']}", flush=True)
    assert out2["nvoid"] > 0
    # level 3 from the atlas's voids
    A2 = int(info2["side"])
    m2 = out2["marca"].reshape(A2, A2, A2)
    reqp2 = np.where(out2["uvoid"] == -1, out2["vf"][11], 0).astype(np.float32)
    win2 = {int(w[0]): w for w in tab2}
    na = c.refine_plan(margin=8)
    assert na >= 1
    c.refine_load(0, K)
    tab3, info3 = c.refine_windows()
    assert info3["depth"] == 2 and info3["n_level"] == 4 * n

    def src2(par):
        w = win2[int(out2["father"][par - 1])]
        return m2, (int(w[2]), int(w[3]), int(w[4])), (int(w[5]), int(w[6]), int(w[7])), int(w[1]), 0

    out3 = _atlas_vs_oracle(c, tab3, info3, m2, src2, reqp2)
    print(f"[refine level 3] windows={len(tab3)} cubes={out3['nvoid']} voids={out3['nvoidt']}", flush=True)
    c.refine_pop()
    c.refine_pop()


def test_refine_several_atlases_equal_one(zel48):
    arr, n, L0, per = zel48
    c, _ = _base_level(arr, n, L0, per)
    assert c.refine_plan(margin=8) == 1
    c.refine_load(0, K)
    tab, info = c.refine_windows()
    A = int(info["side"])
    one = {int(w[0]): {k: _window_cells(v, A, w).copy() for k, v in c.fields().items()} for w in tab}
    c.refine_pop()
    # the base level again (the context's grid is now an atlas)
    c, _ = _base_level(arr, n, L0, per)
    big = max(int(w[1]) for w in tab)
    na = c.refine_plan(margin=8, max_cells=int(big) ** 3 * 2)
    print(f"[refine atlases] {len(tab)} windows in {na} atlases", flush=True)
    assert na > 1
    seen = 0
    for a in range(na):
        c.refine_load(a, K)
        t, inf = c.refine_windows()
        Aa = int(inf["side"])
        assert Aa ** 3 <= big ** 3 * 2
        f = c.fields()
        for w in t:
            ref = one[int(w[0])]
            for key, v in f.items():
                got = _window_cells(v, Aa, w)
                if key == "div":
                    s = (slice(1, int(w[1]) - 1),) * 3
                    assert np.array_equal(got[s], ref[key][s]), key
                else:
                    assert np.array_equal(got, ref[key]), key
            seen += 1
    assert seen == len(tab)


def test_refine_non_periodic_vs_oracle():
    """Mock voids in a non-periodic box: windows clipped to the level's grid."""
    from avism_amd import synth
    p, _, _ = synth.mock_voids(n0=200_000, L0=100.0, nvoids=12, rmin=8.0, rmax=16.0, seed=23)
    arr = p.numpy()
    n, L0, per = 32, 100.0, 0
    c, out1 = _base_level(arr, n, L0, per)
    assert out1["nvoidt"] > 0
    g1 = O.Grid(n, L0, per)
    m1 = O.inner_of_ext(g1, out1["marca"]).reshape(n, n, n)
    reqp1 = np.where(out1["uvoid"] == -1, out1["vf"][11], 0).astype(np.float32)
    assert c.refine_plan(margin=6) >= 1
    c.refine_load(0, K)
    tab, info = c.refine_windows()
    for w in tab:  # inside the level's grid
        assert all(0 <= int(w[2 + a]) and int(w[2 + a]) + int(w[1]) <= 2 * n for a in range(3))
    out2 = _atlas_vs_oracle(c, tab, info, m1, lambda par: (m1, (0, 0, 0), (0, 0, 0), n, 0), reqp1)
    print(f"[refine non-periodic] windows={len(tab)} cubes={out2['nvoid']} voids={out2['nvoidt']}", flush=True)


def test_refine_state_errors(zel48):
    from avism_amd._lib import AvismError
    arr, n, L0, per = zel48
    c = _ctx(n, L0, per)
    c.load_particles(*arr)
    c.particles_to_grid(k=K)
    with pytest.raises(AvismError):  # no catalogue yet
        c.refine_plan()
    with pytest.raises(AvismError):  # no plan
        c.refine_load(0, K)
    gpu_void_outputs(c)
    with pytest.raises(AvismError):
        c.refine_plan(margin=2)
    na = c.refine_plan()
    assert na == 1
    with pytest.raises(AvismError):
        c.refine_load(1, K)
    c.refine_load(0, K)
    with pytest.raises(AvismError):  # the grid is an atlas, not the particles' box
        c.load_particles(*arr)
    with pytest.raises(AvismError):
        c.next_level()
