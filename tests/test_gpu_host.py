"""The Fortran host (host/avism_host.x) as a drop-in for the reference program:
it reads config/voids.dat and input_data/bin_file_{part,grid}NNNNN, runs the
path through the C ABI and writes output_files/{voids,map,cubes}NNNNN.

Grid input (FLAG_DATA=2) of a reference fixture must reproduce the reference's
map (MARCA, delta, divergence) bit for bit and its catalogue (ids, centres,
volumes and radii exactly; GXC, EPS, IP, UMEAN and MTOT within the tolerances
of the void-stage tests).  Particle input (FLAG_DATA=1) must write exactly what
the library returns through the Python binding for the same particles."""
from __future__ import annotations

import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
HOST = ROOT / "host" / "avism_host.x"
TOL = {"gxc": 1e-5, "gyc": 1e-5, "gzc": 1e-5, "eps": 1e-4, "ip": 1e-4, "umean": 1e-5, "mtot": 1e-5}
# catalogue line: IND XC YC ZC GXC GYC GZC VOL REQ UMEAN EPS IP INDP REQP MTOT (avism.f90:2081-2083)
COLS = ["ind", "xc", "yc", "zc", "gxc", "gyc", "gzc", "volnew", "req", "umean", "eps", "ip", "indp", "reqp", "mtot"]
VF_ROW = {"gxc": 0, "gyc": 1, "gzc": 2, "xc": 3, "yc": 4, "zc": 5, "volnew": 6, "umean": 7, "mtot": 8,
          "eps": 9, "ip": 10, "req": 11}


def _record(f, *arrays):
    payload = b"".join(np.ascontiguousarray(a).tobytes() for a in arrays)
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def _run_host(tmp: Path, cfg, args=()):
    from avism_amd.config import write_voids_dat
    if not HOST.exists():
        pytest.fail("host/avism_host.x is not built (make -C host)")
    (tmp / "config").mkdir(exist_ok=True)
    (tmp / "output_files").mkdir(exist_ok=True)
    write_voids_dat(cfg, tmp / "config" / "voids.dat")
    print(f"[host run in {tmp}]", flush=True)
    r = subprocess.run([str(HOST), *args], cwd=tmp, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def _read_map(path: Path, n: int):
    raw = np.fromfile(path, dtype=np.uint8)
    m = n ** 3 * 4
    assert raw.size == 3 * m
    return raw[:m].view(np.int32), raw[m:2 * m].view(np.float32), raw[2 * m:].view(np.float32)


def _read_voids(path: Path):
    """List-directed records (a compiler may wrap them over lines): a 7-value
    file header, a 6-value level header, then 15 values per void."""
    tok = path.read_text().split()
    head, level = tok[:7], tok[7:13]
    rows = np.array([float(t) for t in tok[13:]]).reshape(-1, 15)
    return head, level, rows


def _inner_of_ext(ext: np.ndarray, n: int, low1: int, next_: int):
    e = ext.reshape(next_, next_, next_)  # (K, J, I), x fastest
    o = 1 - low1
    return e[o:o + n, o:o + n, o:o + n].reshape(-1)


def test_host_grid_input_matches_reference(tmp_path, golden_dir):
    from avism_amd.config import VoidsConfig
    from avism_amd._lib import Context
    g = dict(np.load(golden_dir / "grid_p1_n24.npz"))
    n, L0 = int(g["meta_n"]), float(g["meta_L0"])
    cfg = VoidsConfig(first=1, last=1, every=1, ncox=n, ncoy=n, ncoz=n, lado0=L0, flag_period=1,
                      flag_data=2, nhyx=n, nhyy=n, nhyz=n, nvoid_max=200000, dens_thre=-0.6, dens_thre2=10.0,
                      grad_thre=0.25, div_thre=0.0, rmin=1.0, rmin_sub=1.0, flag_write_cubes=1)
    (tmp_path / "input_data").mkdir()
    with open(tmp_path / "input_data" / "bin_file_grid00001", "wb") as f:
        _record(f, np.float32(0.0))
        for k in ("in_delta", "in_vx", "in_vy", "in_vz"):
            _record(f, g[k].astype(np.float32))
    _run_host(tmp_path, cfg)
    gi = Context(0).set_grid(n, L0, 1)
    marca, delta, div = _read_map(tmp_path / "output_files" / "map00001", n)
    assert np.array_equal(marca, _inner_of_ext(g["ref_marca"], n, gi.low1, gi.next))
    assert np.array_equal(delta, np.float32(-1.0) + g["ref_u1co"])
    assert np.array_equal(div, g["ref_div"])
    head, level, rows = _read_voids(tmp_path / "output_files" / "voids00001")
    assert head[:6] == ["1", "0", "0", str(n), str(n), str(n)]
    assert int(level[1]) == int(g["ref_nvoid"]) and int(level[2]) == int(g["ref_nvoidt"])
    ind = rows[:, 0].astype(int)
    main = g["ref_uvoid"] == -1
    ref_ids = g["ref_indice"][main[g["ref_indice"] - 1]]
    assert np.array_equal(ind, ref_ids)
    vf = g["ref_vf"]
    for ci, name in enumerate(COLS):
        if name not in VF_ROW:
            continue
        a = rows[:, ci]
        b = vf[VF_ROW[name]][ind - 1].astype(np.float64)
        if name in TOL:
            scale = np.maximum(np.abs(b), 1.0 if name in ("gxc", "gyc", "gzc") else 1e-30)
            assert np.max(np.abs(a - b) / scale) <= TOL[name] + 1e-6, name  # + text rounding
        else:
            assert np.allclose(a, b, rtol=1e-6, atol=0), name  # list-directed text keeps 9 digits
    cubes = (tmp_path / "output_files" / "cubes00001").read_text().split()
    assert (len(cubes) - 3 - 3) == 14 * int(g["ref_nvoid"])  # two headers, 14 values per cube


@pytest.mark.parametrize("vel", [1, 0])
def test_host_particle_input_matches_binding(tmp_path, vel):
    """BinPart input with velocities (READ_BINARY_PART_1) and without
    (READ_BINARY_PART_2, FLAG_VEL_AVAILABLE = 0: density only, linear-theory
    divergence); the host's map and catalogue equal the binding's."""
    from avism_amd import synth
    from avism_amd.config import VoidsConfig
    from avism_amd._lib import Context
    n, L0 = 32, 64.0
    p = synth.zeldovich(n, L0, seed=17)
    arr = p.numpy()
    cfg = VoidsConfig(first=3, last=3, every=1, ncox=n, ncoy=n, ncoz=n, lado0=L0, flag_period=1, flag_data=1,
                      kneighbours=32, partired=10 ** 8, flag_vel_available=vel)
    (tmp_path / "input_data").mkdir()
    with open(tmp_path / "input_data" / "bin_file_part00003", "wb") as f:
        _record(f, np.int64(arr[0].size), np.float32(0.0))
        for a in (arr if vel else [arr[0], arr[1], arr[2], arr[6]]):
            _record(f, a.astype(np.float32))
    _run_host(tmp_path, cfg)
    c = Context(0)
    c.set_grid(n, L0, 1)
    c.set_cosmology(cfg.ache, cfg.omega0, 0.0)
    c.set_void_params(cfg.dens_thre, cfg.dens_thre2, cfg.grad_thre, cfg.div_thre, cfg.rmin, cfg.rmin_sub,
                      cfg.nvoid_max)
    if vel:
        c.load_particles(*arr)
        c.particles_to_grid(k=32)
    else:
        c.load_particles(arr[0], arr[1], arr[2], None, None, None, arr[6])
        c.particles_to_density(k=32)
    nv = c.find_voids(0)
    c.merge_and_post()
    cat = c.catalogue(nv)
    f = c.fields()
    marca, delta, div = _read_map(tmp_path / "output_files" / "map00003", n)
    assert np.array_equal(marca, c.marca(inner_only=True))
    assert np.array_equal(delta, f["u1co"] - np.float32(1.0))
    assert np.array_equal(div, f["div"])
    _, level, rows = _read_voids(tmp_path / "output_files" / "voids00003")
    assert int(level[1]) == nv and int(level[2]) == int(cat["stats"].nvoidt)
    main_ids = cat["indice"][cat["uvoid"][cat["indice"] - 1] == -1]
    assert np.array_equal(rows[:, 0].astype(int), main_ids)
    assert np.allclose(rows[:, 7], cat["vf"][6][main_ids - 1], rtol=1e-6)


def test_host_two_levels_match_binding(tmp_path):
    """LEVMIN=0, LEVMAX=1 with particle input: one catalogue block and one map
    block per level; the second level's INDP is the cube's FATHER and its
    parent radius the parent's REQ (avism.f90:2073-2083, 2123-2161)."""
    from avism_amd import synth
    from avism_amd.config import VoidsConfig
    from avism_amd._lib import Context
    n, L0 = 16, 40.0
    arr = synth.zeldovich(16, L0, seed=22, rms_disp=2.0).numpy()
    cfg = VoidsConfig(first=4, last=4, every=1, ncox=n, ncoy=n, ncoz=n, levmin=0, levmax=1, lado0=L0,
                      flag_period=1, flag_data=1, kneighbours=16, partired=10 ** 8, rmin_sub=5.0)
    (tmp_path / "input_data").mkdir()
    with open(tmp_path / "input_data" / "bin_file_part00004", "wb") as f:
        _record(f, np.int64(arr[0].size), np.float32(0.0))
        for a in arr:
            _record(f, a.astype(np.float32))
    _run_host(tmp_path, cfg)
    c = Context(0)
    c.set_grid(n, L0, 1)
    c.set_cosmology(cfg.ache, cfg.omega0, 0.0)
    c.set_void_params(cfg.dens_thre, cfg.dens_thre2, cfg.grad_thre, cfg.div_thre, cfg.rmin, cfg.rmin_sub,
                      cfg.nvoid_max)
    c.load_particles(*arr)
    tok = (tmp_path / "output_files" / "voids00004").read_text().split()
    assert tok[:3] == ["2", "0", "1"]
    pos = 7
    raw = np.fromfile(tmp_path / "output_files" / "map00004", dtype=np.uint8)
    mpos = 0
    req_prev = None
    for lev in (0, 1):
        nn = n * 2 ** lev
        if lev:
            assert c.next_level().n == nn
        c.particles_to_grid(k=16)
        nv = c.find_voids(0)
        c.merge_and_post()
        cat = c.catalogue(nv)
        _, _, father = c.cubes(nv)
        f = c.fields()
        head = tok[pos:pos + 6]
        pos += 6
        nvt = int(cat["stats"].nvoidt)
        assert int(head[0]) == lev and int(head[1]) == nv and int(head[2]) == nvt
        rows = np.array([float(t) for t in tok[pos:pos + 15 * nvt]]).reshape(-1, 15)
        pos += 15 * nvt
        main_ids = cat["indice"][cat["uvoid"][cat["indice"] - 1] == -1]
        assert np.array_equal(rows[:, 0].astype(int), main_ids)
        assert np.allclose(rows[:, 8], cat["vf"][11][main_ids - 1], rtol=1e-6)
        if lev == 0:
            assert np.all(rows[:, 12] == 0) and np.all(rows[:, 13] == 0)
        else:
            assert nv > 0 and len(np.unique(father)) >= 2
            assert np.array_equal(rows[:, 12].astype(int), father[main_ids - 1])
            assert np.allclose(rows[:, 13], req_prev[father[main_ids - 1] - 1], rtol=1e-6)
        req_prev = np.where(cat["uvoid"] == -1, cat["vf"][11], 0.0)
        m = nn ** 3 * 4
        blk = raw[mpos:mpos + 3 * m]
        mpos += 3 * m
        assert np.array_equal(blk[:m].view(np.int32), c.marca(inner_only=True))
        assert np.array_equal(blk[m:2 * m].view(np.float32), f["u1co"] - np.float32(1.0))
        assert np.array_equal(blk[2 * m:].view(np.float32), f["div"])
    assert pos == len(tok) and mpos == raw.size


def test_host_survey_matches_binding(tmp_path):
    """Survey mode through the host: input_data/mask.dat (stream), rands.dat
    (sequential records) and a bin_file_part without velocities; the map and
    catalogue equal the binding's avism_survey_density run."""
    from avism_amd import synth
    from avism_amd.config import VoidsConfig
    from avism_amd._lib import Context
    n, L0 = 32, 64.0
    arr = synth.zeldovich(n, L0, seed=19).numpy()
    gal = [arr[0], arr[1], arr[2], arr[6]]
    rng = np.random.default_rng(19)
    nr = 2 * n ** 3
    rands = [rng.uniform(-L0 / 2, L0 / 2, nr).astype(np.float32) for _ in range(3)]
    rands.append(np.full(nr, 2e10, np.float32))
    sm = synth.survey_mask(n, 0.2, seed=19)
    # survey mode is non-periodic (avism.f90:278-281)
    cfg = VoidsConfig(first=3, last=3, every=1, ncox=n, ncoy=n, ncoz=n, lado0=L0, flag_period=0, flag_data=1,
                      kneighbours=16, partired=10 ** 8, flag_vel_available=0, flag_survey=1)
    (tmp_path / "input_data").mkdir()
    with open(tmp_path / "input_data" / "bin_file_part00003", "wb") as f:
        _record(f, np.int64(gal[0].size), np.float32(0.0))
        for a in gal:
            _record(f, a.astype(np.float32))
    with open(tmp_path / "input_data" / "mask.dat", "wb") as f:
        f.write(np.int32(n).tobytes() + np.array([-L0 / 2, L0 / 2, L0], np.float32).tobytes())
        f.write(np.ascontiguousarray(sm, np.int8).tobytes())
    with open(tmp_path / "input_data" / "rands.dat", "wb") as f:
        _record(f, np.int32(nr))
        for a in rands:
            _record(f, a)
    _run_host(tmp_path, cfg)
    c = Context(0)
    c.set_grid(n, L0, 0)
    c.set_cosmology(cfg.ache, cfg.omega0, 0.0)
    c.set_void_params(cfg.dens_thre, cfg.dens_thre2, cfg.grad_thre, cfg.div_thre, cfg.rmin, cfg.rmin_sub,
                      cfg.nvoid_max)
    c.set_mask(sm)
    c.load_particles(gal[0], gal[1], gal[2], None, None, None, gal[3])
    c.load_randoms(*rands)
    c.survey_density(k=16)
    nv = c.find_voids(0)
    c.merge_and_post()
    cat = c.catalogue(nv)
    f = c.fields()
    marca, delta, div = _read_map(tmp_path / "output_files" / "map00003", n)
    assert np.array_equal(marca, c.marca(inner_only=True))
    assert np.array_equal(delta, f["u1co"] - np.float32(1.0))
    assert np.array_equal(div, f["div"])
    _, level, rows = _read_voids(tmp_path / "output_files" / "voids00003")
    assert int(level[1]) == nv and int(level[2]) == int(cat["stats"].nvoidt)
    assert int(cat["stats"].nvoidt) > 0
