"""ctypes bindings to the C oracle (oracle/avism_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

``run_level`` restates the reference driver's single-level flow on the CPU
(avism.f90:1177-2085 for particle input, with reader.f90:523-552 for grid input)
and returns every intermediate the GPU path is compared against.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle_avism.so"

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")


def build(force: bool = False) -> Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


class _VP(C.Structure):
    _fields_ = [("dens_thre", C.c_float), ("dens_thre2", C.c_float), ("grad_thre", C.c_float),
                ("div_thre", C.c_float), ("rmin", C.c_float), ("rmin_sub", C.c_float),
                ("nvoid_max", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    L.orc_grid_new.restype = vp
    L.orc_grid_new.argtypes = [C.c_int32, C.c_float, C.c_int32]
    L.orc_grid_free.argtypes = [vp]
    L.orc_grid_info.argtypes = [vp, _i32p, _f32p, C.c_void_p]
    L.orc_grid_rad.restype = C.c_float
    L.orc_grid_rad.argtypes = [vp, C.c_int32]
    L.orc_knn_new.restype = vp
    L.orc_knn_new.argtypes = [C.c_int64, _f32p, _f32p, _f32p, C.c_float, C.c_int32]
    L.orc_knn_free.argtypes = [vp]
    L.orc_knn_query.restype = C.c_int32
    L.orc_knn_query.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_int32, _i64p, _f32p]
    L.orc_ball_query.restype = C.c_int64
    L.orc_ball_query.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int64, _i64p, _f32p]
    L.orc_convert_units.argtypes = [C.c_int64, _f32p, _f32p, _f32p, _f32p]
    L.orc_ppart_dens.argtypes = [vp, C.c_int64, _f32p, _f32p, _f32p, _f32p, C.c_int32, _f32p, _f32p]
    L.orc_vvel_vw.argtypes = [vp, vp, C.c_int32, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_void_p,
                              _f32p, _f32p, _f32p, _f32p]
    L.orc_vvel_cells.argtypes = [vp, vp, C.c_int32, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_int64, _i64p,
                                 _f32p, _f32p, _f32p]
    L.orc_h_distance.argtypes = [vp, vp, C.c_int32, C.c_void_p, _f32p]
    L.orc_fill_h_zeros.argtypes = [vp, C.c_int64, _f32p, _f32p, _f32p, C.c_int32, _f32p]
    L.orc_ddens_sph.argtypes = [vp, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f64p]
    L.orc_ddens_sph_par.argtypes = [vp, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f64p]
    L.orc_dens_tsc.argtypes = [vp, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.orc_vel_tsc.argtypes = [vp, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_int32, _f32p, _f32p,
                              _f32p]
    L.orc_normalise.restype = C.c_float
    L.orc_normalise.argtypes = [vp, C.c_float, C.c_float, C.c_float, _f32p, _f32p, C.POINTER(C.c_double)]
    L.orc_diver_uniform.argtypes = [C.c_int32, C.c_float, _f32p, _f32p, _f32p, C.c_int32, _f32p]
    L.orc_amr_gas.argtypes = [C.c_int32, C.c_float, C.c_int32, C.c_int32, _f32p, C.c_int32, _i32p, _f32p, _f32p,
                              _f32p, _f32p]
    L.orc_amr_gas.restype = C.c_int32
    L.orc_extend_f32.argtypes = [vp, _f32p, _f32p]
    L.orc_extend_i32.argtypes = [vp, _i32p, _i32p]
    L.orc_indexx.restype = C.c_int32
    L.orc_indexx.argtypes = [C.c_int64, _f32p, _i32p]
    L.orc_voidfind.restype = C.c_int32
    L.orc_voidfind.argtypes = [vp, C.POINTER(_VP), _f32p, _f32p, _i32p, _i8p, C.c_void_p, _i32p, _f32p,
                               _i32p, _i64p, C.c_int64, C.c_void_p, C.c_void_p]
    L.orc_merge_post.restype = C.c_int32
    L.orc_merge_post.argtypes = [vp, C.POINTER(_VP), C.c_int32, _i32p, _f32p, _i32p, _f32p, C.c_float,
                                 C.c_float, _i32p, _i32p, _i32p, _f32p, _i32p, _i32p, _f64p, C.c_int32]
    L.orc_survey.restype = C.c_float
    L.orc_survey.argtypes = [vp, C.c_float, C.c_float, C.c_float, _f32p, _f32p, C.c_void_p, _f32p, _f32p]
    L.orc_div_linear.restype = None
    L.orc_div_linear.argtypes = [C.c_int64, _f32p, C.c_float, C.c_float, C.c_float, _f32p]
    L.orc_cosmo.argtypes = [C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float),
                            C.POINTER(C.c_float)]
    _lib = L
    return L


class Grid:
    """MALLA geometry (grid.f90:3-76) with the PBC extension (avism.f90:780-865)."""

    def __init__(self, n: int, L0: float, periodic: int):
        self._p = lib().orc_grid_new(int(n), float(L0), int(periodic))
        ints = np.zeros(6, np.int32)
        fl = np.zeros(4, np.float32)
        lib().orc_grid_info(self._p, ints, fl, None)
        rad = np.zeros(int(ints[5]), np.float32)
        lib().orc_grid_info(self._p, ints, fl, rad.ctypes.data_as(C.c_void_p))
        self.n, self.periodic, self.nplus, self.low1, self.low2, self.next = (int(v) for v in ints)
        self.L0, self.lado0plus, self.dx, self.rx1 = (float(v) for v in fl)
        self.rad_ext = rad

    def rad(self, i: int) -> float:
        return float(lib().orc_grid_rad(self._p, int(i)))

    def __del__(self):
        try:
            lib().orc_grid_free(self._p)
        except Exception:
            pass


def cosmo(ache, omega0, zeta):
    r = [C.c_float(), C.c_float(), C.c_float()]
    lib().orc_cosmo(ache, omega0, zeta, C.byref(r[0]), C.byref(r[1]), C.byref(r[2]))
    return tuple(v.value for v in r)  # rote, rete, hte


def indexx(arr: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(arr, np.float32)
    out = np.zeros(max(1, a.size), np.int32)
    rc = lib().orc_indexx(a.size, a if a.size else np.zeros(1, np.float32), out)
    if rc < 0:
        raise RuntimeError("NSTACK too small in indexx")
    return out[: a.size]


class KNN:
    def __init__(self, x, y, z, L0, periodic):
        self.x = np.ascontiguousarray(x, np.float32)
        self.y = np.ascontiguousarray(y, np.float32)
        self.z = np.ascontiguousarray(z, np.float32)
        self._p = lib().orc_knn_new(self.x.size, self.x, self.y, self.z, float(L0), int(periodic))

    def query(self, q, k):
        idx = np.zeros(k, np.int64)
        d2 = np.zeros(k, np.float32)
        c = lib().orc_knn_query(self._p, float(q[0]), float(q[1]), float(q[2]), int(k), idx, d2)
        return idx[:c], d2[:c]

    def ball(self, q, r, cap=1 << 20):
        idx = np.zeros(cap, np.int64)
        d2 = np.zeros(cap, np.float32)
        c = lib().orc_ball_query(self._p, float(q[0]), float(q[1]), float(q[2]), float(r), cap, idx, d2)
        return idx[:c], d2[:c]

    def __del__(self):
        try:
            lib().orc_knn_free(self._p)
        except Exception:
            pass


@dataclass
class VoidParams:
    dens_thre: float = -0.6
    dens_thre2: float = 10.0
    grad_thre: float = 0.25
    div_thre: float = 0.0
    rmin: float = 1.0
    rmin_sub: float = 1.0
    nvoid_max: int = 1 << 20

    def c(self):
        return _VP(self.dens_thre, self.dens_thre2, self.grad_thre, self.div_thre, self.rmin,
                   self.rmin_sub, self.nvoid_max)


@dataclass
class LevelResult:
    grid: Grid
    u1co: np.ndarray            # inner N^3 (1+delta)
    div: np.ndarray             # inner N^3
    meandens: float
    nvoid: int = 0
    cubes_i: np.ndarray = None  # [9, nvoid]: inix,iniy,iniz,finx,finy,finz,icx,icy,icz
    cubes_f: np.ndarray = None  # [7, nvoid]: rinx,riny,rinz,rfix,rfiy,rfiz,vol
    father: np.ndarray = None
    indice: np.ndarray = None
    uvoid: np.ndarray = None
    marca: np.ndarray = None    # ext
    vf: np.ndarray = None       # [13, nvoid]
    ncellv: np.ndarray = None
    nvoid2: int = 0
    nvoidt: int = 0
    volt_clean: float = 0.0
    stats: np.ndarray = None
    cand_lin: np.ndarray = None
    cand_r: np.ndarray = None
    extra: dict = field(default_factory=dict)


def fields_sampled(x, y, z, vx, vy, vz, m, n, L0, periodic, parts, cells, k=32):
    """PART_DENS at the particles `parts` and the SPH velocities (u, v, w) at
    the inner cells `cells` (0-based linear indices), by the same routines as
    fields_from_particles (particles.f90:742-930) but evaluated only there: for
    grids too large for the whole field stage on the CPU.  Inputs in file
    units."""
    L = lib()
    g = Grid(n, L0, periodic)
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    vx, vy, vz, m = (np.array(a, np.float32, copy=True) for a in (vx, vy, vz, m))
    L.orc_convert_units(x.size, vx, vy, vz, m)
    knn = KNN(x, y, z, L0, periodic)
    parts = np.ascontiguousarray(parts, np.int64)
    cells = np.ascontiguousarray(cells, np.int64)
    pd_s = np.zeros(parts.size, np.float32)
    hk_s = np.zeros(parts.size, np.float32)
    L.orc_ppart_dens(knn._p, parts.size, x[parts].copy(), y[parts].copy(), z[parts].copy(), m, k, pd_s, hk_s)
    # PART_DENS of every neighbour of the sampled cells, for the SPH weights
    rad = np.array([g.rad(i) for i in range(1, n + 1)], np.float32)
    ix, iy, iz = cells % n, (cells // n) % n, cells // (n * n)
    need = set()
    for a, b, c in zip(rad[ix], rad[iy], rad[iz]):
        idx, d2 = knn.query((a, b, c), k)
        need.update(idx.tolist())
        if d2[-1] <= g.dx * g.dx:  # a dense cell takes the ball of radius DX (particles.f90:868-878)
            ball, _ = knn.ball((a, b, c), g.dx, cap=1 << 14)
            need.update(ball.tolist())
    need = np.array(sorted(need), np.int64)
    pd_n = np.zeros(need.size, np.float32)
    L.orc_ppart_dens(knn._p, need.size, x[need].copy(), y[need].copy(), z[need].copy(), m, k, pd_n,
                     np.zeros(need.size, np.float32))
    pdens = np.ones(x.size, np.float32)
    pdens[need] = pd_n
    u = np.zeros(cells.size, np.float32); v = np.zeros(cells.size, np.float32); w = np.zeros(cells.size, np.float32)
    L.orc_vvel_cells(knn._p, g._p, k, m, pdens, vx, vy, vz, cells.size, cells, u, v, w)
    return dict(part_dens=pd_s, hk=hk_s, u=u, v=v, w=w)


def fields_from_particles(x, y, z, vx, vy, vz, m, n, L0, periodic, k=32, ache=0.678,
                          omega0=0.31, zeta=0.0, vel_interp=1, dens_interp=1, tsc_exact=False, smask=None,
                          parallel_ddens=False):
    """Field stage (avism.f90:1177-1257, 1371-1414). Inputs in file units
    (reader.f90:447-461 conversion applied here).  Returns a dict of arrays.
    tsc_exact: TSC velocity sums in FP64 instead of the reference's REAL*4
    one-thread order.  smask: inner n^3 int8 SMASK (None: all ones); VVEL and
    H_DISTANCE skip masked cells (particles.f90:312, 856), DDENS ignores it.
    vel_interp -1: velocities not available (FLAG_VEL_AVAILABLE = 0,
    avism.f90:1259-1300): density only, DIVERCO from linear theory
    (avism.f90:1646-1650); vx, vy, vz may be None.  parallel_ddens: DDENS on
    every thread with atomic FP64 deposits (the last bits of the cell sums then
    depend on the schedule, as in the reference's OpenMP run)."""
    L = lib()
    g = Grid(n, L0, periodic)
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    if vx is None:
        vx = vy = vz = np.zeros(np.size(x), np.float32)
    vx, vy, vz, m = (np.array(a, np.float32, copy=True) for a in (vx, vy, vz, m))
    L.orc_convert_units(x.size, vx, vy, vz, m)
    npart = x.size
    ncell = n ** 3
    u = np.zeros(ncell, np.float32); v = np.zeros(ncell, np.float32); w = np.zeros(ncell, np.float32)
    hpart = np.zeros(npart, np.float32)
    pdens = np.zeros(npart, np.float32)
    hk = np.zeros(npart, np.float32)
    knn = KNN(x, y, z, L0, periodic)
    sm = None if smask is None else np.ascontiguousarray(smask, np.int8).ctypes.data_as(C.c_void_p)
    if vel_interp == 1:
        L.orc_ppart_dens(knn._p, npart, x, y, z, m, k, pdens, hk)
        L.orc_vvel_vw(knn._p, g._p, k, m, pdens, vx, vy, vz, sm, hpart, u, v, w)
        L.orc_fill_h_zeros(knn._p, npart, x, y, z, k, hpart)
    elif vel_interp == 0:
        L.orc_vel_tsc(g._p, npart, x, y, z, vx, vy, vz, 1 if tsc_exact else 0, u, v, w)
    mass = np.zeros(ncell, np.float32)
    mass64 = np.zeros(ncell, np.float64)
    if dens_interp == 1:
        if vel_interp != 1:
            L.orc_h_distance(knn._p, g._p, k, sm, hpart)
            L.orc_fill_h_zeros(knn._p, npart, x, y, z, k, hpart)
        (L.orc_ddens_sph_par if parallel_ddens else L.orc_ddens_sph)(g._p, npart, x, y, z, hpart, m, mass, mass64)
    else:
        L.orc_dens_tsc(g._p, npart, x, y, z, m, mass)
    u1co = np.zeros(ncell, np.float32)
    s = C.c_double()
    meandens = L.orc_normalise(g._p, ache, omega0, zeta, mass, u1co, C.byref(s))
    div = np.zeros(ncell, np.float32)
    if vel_interp == -1:
        _, rete, hte = cosmo(ache, omega0, zeta)
        L.orc_div_linear(ncell, u1co, omega0, rete, hte, div)
    else:
        L.orc_diver_uniform(n, g.dx, u, v, w, periodic, div)
    return dict(grid=g, u=u, v=v, w=w, hpart=hpart, part_dens=pdens, hk=hk, mass=mass,
                mass64=mass64, u1co=u1co, div=div, meandens=float(meandens), sum64=s.value)


def survey_fields(gal, rand, n, L0, periodic, smask, k=32, dens_interp=1, ache=0.678, omega0=0.31, zeta=0.0):
    """Survey mode (FLAG_SURVEY = 1, avism.f90:1415-1597): galaxies (x, y, z, m)
    without velocities, randoms (x, y, z, m) with KNEIGHBOURS2 = min(8k, 256)
    (avism.f90:1482); D/(alpha R) density and linear-theory DIVERCO."""
    sm = np.ascontiguousarray(smask, np.int8)
    f = fields_from_particles(gal[0], gal[1], gal[2], None, None, None, gal[3], n, L0, periodic, k=k,
                              vel_interp=-1, dens_interp=dens_interp, smask=sm, ache=ache, omega0=omega0,
                              zeta=zeta)
    r = fields_from_particles(rand[0], rand[1], rand[2], None, None, None, rand[3], n, L0, periodic,
                              k=min(8 * k, 256), vel_interp=-1, dens_interp=dens_interp, smask=sm, ache=ache,
                              omega0=omega0, zeta=zeta)
    u1co = np.zeros(n ** 3, np.float32)
    div = np.zeros(n ** 3, np.float32)
    meandens = lib().orc_survey(f["grid"]._p, ache, omega0, zeta, f["mass"], r["mass"],
                                sm.ctypes.data_as(C.c_void_p), u1co, div)
    return dict(grid=f["grid"], u1co=u1co, div=div, meandens=float(meandens), hpart=f["hpart"],
                hrand=r["hpart"], mass=f["mass"], urand=r["mass"])


def fields_from_grid(delta, vx, vy, vz, n, L0, periodic, ache=0.678, omega0=0.31, zeta=0.0):
    """Grid input (reader.f90:541-549; avism.f90:1342 and 1599-1600)."""
    g = Grid(n, L0, periodic)
    u1co = (np.asarray(delta, np.float32) + np.float32(1.0)).astype(np.float32)
    u = (np.asarray(vx, np.float32) / np.float32(299792.458)).astype(np.float32)
    v = (np.asarray(vy, np.float32) / np.float32(299792.458)).astype(np.float32)
    w = (np.asarray(vz, np.float32) / np.float32(299792.458)).astype(np.float32)
    div = np.zeros(n ** 3, np.float32)
    lib().orc_diver_uniform(n, g.dx, u, v, w, periodic, div)
    rote, rete, _ = cosmo(ache, omega0, zeta)
    ul3 = np.float32(np.float32(10.98) * np.float32(10.98)) * np.float32(10.98)
    meandens = np.float32(rote) * np.float32(np.float32(9.1717e18) / ul3)
    return dict(grid=g, u=u, v=v, w=w, u1co=u1co, div=div, meandens=float(meandens))


def void_stage(g: Grid, u1co, div, params: VoidParams, marcap=None, reqp=None, smask=None,
               rete=None, meandens=1.0, stop_after=0, want_candidates=False) -> LevelResult:
    """MARK_ALL .. catalogue (avism.f90:1633-1973) for one level."""
    L = lib()
    ne3 = g.next ** 3
    u_ext = np.zeros(ne3, np.float32)
    d_ext = np.zeros(ne3, np.float32)
    L.orc_extend_f32(g._p, np.ascontiguousarray(u1co, np.float32), u_ext)
    L.orc_extend_f32(g._p, np.ascontiguousarray(div, np.float32), d_ext)
    mp = np.full(g.n ** 3, -1, np.int32) if marcap is None else np.ascontiguousarray(marcap, np.int32)
    mp_ext = np.zeros(ne3, np.int32)
    L.orc_extend_i32(g._p, mp, mp_ext)
    if smask is None:
        sm = np.ones(ne3, np.int8)
    elif np.asarray(smask).size == g.n ** 3:  # inner SMASK: the periodic shell stays 0 (avism.f90:866-904)
        sm = np.zeros((g.next,) * 3, np.int8)
        o = 1 - g.low1
        sm[o:o + g.n, o:o + g.n, o:o + g.n] = np.asarray(smask, np.int8).reshape((g.n,) * 3)
        sm = sm.ravel()
    else:
        sm = np.ascontiguousarray(smask, np.int8)
    NV = params.nvoid_max
    ci = np.zeros(9 * NV, np.int32)
    cf = np.zeros(7 * NV, np.float32)
    father = np.zeros(NV, np.int32)
    stats = np.zeros(4, np.int64)
    reqp_arr = None if reqp is None else np.ascontiguousarray(reqp, np.float32)
    cap = ne3 if want_candidates else 0
    cand_lin = np.zeros(max(cap, 1), np.int64)
    cand_r = np.zeros(max(cap, 1), np.int32)
    nvoid = L.orc_voidfind(g._p, C.byref(params.c()), u_ext, d_ext, mp_ext, sm,
                           None if reqp_arr is None else reqp_arr.ctypes.data_as(C.c_void_p),
                           ci, cf, father, stats, cap,
                           cand_lin.ctypes.data_as(C.c_void_p) if cap else None,
                           cand_r.ctypes.data_as(C.c_void_p) if cap else None)
    if nvoid < 0:
        raise RuntimeError(f"orc_voidfind failed: {nvoid}")
    res = LevelResult(grid=g, u1co=np.asarray(u1co, np.float32), div=np.asarray(div, np.float32),
                      meandens=meandens, nvoid=nvoid)
    res.stats = stats
    if want_candidates:
        res.cand_lin = cand_lin[: stats[0]]
        res.cand_r = cand_r[: stats[0]]
    res.cubes_i = ci.reshape(9, NV)[:, :nvoid].copy()
    res.cubes_f = cf.reshape(7, NV)[:, :nvoid].copy()
    res.father = father[:nvoid].copy()
    if stop_after < 0:
        return res
    if rete is None:
        rete = cosmo(0.678, 0.31, 0.0)[1]
    nv = max(nvoid, 1)
    indice = np.zeros(nv, np.int32)
    uvoid = np.zeros(nv, np.int32)
    marca = np.zeros(ne3, np.int32)
    vf = np.zeros(13 * nv, np.float32)
    ncellv = np.zeros(nv, np.int32)
    si = np.zeros(4, np.int32)
    sd = np.zeros(2, np.float64)
    rc = L.orc_merge_post(g._p, C.byref(params.c()), nvoid, ci, cf, father, u_ext, float(meandens),
                          float(rete), indice, uvoid, marca, vf, ncellv, si, sd, int(stop_after))
    if rc < 0:
        raise RuntimeError(f"orc_merge_post failed: {rc}")
    res.indice = indice[:nvoid]
    res.uvoid = uvoid[:nvoid]
    res.marca = marca
    res.vf = vf.reshape(13, nv)[:, :nvoid] if nvoid else np.zeros((13, 0), np.float32)
    res.ncellv = ncellv[:nvoid]
    res.nvoid2 = int(si[0])
    res.nvoidt = int(si[1])
    res.volt_clean = float(sd[0])
    return res


def inner_of_ext(g: Grid, a_ext: np.ndarray) -> np.ndarray:
    """The inner cells 1..N of an ext-grid array (x fastest)."""
    ne, n, o = g.next, g.n, 1 - g.low1
    return np.ascontiguousarray(np.asarray(a_ext).reshape(ne, ne, ne)[o:o + n, o:o + n, o:o + n]).ravel()


def next_level_inputs(g: Grid, marca_ext, uvoid, req):
    """Hand-off to the next void-in-void level (avism.f90:2123-2161): MARCAP on
    the grid of 2N, MARCAP(IX,JY,KZ) = MARCA(INT((IX-1)/2.)+1, ...), and
    REQP(IND) = REQ(IND) for the main voids (UVOID = -1), 0 for the others."""
    n = g.n
    mi = inner_of_ext(g, marca_ext).reshape(n, n, n)
    up = mi.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)
    reqp = np.where(np.asarray(uvoid) == -1, np.asarray(req, np.float32), np.float32(0.0)).astype(np.float32)
    return np.ascontiguousarray(up, np.int32).ravel(), reqp


def amr_gas(nb: int, L0: float, periodic: int, ir: int, base, pint, preal, pdata):
    """MASCLET AMR gas onto the uniform grid of level ir (SMOOTH_2, DIVER_UNIFORM,
    DIVER_FINA_GAS, VMESH; orc_amr_gas): (U1GCO, DIVERGCO) of n^3, n = nb 2^ir."""
    n = nb << ir if ir >= 0 else nb >> (-ir)
    b = np.ascontiguousarray(np.concatenate([np.ravel(a) for a in base]), np.float32)
    u = np.zeros(n ** 3, np.float32)
    d = np.zeros(n ** 3, np.float32)
    rc = lib().orc_amr_gas(nb, float(L0), int(periodic), int(ir), b, len(pint) // 6,
                           np.ascontiguousarray(pint, np.int32), np.ascontiguousarray(preal, np.float32),
                           np.ascontiguousarray(pdata, np.float32), u, d)
    if rc != 0:
        raise RuntimeError(f"orc_amr_gas failed: {rc}")
    return u, d


# ----------------------------------------------------------------- reference
REF_DIR = HERE / "_ref"


def run_reference_amr(jobdir: Path, nb: int, L0: float, periodic: int, ir: int, base, pint, preal, pdata,
                      timeout=300):
    """The reference's SMOOTH_2 / DIVER_UNIFORM or DIVER_FINA_GAS + VMESH run by
    the harness (mode 4) on MASCLET-like gas: (U1GCO, DIVERGCO)."""
    jobdir = Path(jobdir)
    jobdir.mkdir(parents=True, exist_ok=True)
    npa = len(pint) // 6
    with open(jobdir / "amr_base.bin", "wb") as fh:
        for a in base:
            fh.write(np.ascontiguousarray(a, np.float32).tobytes())
    (jobdir / "amr_pint.bin").write_bytes(np.ascontiguousarray(pint, np.int32).tobytes())
    (jobdir / "amr_preal.bin").write_bytes(np.ascontiguousarray(preal, np.float32).tobytes())
    (jobdir / "amr_patch.bin").write_bytes(np.ascontiguousarray(pdata, np.float32)[:max(len(pdata), 0)].tobytes())
    (jobdir / "amr.in").write_text(f"{ir} {npa}\n")
    (jobdir / "job.in").write_text(f"4\n{nb} {L0!r} {periodic}\n-0.6 10.0 0.25 0.0 1.0 1.0 100\n"
                                   f"0.678 0.31 0.0\n32 1 1 0\n")
    exe = REF_DIR / f"ref_harness_p{int(periodic)}.x"
    r = subprocess.run([str(exe), str(jobdir)], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"reference AMR harness failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
    n = nb << ir if ir >= 0 else nb >> (-ir)
    out = np.fromfile(jobdir / "out_amr.bin", np.float32)
    return out[: n ** 3].copy(), out[n ** 3: 2 * n ** 3].copy()


def run_reference_read_masclet(jobdir: Path, nb: int, it: int, npalev: int, nlevels: int, npmax: int, timeout=300):
    """The reference's READ_MASCLET (reader.f90:86-405) on <jobdir>/simu_masclet/
    files of iteration it: dict(ndxyz0, nl, npartt, gas = [U1G, U2G, U3G, U4G],
    patches = [(ints6, reals3, [U11G, U12G, U13G, U14G])], particles = 7 arrays)."""
    jobdir = Path(jobdir)
    (jobdir / "amr.in").write_text(f"{it} {npalev} {nlevels}\n")
    (jobdir / "job.in").write_text(f"5\n{nb} 1.0 0\n-0.6 10.0 0.25 0.0 1.0 1.0 100\n0.678 0.31 0.0\n32 1 1 {npmax}\n")
    exe = REF_DIR / "ref_harness_p0.x"
    r = subprocess.run([str(exe), str(jobdir.resolve())], capture_output=True, text=True, timeout=timeout,
                       cwd=str(jobdir))
    if r.returncode != 0:
        raise RuntimeError(f"reference READ_MASCLET failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
    head = np.fromfile(jobdir / "out_masclet_head.bin", np.int32)
    ndxyz0, nl, npartt, npatch = (int(v) for v in head[:4])
    n3 = nb ** 3
    gas = np.fromfile(jobdir / "out_masclet_gas.bin", np.float32)
    raw = (jobdir / "out_masclet_patch.bin").read_bytes()
    patches, off = [], 0
    for _ in range(npatch):
        ints = np.frombuffer(raw, np.int32, 6, off)
        reals = np.frombuffer(raw, np.float32, 3, off + 24)
        m = int(ints[0]) * int(ints[1]) * int(ints[2])
        f = np.frombuffer(raw, np.float32, 4 * m, off + 36)
        patches.append((ints.copy(), reals.copy(), [f[q * m:(q + 1) * m].copy() for q in range(4)]))
        off += 36 + 16 * m
    part = np.fromfile(jobdir / "out_masclet_part.bin", np.float32)
    return dict(ndxyz0=ndxyz0, nl=nl, npartt=npartt, gas=[gas[q * n3:(q + 1) * n3] for q in range(4)],
                patches=patches, particles=[part[q * npartt:(q + 1) * npartt] for q in range(7)])


def ref_available(periodic: int) -> bool:
    return (REF_DIR / f"ref_harness_p{int(periodic)}.x").exists()


def ref_fields_openmp() -> bool:
    """True when oracle/_ref was built with the field routines (particles.f90)
    under OpenMP, as the reference's own Makefile builds them."""
    return (REF_DIR / "omp_fields").exists()


def run_reference_indexx(jobdir: Path, keys: np.ndarray) -> np.ndarray:
    """INDEXX (num.f90:2-84) run by the reference routine itself (1-based output)."""
    jobdir = Path(jobdir)
    jobdir.mkdir(parents=True, exist_ok=True)
    keys = np.ascontiguousarray(keys, np.float32)
    (jobdir / "keys.bin").write_bytes(keys.tobytes())
    (jobdir / "job.in").write_text(f"3\n1 1.0 0\n0 0 0 0 0 0 1\n0 0 0\n0 0 0 {keys.size}\n")
    subprocess.run([str(REF_DIR / "ref_harness_p0.x"), str(jobdir)], check=True, timeout=300,
                   stdout=subprocess.DEVNULL)
    return np.fromfile(jobdir / "out_indexx.bin", np.int32)


def run_reference(jobdir: Path, mode: int, n: int, L0: float, periodic: int, params: VoidParams,
                  ache=0.678, omega0=0.31, zeta=0.0, k=32, vel_interp=1, dens_interp=1,
                  parts=None, fields=None, marcap=None, reqp=None, smask=None, rands=None, timeout=600) -> dict:
    """Run the reference routines (oracle/_ref) on one level and read the outputs.
    marcap (inner n^3 int32) and reqp (float32, ids 1..len) are a parent level's
    MARCAP and REQP (avism.f90:2123-2161); None runs the first level."""
    import resource

    jobdir = Path(jobdir)
    jobdir.mkdir(parents=True, exist_ok=True)
    for f in ("marcap.bin", "reqp.bin", "smask.bin", "rands.bin"):
        (jobdir / f).unlink(missing_ok=True)
    if rands is not None:  # survey mode: NRAND (int32), then x, y, z, m
        with open(jobdir / "rands.bin", "wb") as fh:
            fh.write(np.int32(rands[0].size).tobytes())
            for a in rands:
                fh.write(np.ascontiguousarray(a, np.float32).tobytes())
    if smask is not None:
        sm = np.ascontiguousarray(smask, np.int8)
        assert sm.size == n ** 3
        (jobdir / "smask.bin").write_bytes(sm.tobytes())
    if marcap is not None:
        mp = np.ascontiguousarray(marcap, np.int32)
        assert mp.size == n ** 3
        rq = np.zeros(params.nvoid_max, np.float32)
        if reqp is not None:
            rq[: len(reqp)] = np.asarray(reqp, np.float32)
        (jobdir / "marcap.bin").write_bytes(mp.tobytes())
        (jobdir / "reqp.bin").write_bytes(rq.tobytes())
    npart = 0
    if mode == 1:
        arrs = [np.ascontiguousarray(a, np.float32) for a in parts]
        npart = arrs[0].size
        with open(jobdir / "parts.bin", "wb") as fh:
            for a in arrs:
                fh.write(a.tobytes())
    else:
        with open(jobdir / "fields.bin", "wb") as fh:
            for a in fields:
                fh.write(np.ascontiguousarray(a, np.float32).tobytes())
    (jobdir / "job.in").write_text(
        f"{mode}\n{n} {L0!r} {periodic}\n"
        f"{params.dens_thre!r} {params.dens_thre2!r} {params.grad_thre!r} {params.div_thre!r} "
        f"{params.rmin!r} {params.rmin_sub!r} {params.nvoid_max}\n"
        f"{ache!r} {omega0!r} {zeta!r}\n{k} {vel_interp} {dens_interp} {npart}\n")

    def unlimit():  # the reference needs a large stack (run.sh:16); raise the soft limit to the hard one
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        try:
            resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))
        except (ValueError, OSError):
            pass

    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"), OMP_STACKSIZE="512m")
    exe = REF_DIR / f"ref_harness_p{int(periodic)}.x"
    subprocess.run([str(exe), str(jobdir)], check=True, timeout=timeout, preexec_fn=unlimit, env=env,
                   stdout=subprocess.DEVNULL)
    head = (jobdir / "out_head.txt").read_text().split()
    nvoid, nvoid2, nvoidt, low1, low2, nplus = (int(v) for v in head[:6])
    volt, meandens = float(head[6]), float(head[7])
    ne = low2 - low1 + 1
    out = dict(nvoid=nvoid, nvoid2=nvoid2, nvoidt=nvoidt, volt_clean=volt, meandens=meandens)
    fl = np.fromfile(jobdir / "out_fields.bin", np.float32)
    out["u1co"], out["div"] = fl[: n ** 3], fl[n ** 3:]
    out["marca"] = np.fromfile(jobdir / "out_marca.bin", np.int32)
    assert out["marca"].size == ne ** 3
    vel = np.fromfile(jobdir / "out_vel.bin", np.float32)
    out["u"], out["v"], out["w"] = vel[: n ** 3], vel[n ** 3: 2 * n ** 3], vel[2 * n ** 3:]
    if mode == 1:
        out["mass"] = np.fromfile(jobdir / "out_mass.bin", np.float32)
        if (jobdir / "out_rands.bin").exists():
            out["urand"] = np.fromfile(jobdir / "out_rands.bin", np.float32)
        pp = np.fromfile(jobdir / "out_parts.bin", np.float32)
        out["part_dens"], out["hpart"] = pp[:npart], pp[npart:]
    raw = open(jobdir / "out_cubes.bin", "rb").read()
    if nvoid > 0:
        ints = np.frombuffer(raw[: 13 * nvoid * 4], np.int32).reshape(13, nvoid)
        flts = np.frombuffer(raw[13 * nvoid * 4:], np.float32).reshape(19, nvoid)
        out["cubes_i"] = ints[:9].copy()
        out["father"], out["indice"], out["uvoid"], out["ncellv"] = (ints[9].copy(), ints[10].copy(),
                                                                     ints[11].copy(), ints[12].copy())
        out["cubes_f"] = flts[:7].copy()
        out["vf"] = flts[7:].copy()  # gxc..req in the order of orc_merge_post's vf[0:12]
    return out
