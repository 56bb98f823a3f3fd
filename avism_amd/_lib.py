"""ctypes bindings to libavism_hip.so (C ABI in include/avism.h).

The library is loaded from this package directory.  A missing library is an
error: the package has no CPU fallback for any stage.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libavism_hip.so"

AVISM_OK = 0
AVISM_HOST, AVISM_DEVICE, AVISM_HOST_ASYNC = 0, 1, 2
GROW_PLAIN, GROW_PROPOSAL, GROW_LEARNED = 0, 1, 2
CNN_NPARAM, CNN_NCLASS, CNN_CACHE, CNN_PATCH = 19064, 8, 7920, 2187

ERRORS = {-1: "AVISM_ERR_ARG", -2: "AVISM_ERR_HIP", -3: "AVISM_ERR_INPUT", -4: "AVISM_ERR_LIMIT",
          -5: "AVISM_ERR_STATE", -6: "AVISM_ERR_UNSUPPORTED"}


class AvismError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class GridInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("periodic", C.c_int32), ("nplus", C.c_int32), ("low1", C.c_int32),
                ("low2", C.c_int32), ("next", C.c_int32), ("L0", C.c_float), ("lado0plus", C.c_float),
                ("dx", C.c_float), ("rx1", C.c_float)]


class VoidParams(C.Structure):
    _fields_ = [("dens_thre", C.c_float), ("dens_thre2", C.c_float), ("grad_thre", C.c_float),
                ("div_thre", C.c_float), ("rmin", C.c_float), ("rmin_sub", C.c_float),
                ("nvoid_max", C.c_int32)]


class RefineInfo(C.Structure):
    _fields_ = [("depth", C.c_int32), ("n_level", C.c_int32), ("natlas", C.c_int32), ("atlas", C.c_int32),
                ("side", C.c_int32), ("nwindows", C.c_int32), ("window_cells", C.c_int64), ("dx", C.c_float),
                ("rx1_level", C.c_float), ("dx_atlas", C.c_float), ("rx1_atlas", C.c_float),
                ("meandens", C.c_float)]


class LevelStats(C.Structure):
    _fields_ = [("ncand", C.c_int64), ("ncubes", C.c_int32), ("nvoid2", C.c_int32),
                ("nvoidt", C.c_int32), ("nperiodic", C.c_int32), ("volt_clean", C.c_double),
                ("meandens", C.c_float), ("ff", C.c_float)]


EXPORTS = [
    "avism_create", "avism_destroy", "avism_last_error", "avism_set_timing", "avism_get_timings",
    "avism_set_grid", "avism_get_grid", "avism_set_cosmology", "avism_set_void_params",
    "avism_set_parent_map", "avism_set_parent_radii", "avism_set_mask", "avism_next_level", "avism_load_particles",
    "avism_particles_to_grid",
    "avism_particles_to_density",
    "avism_load_randoms",
    "avism_survey_density",
    "avism_load_grid_fields", "avism_load_void_fields", "avism_get_fields", "avism_get_particle_fields", "avism_find_voids",
    "avism_set_proposals", "avism_cnn_propose", "avism_get_candidates", "avism_get_cubes", "avism_indexx",
    "avism_merge_and_post", "avism_get_catalogue", "avism_get_marca", "avism_run_level",
    "avism_cnn_set_params", "avism_cnn_patches", "avism_cnn_forward", "avism_cnn_backward", "avism_cnn_adam",
    "avism_sync", "avism_get_counters", "avism_indexx_noalloc",
    "avism_batch_set", "avism_batch_load", "avism_batch_fields", "avism_batch_voids", "avism_batch_patch",
    "avism_batch_get_fields", "avism_batch_info",
    "avism_refine_plan", "avism_refine_load", "avism_refine_windows", "avism_refine_pop",
    "avism_amr_set", "avism_amr_fields",
]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m avism_amd.build` "
                           "(there is no CPU fallback)")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 (same
    # soname as ROCm's).  Importing torch first lets the library bind to that
    # copy; loading ROCm's copy first would leave torch with a second runtime
    # that finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
    sig = {
        "avism_create": ([C.POINTER(vp), i32], C.c_int),
        "avism_destroy": ([vp], C.c_int),
        "avism_last_error": ([vp], C.c_char_p),
        "avism_set_timing": ([vp, i32], C.c_int),
        "avism_get_timings": ([vp, C.c_char_p, i32, vp, C.POINTER(i32)], C.c_int),
        "avism_set_grid": ([vp, i32, f32, i32], C.c_int),
        "avism_get_grid": ([vp, C.POINTER(GridInfo), vp], C.c_int),
        "avism_set_cosmology": ([vp, f32, f32, f32], C.c_int),
        "avism_set_void_params": ([vp, C.POINTER(VoidParams)], C.c_int),
        "avism_set_parent_map": ([vp, vp, i32], C.c_int),
        "avism_set_parent_radii": ([vp, vp, i32, i32], C.c_int),
        "avism_set_mask": ([vp, vp, i32], C.c_int),
        "avism_next_level": ([vp], C.c_int),
        "avism_load_particles": ([vp, i64, vp, vp, vp, vp, vp, vp, vp, i32], C.c_int),
        "avism_particles_to_grid": ([vp, i32, i32, i32], C.c_int),
        "avism_particles_to_density": ([vp, i32, i32], C.c_int),
        "avism_load_randoms": ([vp, C.c_int64, vp, vp, vp, vp, i32], C.c_int),
        "avism_survey_density": ([vp, i32, i32], C.c_int),
        "avism_load_grid_fields": ([vp, vp, vp, vp, vp, i32], C.c_int),
        "avism_load_void_fields": ([vp, vp, vp, f32, i32], C.c_int),
        "avism_get_fields": ([vp, vp, vp, vp, vp, vp, i32], C.c_int),
        "avism_get_particle_fields": ([vp, vp, vp], C.c_int),
        "avism_find_voids": ([vp, i32, C.POINTER(i32)], C.c_int),
        "avism_set_proposals": ([vp, vp, i64, i32], C.c_int),
        "avism_cnn_propose": ([vp, vp, i64, i32], C.c_int),
        "avism_get_candidates": ([vp, vp, vp, i64, C.POINTER(i64)], C.c_int),
        "avism_get_cubes": ([vp, vp, vp, vp], C.c_int),
        "avism_indexx": ([vp, vp, i64, vp, i32], C.c_int),
        "avism_merge_and_post": ([vp, C.POINTER(i32)], C.c_int),
        "avism_get_catalogue": ([vp, vp, vp, vp, vp, C.POINTER(LevelStats)], C.c_int),
        "avism_get_marca": ([vp, vp, i32, i32], C.c_int),
        "avism_run_level": ([vp, i32, i32, i32, i32, C.POINTER(LevelStats)], C.c_int),
        "avism_cnn_set_params": ([vp, vp, i32], C.c_int),
        "avism_cnn_patches": ([vp, vp, i64, vp], C.c_int),
        "avism_cnn_forward": ([vp, vp, i64, vp, vp, vp], C.c_int),
        "avism_cnn_backward": ([vp, vp, i64, vp, vp, vp, vp, vp, vp], C.c_int),
        "avism_cnn_adam": ([vp, vp, vp, vp, vp, f32, i32], C.c_int),
        "avism_sync": ([vp], C.c_int),
        "avism_get_counters": ([vp, vp, i32], C.c_int),
        "avism_indexx_noalloc": ([vp, vp, i64, vp], C.c_int),
        "avism_batch_set": ([vp, i32, vp, vp, vp, vp], C.c_int),
        "avism_batch_load": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, i32], C.c_int),
        "avism_batch_fields": ([vp, i32], C.c_int),
        "avism_batch_voids": ([vp, i32, i32], C.c_int),
        "avism_batch_patch": ([vp, i32, C.POINTER(vp)], C.c_int),
        "avism_batch_get_fields": ([vp, i32, vp, vp, vp, vp, vp, vp, vp, i32], C.c_int),
        "avism_batch_info": ([vp, C.c_char_p, i32, C.POINTER(i64), vp], C.c_int),
        "avism_refine_plan": ([vp, i32, i64, C.POINTER(i32)], C.c_int),
        "avism_refine_load": ([vp, i32, i32], C.c_int),
        "avism_refine_windows": ([vp, vp, i32, C.POINTER(i32), C.POINTER(RefineInfo)], C.c_int),
        "avism_refine_pop": ([vp], C.c_int),
        "avism_amr_set": ([vp, i32, vp, i32, vp, vp, vp, i32], C.c_int),
        "avism_amr_fields": ([vp, i32, i32, i32], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data_as(C.c_void_p)
    # torch tensor (device or host)
    return C.c_void_p(a.data_ptr())


class Context:
    """One GPU context: owns device memory and a stream (avism_create)."""

    def __init__(self, device: int = 0):
        L = lib()
        self._h = C.c_void_p()
        rc = L.avism_create(C.byref(self._h), int(device))
        if rc != AVISM_OK:
            raise AvismError(rc, "avism_create failed (no HIP device?)")

    def _chk(self, rc: int):
        if rc != AVISM_OK:
            raise AvismError(rc, lib().avism_last_error(self._h).decode())

    def close(self):
        if self._h and not getattr(self, "_borrowed", False):
            lib().avism_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up
    def set_grid(self, n: int, L0: float, periodic: int) -> GridInfo:
        self._chk(lib().avism_set_grid(self._h, int(n), float(L0), int(periodic)))
        return self.grid()

    def grid(self, with_rad=False):
        gi = GridInfo()
        self._chk(lib().avism_get_grid(self._h, C.byref(gi), None))
        if with_rad:
            rad = np.zeros(gi.next, np.float32)
            self._chk(lib().avism_get_grid(self._h, C.byref(gi), _ptr(rad)))
            return gi, rad
        return gi

    def set_cosmology(self, ache=0.678, omega0=0.31, zeta=0.0):
        self._chk(lib().avism_set_cosmology(self._h, float(ache), float(omega0), float(zeta)))

    def set_void_params(self, dens_thre=-0.6, dens_thre2=10.0, grad_thre=0.25, div_thre=0.0,
                        rmin=1.0, rmin_sub=1.0, nvoid_max=30000000):
        p = VoidParams(dens_thre, dens_thre2, grad_thre, div_thre, rmin, rmin_sub, int(nvoid_max))
        self._chk(lib().avism_set_void_params(self._h, C.byref(p)))

    def set_parent_map(self, marcap_inner):
        """MARCAP of the previous level on the inner grid (None: -1 everywhere)."""
        if marcap_inner is None:
            self._chk(lib().avism_set_parent_map(self._h, None, AVISM_HOST))
            return
        a = np.ascontiguousarray(marcap_inner, np.int32).ravel()
        gi = self.grid()
        if a.size != gi.n ** 3:
            raise ValueError(f"parent map has {a.size} cells, the grid {gi.n ** 3}")
        self._chk(lib().avism_set_parent_map(self._h, _ptr(a), AVISM_HOST))

    def set_parent_radii(self, reqp):
        """REQP of the previous level, reqp[id-1] for parent ids 1..len (None: no test)."""
        if reqp is None:
            self._chk(lib().avism_set_parent_radii(self._h, None, 0, AVISM_HOST))
            return
        a = np.ascontiguousarray(reqp, np.float32).ravel()
        self._chk(lib().avism_set_parent_radii(self._h, _ptr(a) if a.size else _ptr(np.zeros(1, np.float32)),
                                               int(a.size), AVISM_HOST))

    def set_mask(self, smask_inner):
        """SMASK on the inner grid (int8, 1 = observed; None: no mask)."""
        if smask_inner is None:
            self._chk(lib().avism_set_mask(self._h, None, AVISM_HOST))
            return
        a = np.ascontiguousarray(smask_inner, np.int8).ravel()
        gi = self.grid()
        if a.size != gi.n ** 3:
            raise ValueError(f"mask has {a.size} cells, the grid {gi.n ** 3}")
        self._mask = a
        self._chk(lib().avism_set_mask(self._h, _ptr(a), AVISM_HOST))

    def next_level(self) -> GridInfo:
        """Void-in-void hand-off: MARCAP, REQP and the grid of 2N (avism.f90:2123-2161)."""
        self._chk(lib().avism_next_level(self._h))
        return self.grid()

    # ---- refined levels in windows around the parent voids
    def refine_plan(self, margin: int = 8, max_cells: int = 0) -> int:
        """Plan the next level's windows from this level's voids; the atlas count."""
        na = C.c_int32()
        self._chk(lib().avism_refine_plan(self._h, int(margin), int(max_cells), C.byref(na)))
        return na.value

    def refine_load(self, atlas: int, k: int = 32):
        """Atlas `atlas` of the top plan becomes this context's grid, with its fields."""
        self._chk(lib().avism_refine_load(self._h, int(atlas), int(k)))

    def refine_windows(self):
        """(windows, info): windows[i] = parent, side, origin x, y, z in the level's
        cells, origin x, y, z in the atlas; info: RefineInfo as a dict."""
        n = C.c_int32()
        info = RefineInfo()
        self._chk(lib().avism_refine_windows(self._h, None, 0, C.byref(n), C.byref(info)))
        tab = np.zeros((max(n.value, 1), 8), np.int32)
        self._chk(lib().avism_refine_windows(self._h, _ptr(tab), int(n.value), C.byref(n), None))
        return tab[: n.value], {f: getattr(info, f) for f, _ in RefineInfo._fields_}

    def refine_pop(self):
        self._chk(lib().avism_refine_pop(self._h))

    # ---- MASCLET AMR input
    def amr_set(self, nhyx: int, base, pint=None, preal=None, pdata=None):
        """The gas of a MASCLET snapshot: base = [delta, vx, vy, vz] on nhyx^3
        cells; level-1 patches as (pint, preal, pdata) (synth.amr_pack)."""
        b = np.ascontiguousarray(np.concatenate([np.ravel(a) for a in base]), np.float32)
        npa = 0 if pint is None else len(pint) // 6
        arrs = [None if a is None else np.ascontiguousarray(a, t)
                for a, t in ((pint, np.int32), (preal, np.float32), (pdata, np.float32))]
        self._keep_amr = (b, *arrs)
        self._chk(lib().avism_amr_set(self._h, int(nhyx), _ptr(b), int(npa), *[_ptr(a) for a in arrs], AVISM_HOST))

    def amr_fields(self, ir: int, flag_dens: int = 2, flag_div: int = 2):
        """U1CO and DIVERCO of level ir from the gas (and the dark matter of the
        last particles_to_grid on this grid, per FLAG_DENS / FLAG_DIV)."""
        self._chk(lib().avism_amr_fields(self._h, int(ir), int(flag_dens), int(flag_div)))

    def set_timing(self, on: bool):
        self._chk(lib().avism_set_timing(self._h, int(bool(on))))

    def timings(self):
        """(stage, ms) of every stage timed since the last call, in order."""
        cap, nbytes = 256, 1 << 14
        while True:
            n = C.c_int32(cap)
            ms = np.zeros(cap, np.float32)
            names = C.create_string_buffer(nbytes)
            rc = lib().avism_get_timings(self._h, names, len(names), _ptr(ms), C.byref(n))
            if rc == -4:  # AVISM_ERR_LIMIT: buffers too small, nothing consumed
                cap, nbytes = max(cap, n.value), max(2 * nbytes, 48 * max(cap, n.value))
                continue
            self._chk(rc)
            keys = names.value.decode().split(";")[: n.value]
            assert len(keys) == n.value
            return list(zip(keys, ms[: n.value].tolist()))

    # ---- field stage
    def load_particles(self, x, y, z, vx, vy, vz, m, where=None):
        """vx, vy, vz may all be None: particles without velocities."""
        arrs = [x, y, z, vx, vy, vz, m]
        if where is None:
            where = AVISM_HOST if isinstance(x, np.ndarray) else (
                AVISM_DEVICE if getattr(x, "is_cuda", False) else AVISM_HOST)
        if where == AVISM_HOST:
            arrs = [None if a is None else np.ascontiguousarray(a if isinstance(a, np.ndarray) else a.numpy(),
                                                                np.float32) for a in arrs]
        n = int(arrs[0].shape[0]) if isinstance(arrs[0], np.ndarray) else int(arrs[0].numel())
        self._keep = arrs
        self._chk(lib().avism_load_particles(self._h, n, *[None if a is None else _ptr(a) for a in arrs],
                                             int(where)))

    def particles_to_grid(self, k=32, vel_interp=1, dens_interp=1):
        self._chk(lib().avism_particles_to_grid(self._h, int(k), int(vel_interp), int(dens_interp)))

    def particles_to_density(self, k=32, dens_interp=1):
        """FLAG_VEL_AVAILABLE = 0: density only, DIVERCO from linear theory."""
        self._chk(lib().avism_particles_to_density(self._h, int(k), int(dens_interp)))

    def load_randoms(self, x, y, z, m):
        """Survey randoms (host arrays)."""
        arrs = [np.ascontiguousarray(a, np.float32) for a in (x, y, z, m)]
        self._keep_rand = arrs
        self._chk(lib().avism_load_randoms(self._h, int(arrs[0].size), *[_ptr(a) for a in arrs], AVISM_HOST))

    def survey_density(self, k=32, dens_interp=1):
        """FLAG_SURVEY = 1: D/(alpha R) density, linear-theory DIVERCO."""
        self._chk(lib().avism_survey_density(self._h, int(k), int(dens_interp)))

    def load_grid_fields(self, delta, vx, vy, vz):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (delta, vx, vy, vz)]
        self._chk(lib().avism_load_grid_fields(self._h, *[_ptr(a) for a in arrs], AVISM_HOST))

    def load_void_fields(self, u1co, div, meandens):
        a = np.ascontiguousarray(u1co, np.float32)
        b = np.ascontiguousarray(div, np.float32)
        self._chk(lib().avism_load_void_fields(self._h, _ptr(a), _ptr(b), float(meandens), AVISM_HOST))

    def fields(self):
        n = self.grid().n
        out = {k: np.zeros(n ** 3, np.float32) for k in ("u1co", "div", "u", "v", "w")}
        self._chk(lib().avism_get_fields(self._h, _ptr(out["u1co"]), _ptr(out["div"]), _ptr(out["u"]),
                                         _ptr(out["v"]), _ptr(out["w"]), AVISM_HOST))
        return out

    def fields_async(self, u1co=None, div=None):
        """Queue the download of U1CO and DIVERCO into caller-owned pinned host
        tensors (torch, pin_memory=True) behind the work queued so far; they are
        complete after sync()."""
        n3 = self.grid().n ** 3
        for t in (u1co, div):
            if t is not None and (not t.is_pinned() or t.numel() != n3 or t.element_size() != 4
                                  or not t.is_contiguous()):
                raise ValueError("fields_async needs contiguous pinned 4-byte host tensors of n^3 cells")
        self._chk(lib().avism_get_fields(self._h, _ptr(u1co), _ptr(div), None, None, None, AVISM_HOST_ASYNC))

    def sync(self):
        self._chk(lib().avism_sync(self._h))

    def counters(self):
        """Work counters: DDENS (particle, cell) pairs, PPART queries, VVEL
        queries; of the last find_voids, the face cells the growth examined and
        the candidates grown by one wave each."""
        out = np.zeros(8, np.int64)
        self._chk(lib().avism_get_counters(self._h, _ptr(out), 8))
        return dict(ddens_pairs=int(out[0]), ppart_queries=int(out[1]), vvel_queries=int(out[2]),
                    growth_cells=int(out[3]), growth_wave_candidates=int(out[4]))

    def particle_fields(self, np_total: int):
        pd = np.zeros(np_total, np.float32)
        hp = np.zeros(np_total, np.float32)
        self._chk(lib().avism_get_particle_fields(self._h, _ptr(pd), _ptr(hp)))
        return pd, hp

    # ---- void stage
    def find_voids(self, growth_mode=GROW_PLAIN) -> int:
        n = C.c_int32()
        self._chk(lib().avism_find_voids(self._h, int(growth_mode), C.byref(n)))
        return n.value

    def set_proposals(self, r_hat):
        a = np.ascontiguousarray(r_hat, np.int8)
        self._chk(lib().avism_set_proposals(self._h, _ptr(a), int(a.size), AVISM_HOST))

    def candidates(self):
        n = C.c_int64()
        self._chk(lib().avism_get_candidates(self._h, None, None, 0, C.byref(n)))
        lin = np.zeros(max(n.value, 1), np.int64)
        r = np.zeros(max(n.value, 1), np.int32)
        self._chk(lib().avism_get_candidates(self._h, _ptr(lin), _ptr(r), n.value, C.byref(n)))
        return lin[: n.value], r[: n.value]

    def cubes(self, ncubes: int):
        ci = np.zeros(9 * max(ncubes, 1), np.int32)
        cf = np.zeros(7 * max(ncubes, 1), np.float32)
        fa = np.zeros(max(ncubes, 1), np.int32)
        self._chk(lib().avism_get_cubes(self._h, _ptr(ci), _ptr(cf), _ptr(fa)))
        return ci.reshape(9, -1)[:, :ncubes], cf.reshape(7, -1)[:, :ncubes], fa[:ncubes]

    def indexx(self, keys):
        a = np.ascontiguousarray(keys, np.float32)
        out = np.zeros(max(a.size, 1), np.int32)
        self._chk(lib().avism_indexx(self._h, _ptr(a), int(a.size), _ptr(out), AVISM_HOST))
        return out[: a.size]

    # ---- merge + post
    def merge_and_post(self) -> int:
        n = C.c_int32()
        self._chk(lib().avism_merge_and_post(self._h, C.byref(n)))
        return n.value

    def catalogue(self, ncubes: int):
        nv = max(ncubes, 1)
        ind = np.zeros(nv, np.int32)
        uv = np.zeros(nv, np.int32)
        nc = np.zeros(nv, np.int32)
        vf = np.zeros(12 * nv, np.float32)
        st = LevelStats()
        self._chk(lib().avism_get_catalogue(self._h, _ptr(ind), _ptr(uv), _ptr(nc), _ptr(vf), C.byref(st)))
        return dict(indice=ind[:ncubes], uvoid=uv[:ncubes], ncellv=nc[:ncubes],
                    vf=vf.reshape(12, nv)[:, :ncubes], stats=st)

    def stats(self) -> LevelStats:
        st = LevelStats()
        self._chk(lib().avism_get_catalogue(self._h, None, None, None, None, C.byref(st)))
        return st

    def marca(self, inner_only=False, out=None):
        """MARCA (ext grid, or the inner grid for the map file); out: a
        caller-owned host array or tensor of that many int32 to fill."""
        gi = self.grid()
        n = gi.n ** 3 if inner_only else gi.next ** 3
        if out is None:
            out = np.zeros(n, np.int32)
        elif (out.numel() if hasattr(out, "numel") else out.size) != n:
            raise ValueError("MARCA output has the wrong size")
        self._chk(lib().avism_get_marca(self._h, _ptr(out), int(bool(inner_only)), AVISM_HOST))
        return out

    # ---- batches of independent patches (each patch: one grid, its particles and mask)
    def batch_set(self, n, L0, nps, periodic=None):
        """Layout: patch p has n[p]^3 cells over a box of L0[p] and nps[p] particles."""
        n = np.ascontiguousarray(n, np.int32)
        L0 = np.ascontiguousarray(L0, np.float32)
        nps = np.ascontiguousarray(nps, np.int64)
        per = None if periodic is None else np.ascontiguousarray(periodic, np.int32)
        self._batch_layout = (n.copy(), L0.copy(), nps.copy())
        self._chk(lib().avism_batch_set(self._h, int(n.size), _ptr(n), _ptr(L0), _ptr(per), _ptr(nps)))

    def batch_load(self, x, y, z, vx, vy, vz, m, smask=None, where=None):
        """Every patch's particles back to back (numpy or torch, host or device);
        smask: the patches' inner masks back to back (int8) or None."""
        arrs = [x, y, z, vx, vy, vz, m]
        if where is None:
            where = AVISM_DEVICE if getattr(x, "is_cuda", False) else AVISM_HOST
        if where == AVISM_HOST:
            arrs = [None if a is None else np.ascontiguousarray(a if isinstance(a, np.ndarray) else a.numpy(),
                                                                np.float32) for a in arrs]
            if smask is not None:
                smask = np.ascontiguousarray(smask if isinstance(smask, np.ndarray) else smask.numpy(), np.int8)
        self._keep_batch = (arrs, smask)
        self._chk(lib().avism_batch_load(self._h, *[_ptr(a) for a in arrs], _ptr(smask), int(where)))

    def batch_fields(self, k=32):
        self._chk(lib().avism_batch_fields(self._h, int(k)))

    def batch_voids(self, growth_mode=GROW_PLAIN, threads=8):
        self._chk(lib().avism_batch_voids(self._h, int(growth_mode), int(threads)))

    def batch_patch(self, p: int) -> "Context":
        """Patch p's context (owned by this one) for the usual getters."""
        h = C.c_void_p()
        self._chk(lib().avism_batch_patch(self._h, int(p), C.byref(h)))
        sub = Context.__new__(Context)
        sub._h = h
        sub._borrowed = True
        sub._parent = self
        return sub

    def batch_fields_of(self, p: int):
        """Patch p's inner fields and, in its input order, PART_DENS and HPART."""
        n, _, nps = self._batch_layout
        n3, npp = int(n[p]) ** 3, int(nps[p])
        out = {k: np.zeros(n3, np.float32) for k in ("u1co", "div", "u", "v", "w")}
        out["part_dens"] = np.zeros(npp, np.float32)
        out["hpart"] = np.zeros(npp, np.float32)
        self._chk(lib().avism_batch_get_fields(self._h, int(p), *[_ptr(out[k]) for k in
                                                                  ("u1co", "div", "u", "v", "w", "part_dens",
                                                                   "hpart")], AVISM_HOST))
        return out

    def batch_info(self):
        import json
        buf = C.create_string_buffer(1 << 16)
        launches = C.c_int64()
        nb = len(self._batch_layout[0])
        md = np.zeros(nb, np.float32)
        self._chk(lib().avism_batch_info(self._h, buf, len(buf), C.byref(launches), _ptr(md)))
        txt = buf.value.decode()
        return dict(tuning=json.loads(txt) if txt else None, graph_launches=int(launches.value), meandens=md)

    def run_level(self, k=32, vel_interp=1, dens_interp=1, growth_mode=GROW_PLAIN) -> LevelStats:
        st = LevelStats()
        self._chk(lib().avism_run_level(self._h, int(k), int(vel_interp), int(dens_interp),
                                        int(growth_mode), C.byref(st)))
        return st

    # ---- half-width proposal network (device tensors; shapes checked here
    # before any launch)
    @staticmethod
    def _dev(t, dtype, shape, name):
        import torch
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise ValueError(f"{name} must be a device tensor")
        if t.dtype != dtype or tuple(t.shape) != tuple(shape) or not t.is_contiguous():
            raise ValueError(f"{name}: expected contiguous {dtype} {tuple(shape)}, got {t.dtype} "
                             f"{tuple(t.shape)}")
        return C.c_void_p(t.data_ptr())

    def cnn_set_params(self, params):
        """Network used by GROW_LEARNED (numpy fp32 or device tensor, CNN_NPARAM values)."""
        if isinstance(params, np.ndarray):
            a = np.ascontiguousarray(params, np.float32)
            if a.size != CNN_NPARAM:
                raise ValueError("wrong parameter count")
            self._chk(lib().avism_cnn_set_params(self._h, _ptr(a), AVISM_HOST))
        else:
            import torch
            p = self._dev(params, torch.float32, (CNN_NPARAM,), "params")
            self._chk(lib().avism_cnn_set_params(self._h, p, AVISM_DEVICE))

    def cnn_propose(self) -> np.ndarray:
        """The installed network's proposal classes for the current candidates
        (processing order), as GROW_LEARNED computes them."""
        nc = C.c_int64()
        self._chk(lib().avism_get_candidates(self._h, None, None, 0, C.byref(nc)))
        out = np.zeros(nc.value, np.int8)
        self._chk(lib().avism_cnn_propose(self._h, _ptr(out), int(nc.value), AVISM_HOST))
        return out

    def cnn_patches(self, ranks):
        """bf16 feature patches [n, 3, 9, 9, 9] (as int16) of candidates at processing ranks."""
        import torch
        ranks = ranks.to(torch.int64).contiguous()
        n = int(ranks.numel())
        nc = C.c_int64()
        self._chk(lib().avism_get_candidates(self._h, None, None, 0, C.byref(nc)))
        if n and (int(ranks.min()) < 0 or int(ranks.max()) >= nc.value):
            raise ValueError("candidate rank out of range")
        out = torch.empty((n, 3, 9, 9, 9), dtype=torch.int16, device=ranks.device)
        if n:
            self._chk(lib().avism_cnn_patches(self._h, self._dev(ranks, torch.int64, (n,), "ranks"), n,
                                              C.c_void_p(out.data_ptr())))
        return out

    def cnn_forward(self, params, patches, want_cache=True):
        import torch
        n = int(patches.shape[0])
        p = self._dev(params, torch.float32, (CNN_NPARAM,), "params")
        x = self._dev(patches, torch.int16, (n, 3, 9, 9, 9), "patches")
        logits = torch.empty((n, CNN_NCLASS), dtype=torch.float32, device=patches.device)
        cache = torch.empty((n, CNN_CACHE), dtype=torch.int16, device=patches.device) if want_cache else None
        self._chk(lib().avism_cnn_forward(self._h, p, n, x, C.c_void_p(logits.data_ptr()),
                                          C.c_void_p(cache.data_ptr()) if cache is not None else None))
        return logits, cache

    def cnn_backward(self, params, patches, cache, logits, labels):
        import torch
        n = int(patches.shape[0])
        p = self._dev(params, torch.float32, (CNN_NPARAM,), "params")
        x = self._dev(patches, torch.int16, (n, 3, 9, 9, 9), "patches")
        ca = self._dev(cache, torch.int16, (n, CNN_CACHE), "cache")
        lg = self._dev(logits, torch.float32, (n, CNN_NCLASS), "logits")
        lb = self._dev(labels, torch.int32, (n,), "labels")
        grads = torch.empty(CNN_NPARAM, dtype=torch.float32, device=patches.device)
        loss = torch.empty(1, dtype=torch.float32, device=patches.device)
        self._chk(lib().avism_cnn_backward(self._h, p, n, x, ca, lg, lb, C.c_void_p(grads.data_ptr()),
                                           C.c_void_p(loss.data_ptr())))
        return grads, loss

    def cnn_adam(self, params, grads, m, v, lr, step):
        import torch
        ptrs = [self._dev(t, torch.float32, (CNN_NPARAM,), nm)
                for t, nm in ((params, "params"), (grads, "grads"), (m, "m"), (v, "v"))]
        self._chk(lib().avism_cnn_adam(self._h, *ptrs, float(lr), int(step)))
