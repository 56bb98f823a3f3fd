"""Host-side checks that need no GPU: the C-ABI library loads and exports every
entry point include/avism.h declares, it fails loudly without a device, and the
voids.dat reader/writer keeps the reference's positional layout."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF_VOIDS_DAT = Path("/root/reference/config/voids.dat")


def header_functions():
    text = (ROOT / "include" / "avism.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(avism_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = header_functions()
    for required in ("avism_create", "avism_destroy", "avism_last_error", "avism_set_grid",
                     "avism_particles_to_grid", "avism_find_voids", "avism_merge_and_post",
                     "avism_get_catalogue", "avism_get_marca", "avism_cnn_forward", "avism_cnn_backward"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from avism_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == header_functions()


def test_binding_signatures_load():
    from avism_amd import _lib
    L = _lib.lib()
    for n in _lib.EXPORTS:
        assert getattr(L, n).argtypes is not None, n


def test_create_fails_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from avism_amd._lib import AvismError, Context
    with pytest.raises(AvismError):
        Context(0)


def test_voids_dat_roundtrip(tmp_path):
    from avism_amd.config import VoidsConfig, read_voids_dat, write_voids_dat
    cfg = VoidsConfig(ncox=512, ncoy=512, ncoz=512, lado0=1000.0, dens_thre=-0.7, kneighbours=32,
                      nvoid_max=1234567, massdm=3.5e9)
    p = tmp_path / "voids.dat"
    write_voids_dat(cfg, p)
    assert len(p.read_text().splitlines()) == 73
    assert read_voids_dat(p) == cfg


@pytest.mark.skipif(not REF_VOIDS_DAT.exists(), reason="reference configuration not mounted")
def test_reads_reference_voids_dat(tmp_path):
    from avism_amd.config import read_voids_dat, write_voids_dat
    ref = read_voids_dat(REF_VOIDS_DAT)
    # thresholds of the shipped configuration (config/voids.dat:23-33)
    assert (ref.dens_thre, ref.dens_thre2, ref.grad_thre, ref.div_thre, ref.rmin) == (-0.6, 10.0, 0.25, 0.0, 1.0)
    p = tmp_path / "again.dat"
    write_voids_dat(ref, p)
    assert read_voids_dat(p) == ref
    assert len(p.read_text().splitlines()) == len(REF_VOIDS_DAT.read_text().splitlines())


def test_validate_rejects_what_the_reference_stops_on():
    from avism_amd.config import ConfigError, VoidsConfig
    VoidsConfig().validate()
    for bad in (dict(flag_data=4), dict(flag_vel_interp=2), dict(flag_survey=1, flag_period=1),
                dict(kneighbours=-1)):
        with pytest.raises(ConfigError):
            VoidsConfig(**bad).validate()


def test_binding_constants_match_header():
    from avism_amd import _lib
    text = (ROOT / "include" / "avism.h").read_text()
    defs = dict(re.findall(r"#define\s+(AVISM_[A-Z0-9_]+)\s+(-?\d+)", text))
    assert int(defs["AVISM_CNN_NPARAM"]) == _lib.CNN_NPARAM
    assert int(defs["AVISM_CNN_NCLASS"]) == _lib.CNN_NCLASS
    assert int(defs["AVISM_CNN_CACHE"]) == _lib.CNN_CACHE == 343 * 16 + 125 * 16 + 27 * 16
    assert int(defs["AVISM_CNN_PATCH"]) == _lib.CNN_PATCH == 3 * 9 ** 3
    assert (int(defs["AVISM_GROW_PLAIN"]), int(defs["AVISM_GROW_PROPOSAL"]), int(defs["AVISM_GROW_LEARNED"])) == (
        _lib.GROW_PLAIN, _lib.GROW_PROPOSAL, _lib.GROW_LEARNED)
    import cnn_ref
    assert cnn_ref.NPARAM == _lib.CNN_NPARAM


def test_fortran_host_builds_against_the_header(tmp_path):
    """The ISO_C_BINDING module compiles (amdflang or gfortran, when present)."""
    import shutil
    import subprocess
    fc = shutil.which("amdflang") or shutil.which("gfortran")
    if fc is None:
        pytest.skip("no Fortran compiler")
    r = subprocess.run([fc, "-c", str(ROOT / "host" / "avism_c.f90"), "-o", str(tmp_path / "a.o")],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
