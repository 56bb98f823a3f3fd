"""voids.dat reader/writer.

The reference reads ``config/voids.dat`` line by line with list-directed
``READ(1,*)`` statements (avism.f90:173-254): every value line is preceded by one
label line, and three header lines open each block.  This module keeps that exact
positional layout (73 lines, matching config/voids.dat:1-73) so a configuration
written here is read unchanged by the reference and vice versa.
"""
from __future__ import annotations

import dataclasses
import re
from pathlib import Path
from typing import List

__all__ = ["VoidsConfig", "read_voids_dat", "write_voids_dat", "ConfigError"]


class ConfigError(ValueError):
    """Raised where the reference would STOP on a bad voids.dat value."""


@dataclasses.dataclass
class VoidsConfig:
    # general block (avism.f90:184-197)
    first: int = 1
    last: int = 1
    every: int = 1
    ncox: int = 128
    ncoy: int = 128
    ncoz: int = 128
    levmin: int = 0
    levmax: int = 0
    ache: float = 0.678
    omega0: float = 0.31
    lado0: float = 147.5
    flag_period: int = 1
    # void finder block (avism.f90:201-213)
    nvoid_max: int = 30000000
    dens_thre: float = -0.6
    dens_thre2: float = 10.0
    grad_thre: float = 0.25
    div_thre: float = 0.0
    rmin: float = 1.0
    rmin_sub: float = 1.0
    # data block (avism.f90:217-227)
    flag_data: int = 1
    nhyx: int = 128
    nhyy: int = 128
    nhyz: int = 128
    files_per_snap: int = 16
    parttypex: int = 2
    massdm: float = 4.52e10
    flag_gadget_reader: int = 1
    flag_survey: int = 0
    # particle block (avism.f90:231-239)
    partired: int = 800000000
    flag_vel_available: int = 1
    flag_vel_interp: int = 1
    flag_dens_interp: int = 1
    kneighbours: int = 256
    # output block (avism.f90:243-244)
    flag_write_cubes: int = 0
    # MASCLET block (avism.f90:251-255)
    npalev: int = 100000
    nlevels: int = 10
    namrx: int = 64
    namry: int = 64
    namrz: int = 64
    flag_div: int = 1
    flag_dens: int = 1

    def validate(self) -> None:
        """The consistency checks of avism.f90:260-330 that apply to this path."""
        if self.flag_data < 0 or self.flag_data > 3:
            raise ConfigError("FLAG_DATA must be 0,1,2,3")
        if self.kneighbours < 0:
            raise ConfigError("KNEIGHBOURS must be positive")
        if self.flag_vel_available not in (0, 1):
            raise ConfigError("FLAG_VEL_AVAILABLE must be 0 or 1")
        if self.flag_vel_interp not in (0, 1):
            raise ConfigError("FLAG_VEL_INTERP must be 0 or 1")
        if self.flag_dens_interp not in (0, 1):
            raise ConfigError("FLAG_DENS_INTERP must be 0 or 1")
        if self.flag_write_cubes not in (0, 1):
            raise ConfigError("FLAG_WRITE_CUBES must be 0 or 1")
        if self.flag_survey not in (0, 1):
            raise ConfigError("FLAG_SURVEY must be 0 or 1")
        if self.flag_survey == 1 and self.flag_period == 1:
            raise ConfigError("PBCs considered in voids.dat but SURVEY/MASKED conditions activated")
        if self.flag_data in (0, 3):
            if self.flag_div < 0 or self.flag_div > 2:
                raise ConfigError("FLAG_DIV must be 0, 1 or 2")
            if self.flag_dens < 0 or self.flag_dens > 2:
                raise ConfigError("FLAG_DENS must be 0, 1 or 2")


# (label, fields) per value line, in file order.  Labels follow config/voids.dat.
_LAYOUT: List[tuple] = [
    ("files: first, last, every", ("first", "last", "every")),
    ("cells for the coarser grid", ("ncox", "ncoy", "ncoz")),
    ("min and max levels of the hierarchies to be used by voidfind", ("levmin", "levmax")),
    ("Hubble constant (h), omega matter", ("ache", "omega0")),
    ("box size", ("lado0",)),
    ("periodic boundary conditions -> 0: no, 1: yes", ("flag_period",)),
    None,  # block header
    ("maximum number of voids for dimensioning", ("nvoid_max",)),
    ("density contrast threshold for centers", ("dens_thre",)),
    ("density contrast threshold for the edges", ("dens_thre2",)),
    ("density gradient threshold", ("grad_thre",)),
    ("velocity divergence threshold", ("div_thre",)),
    ("min void radius (in Mpc)", ("rmin",)),
    ("min void radius to look for subvoids (in Mpc)", ("rmin_sub",)),
    None,
    ("type of data -> 0:MASCLET, 1:BinPart, 2:BinGrid, 3:GADGET-like", ("flag_data",)),
    ("IF MASCLET or Grid data: NX, NY, NZ (INPUT coarse grid size)", ("nhyx", "nhyy", "nhyz")),
    ("IF GADGET-like: files per snapshot, PartType (1:gas, 2:dm), DM mass (Msun)",
     ("files_per_snap", "parttypex", "massdm")),
    ("GADGET-like reader: 0:IllustrisTNG, 1:FLAMINGO-DMO", ("flag_gadget_reader",)),
    ("IF SURVEY -> 0:no, 1:yes", ("flag_survey",)),
    None,
    ("maximum number of particles to read (DIMENSIONING)", ("partired",)),
    ("particle velocity is available -> 0: no, 1: yes", ("flag_vel_available",)),
    ("interpolation scheme for particle velocity field (0:TSC, 1:SPH)", ("flag_vel_interp",)),
    ("interpolation scheme for particle density field (0:TSC, 1:SPH)", ("flag_dens_interp",)),
    ("Number of nearest neighbours to use to transform into grid", ("kneighbours",)),
    None,
    ("whether to write cubes that are merged to form voids -> 0: no, 1: yes", ("flag_write_cubes",)),
    None,
    ("NPALEV, NLEVELS, NAMRX, NAMRY, NAMRZ", ("npalev", "nlevels", "namrx", "namry", "namrz")),
    ("which velocity for divergence -> 0: all, 1: DM, 2: gas", ("flag_div",)),
    ("which density field to use -> 0: all, 1: DM, 2: gas", ("flag_dens",)),
]

_BLOCK_TITLES = [
    "General parameters block",
    "Void finder input parameters",
    "Type of data to process",
    "Particle data handling parameters",
    "Output management parameters",
    "MASCLET ONLY (skip if another type of input is used)",
]

_STARS = "*" * 79


def _parse_value(tok: str):
    t = tok.strip().replace("d", "e").replace("D", "E")
    try:
        return int(t)
    except ValueError:
        return float(t)


def read_voids_dat(path) -> VoidsConfig:
    """Positional read mirroring avism.f90:179-253 (one READ per line)."""
    lines = Path(path).read_text().splitlines()
    cfg = VoidsConfig()
    pos = 0

    def take() -> str:
        nonlocal pos
        if pos >= len(lines):
            raise ConfigError(f"voids.dat ended early at line {pos + 1}")
        s = lines[pos]
        pos += 1
        return s

    # opening banner: 5 lines, the 5th being the first label (avism.f90:179-183)
    for _ in range(4):
        take()
    for item in _LAYOUT:
        if item is None:
            for _ in range(3):  # block header lines
                take()
            continue
        _label, fields = item
        take()  # label line
        toks = [t for t in re.split(r"[,\s]+", take().strip()) if t]
        if len(toks) < len(fields):
            raise ConfigError(f"voids.dat line {pos}: expected {len(fields)} values")
        for name, tok in zip(fields, toks):
            ftype = type(getattr(cfg, name))
            val = _parse_value(tok)
            setattr(cfg, name, ftype(val))
    return cfg


def write_voids_dat(cfg: VoidsConfig, path) -> None:
    out = [
        "*************PARAMETERS FOR AVISM.F90**********",
        _STARS,
        "*       " + _BLOCK_TITLES[0],
        _STARS,
    ]
    block = 1
    for item in _LAYOUT:
        if item is None:
            out += [_STARS, "*       " + _BLOCK_TITLES[block], _STARS]
            block += 1
            continue
        label, fields = item
        out.append(label + " " + "-" * max(3, 78 - len(label)) + ">")
        vals = []
        for name in fields:
            v = getattr(cfg, name)
            vals.append(repr(float(v)) if isinstance(v, float) else str(v))
        out.append(",".join(vals))
    Path(path).write_text("\n".join(out) + "\n")
