"""MASCLET snapshot files (FLAG_DATA = 0), the input READ_MASCLET reads
(reader.f90:86-405, file names from NOMFILE_MASCLET, reader.f90:2-27):

* ``gridsNNNNN`` (text, list-directed): ``IRR T NL MAP``, ``ZETA``,
  ``IR NDXYZ0 NST0``, then per level IR = 1..NL ``IRR NPATCH NPART NPARTST``,
  one line, and per patch ``PATCHNX PATCHNY PATCHNZ``, ``PATCHX PATCHY PATCHZ``,
  ``PATCHRX PATCHRY PATCHRZ``, ``PARE``;
* ``clusNNNNN`` (Fortran unformatted sequential): a header record, the base
  grid's gas density contrast and velocities (NHYX^3 each), 6 records not
  read, then per patch of levels 1..NL2 its density contrast and velocities
  followed by 7 records not read;
* ``cldmNNNNN``: a header record, the base grid's dark-matter density
  contrast, the base particles' x, y, z, vx, vy, vz (NDXYZ0 each, mass MAP
  from the grids file), one record not read; per level IR = 1..NL, each
  patch's dark-matter density contrast (read up to NL2), then that level's
  particles' x, y, z, vx, vy, vz, mass and one record not read.

``read`` returns what avism_amr_set and avism_load_particles take (the gas of
the base grid and level-1 patches; the particles of every level in MASCLET
units); ``write`` writes such a snapshot (for tests and synthetic runs).
Records carry 4-byte length markers, as gfortran and flang write them.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

__all__ = ["read", "write", "names"]


def names(it: int):
    nom = f"{int(it):05d}"
    return f"clus{nom}", f"cldm{nom}", f"grids{nom}", f"clst{nom}"


def _records(path: Path):
    data = Path(path).read_bytes()
    off = 0
    out = []
    while off < len(data):
        n = int(np.frombuffer(data, np.int32, 1, off)[0])
        out.append(data[off + 4: off + 4 + n])
        n2 = int(np.frombuffer(data, np.int32, 1, off + 4 + n)[0])
        if n2 != n:
            raise ValueError(f"{path}: record markers differ ({n} vs {n2})")
        off += 8 + n
    return out


def _put(fh, *arrays):
    payload = b"".join(np.ascontiguousarray(a).tobytes() for a in arrays)
    m = np.int32(len(payload)).tobytes()
    fh.write(m + payload + m)


def write(dirpath, it: int, nhyx: int, base, patches, dm_base, dm_levels=(), t: float = 0.0, zeta: float = 0.0,
          map_mass: float = 1.0):
    """base: [delta, vx, vy, vz] of nhyx^3 (x fastest); patches: level-1 dicts
    (n, x, r, fields) as synth.amr_gas makes them; dm_base: (x, y, z, vx, vy,
    vz) of the base particles (mass map_mass); dm_levels: per level >= 1 a
    tuple (x, y, z, vx, vy, vz, m) of its particles (level 1 is that of the
    patches)."""
    d = Path(dirpath)
    d.mkdir(parents=True, exist_ok=True)
    f_clus, f_cldm, f_grids, _ = names(it)
    nl = max(1 if patches else 0, len(dm_levels))
    npart = [len(lv[0]) for lv in dm_levels] + [0] * (nl - len(dm_levels))
    lines = [f"{it} {t!r} {nl} {map_mass!r}", f"{zeta!r}", f"0 {len(dm_base[0])} 0"]
    for ir in range(1, nl + 1):
        plist = patches if ir == 1 else []
        lines.append(f"{ir} {len(plist)} {npart[ir - 1]} 0")
        lines.append("0")
        for p in plist:
            lines.append(" ".join(str(v) for v in p["n"]))
            lines.append(" ".join(str(v) for v in p["x"]))
            lines.append(" ".join(repr(float(np.float32(v))) for v in p["r"]))
            lines.append("0")
    (d / f_grids).write_text("\n".join(lines) + "\n")
    z4 = np.zeros(1, np.float32)
    with open(d / f_clus, "wb") as fh:
        _put(fh, np.int32(it))
        for a in base:
            _put(fh, np.asarray(a, np.float32))
        for _ in range(6):
            _put(fh, z4)
        for p in patches:
            for a in p["fields"]:
                _put(fh, np.asarray(a, np.float32))
            for _ in range(7):
                _put(fh, z4)
    with open(d / f_cldm, "wb") as fh:
        _put(fh, np.int32(it))
        _put(fh, np.zeros(nhyx ** 3, np.float32))
        for a in dm_base:
            _put(fh, np.asarray(a, np.float32))
        _put(fh, z4)
        for ir in range(1, nl + 1):
            plist = patches if ir == 1 else []
            for p in plist:
                _put(fh, np.zeros(int(np.prod(p["n"])), np.float32))
            lv = dm_levels[ir - 1] if ir - 1 < len(dm_levels) else (np.zeros(0, np.float32),) * 7
            for a in lv:
                _put(fh, np.asarray(a, np.float32))
            _put(fh, z4)


def read(dirpath, it: int, nhyx: int, nl2: int = 1):
    """The snapshot's base gas, its level-1 patches (synth.amr_gas's form) and
    the dark-matter particles of every level: dict(base, patches, particles =
    (x, y, z, vx, vy, vz, m), t, zeta, nl)."""
    d = Path(dirpath)
    f_clus, f_cldm, f_grids, _ = names(it)
    tok = (d / f_grids).read_text().split("\n")
    head = tok[0].split()
    nl, map_mass = int(head[2]), float(head[3])
    t = float(head[1])
    zeta = float(tok[1].split()[0])
    ndxyz0 = int(tok[2].split()[1])
    row = 3
    levels = []
    for ir in range(1, nl + 1):
        w = tok[row].split()
        npatch, npart = int(w[1]), int(w[2])
        row += 2
        plist = []
        for _ in range(npatch):
            n = tuple(int(v) for v in tok[row].split())
            x = tuple(int(v) for v in tok[row + 1].split())
            r = tuple(float(v) for v in tok[row + 2].split())
            row += 4
            plist.append(dict(n=n, x=x, r=r))
        levels.append((plist, npart))
    n3 = nhyx ** 3
    rec = _records(d / f_clus)
    base = [np.frombuffer(rec[1 + q], np.float32, n3).copy() for q in range(4)]
    k = 1 + 4 + 6
    for ir, (plist, _) in enumerate(levels, start=1):
        if ir > nl2:
            break
        for p in plist:
            m = int(np.prod(p["n"]))
            p["fields"] = [np.frombuffer(rec[k + q], np.float32, m).copy() for q in range(4)]
            k += 4 + 7
    rec = _records(d / f_cldm)
    parts = [np.frombuffer(rec[2 + q], np.float32, ndxyz0).copy() for q in range(6)]
    parts.append(np.full(ndxyz0, map_mass, np.float32))
    k = 2 + 6 + 1
    for plist, npart in levels:
        k += len(plist)
        for q in range(7):
            parts[q] = np.concatenate([parts[q], np.frombuffer(rec[k + q], np.float32, npart)])
        k += 7 + 1
    patches = levels[0][0] if levels else []
    return dict(base=base, patches=patches, particles=tuple(parts), t=t, zeta=zeta, nl=nl)
