"""Seeded synthetic inputs for the benchmark and test configurations.

There is no network here, so every workload is generated:

* ``zeldovich``   -- lattice particles displaced by a Gaussian random field with
  P(k) ~ k^slope exp(-k^2/kc^2) (Zel'dovich approximation), rms displacement
  fixed, velocities f*H0*psi, periodic box, positions wrapped to [-L/2, L/2).
* ``mock_voids``  -- uniform particles in a non-periodic box thinned inside
  spherical voids with delta(r) = -0.85 (1 - (r/R)^3) and a linear radial outflow
  v = -(1/3) f H0 delta(r) r plus Gaussian velocity noise (the reference's test1
  idea, test1/mock_voids.f90, with our own generator and profile).
* ``grid_fields`` -- a log-normal density contrast and its linear-theory velocity
  field on a grid, for grid-input runs (reader.f90:523-552).

Arrays are float32 (positions Mpc/h, velocities km/s, masses Msun), as the
reference's binary particle format stores them (reader.f90:447-453).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

__all__ = ["Particles", "zeldovich", "mock_voids", "grid_fields", "particle_mass", "survey_mask", "patch_batch",
           "amr_gas"]

RHO_CRIT = 2.775e11  # h^2 Msun / Mpc^3


@dataclass
class Particles:
    x: torch.Tensor
    y: torch.Tensor
    z: torch.Tensor
    vx: torch.Tensor
    vy: torch.Tensor
    vz: torch.Tensor
    m: torch.Tensor
    L0: float
    periodic: int

    @property
    def n(self) -> int:
        return int(self.x.numel())

    def to(self, device) -> "Particles":
        f = lambda t: t.to(device)
        return Particles(f(self.x), f(self.y), f(self.z), f(self.vx), f(self.vy), f(self.vz),
                         f(self.m), self.L0, self.periodic)

    def numpy(self):
        return [t.detach().cpu().numpy() for t in (self.x, self.y, self.z, self.vx, self.vy, self.vz, self.m)]


def particle_mass(L0: float, np_total: int, omega_m: float = 0.31) -> float:
    return omega_m * RHO_CRIT * L0 ** 3 / np_total


def _kgrid(n: int, L: float, device):
    k1 = torch.fft.fftfreq(n, d=L / n, device=device) * (2 * math.pi)
    kr = torch.fft.rfftfreq(n, d=L / n, device=device) * (2 * math.pi)
    kx = k1.view(n, 1, 1)
    ky = k1.view(1, n, 1)
    kz = kr.view(1, 1, -1)
    return kx, ky, kz


def _gaussian_k(n: int, L: float, seed: int, slope: float, kc: float, device):
    """Fourier amplitudes of a real Gaussian field with P(k) ~ k^slope exp(-k^2/kc^2).

    Index order of the returned arrays is (x, y, z); callers flatten them in
    Fortran order (x fastest)."""
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    white = torch.randn((n, n, n), generator=g, device=device, dtype=torch.float32)
    wk = torch.fft.rfftn(white)
    del white
    kx, ky, kz = _kgrid(n, L, device)
    k2 = kx * kx + ky * ky + kz * kz
    k2c = k2.clone()
    k2c[0, 0, 0] = 1.0
    amp = torch.pow(k2c, slope / 4.0) * torch.exp(-k2c / (2.0 * kc * kc))
    amp[0, 0, 0] = 0.0
    return wk * amp, (kx, ky, kz, k2c)


def _fortran_flat(a: torch.Tensor) -> torch.Tensor:
    """(x, y, z)-indexed array -> flat, x fastest."""
    return a.permute(2, 1, 0).contiguous().view(-1)


def zeldovich(n_side: int, L0: float, seed: int, kc: float = 0.5, slope: float = -1.5,
              rms_disp: float = 3.0, f: float = 0.5, H0: float = 100.0,
              device="cpu", omega_m: float = 0.31) -> Particles:
    n = n_side
    dk, (kx, ky, kz, k2) = _gaussian_k(n, L0, seed, slope, kc, device)
    psi = []
    for kj in (kx, ky, kz):
        comp = torch.fft.irfftn(dk * (1j * kj / k2), s=(n, n, n))
        psi.append(comp)
    del dk
    rms = torch.sqrt(sum((p.double() ** 2).mean() for p in psi)).item()
    scale = rms_disp / max(rms, 1e-30)
    q = (torch.arange(n, device=device, dtype=torch.float64) + 0.5) * (L0 / n) - L0 / 2.0
    qs = (q.view(n, 1, 1), q.view(1, n, 1), q.view(1, 1, n))
    out = []
    vel = []
    for j in range(3):
        d = psi[j].double() * scale
        pos = qs[j] + d
        pos = torch.remainder(pos + L0 / 2.0, L0) - L0 / 2.0
        out.append(_fortran_flat(pos.float()))
        vel.append(_fortran_flat((d * (f * H0)).float()))
    m = torch.full_like(out[0], particle_mass(L0, n ** 3, omega_m))
    return Particles(out[0], out[1], out[2], vel[0], vel[1], vel[2], m, float(L0), 1)


def mock_voids(n0: int = 1_000_000, L0: float = 200.0, nvoids: int = 40, rmin: float = 8.0,
               rmax: float = 25.0, seed: int = 1, f: float = 0.5, H0: float = 100.0,
               sigma_v: float = 50.0, omega_m: float = 0.31):
    """Returns (Particles, centres[nvoids,3], radii[nvoids]) -- non-periodic box."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-L0 / 2, L0 / 2, size=(nvoids, 3))
    radii = rng.uniform(rmin, rmax, size=nvoids)
    pos = rng.uniform(-L0 / 2, L0 / 2, size=(n0, 3))
    delta = np.zeros(n0)
    vel = np.zeros((n0, 3))
    for c, R in zip(centres, radii):
        d = pos - c
        r = np.sqrt((d * d).sum(1))
        inside = r < R
        dv = -0.85 * (1.0 - (r[inside] / R) ** 3)
        # overlapping voids: keep the deepest contrast
        upd = dv < delta[inside]
        idx = np.nonzero(inside)[0][upd]
        delta[idx] = dv[upd]
        vel[idx] = (-(1.0 / 3.0) * f * H0 * dv[upd])[:, None] * d[idx]
    keep = rng.uniform(size=n0) < (1.0 + delta)
    pos = pos[keep]
    vel = vel[keep] + rng.normal(0.0, sigma_v, size=vel[keep].shape)
    npk = pos.shape[0]
    m = np.full(npk, particle_mass(L0, npk, omega_m), dtype=np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
    parts = Particles(t(pos[:, 0]), t(pos[:, 1]), t(pos[:, 2]), t(vel[:, 0]), t(vel[:, 1]),
                      t(vel[:, 2]), t(m), float(L0), 0)
    return parts, centres, radii


def survey_mask(n: int, empty_frac: float = 0.2, seed: int = 0, rmin: float = 1.5, rmax: float = 5.0) -> np.ndarray:
    """A seeded survey-like mask on an n^3 grid (int8, x fastest, 1 = observed):
    random balls of radius rmin..rmax cells are cut out until empty_frac of the
    cells are empty (the shape the reference reads as SMASK, avism.f90:403-418)."""
    rng = np.random.default_rng(seed)
    m = np.ones((n, n, n), np.int8)  # [z, y, x]
    empty, target = 0, empty_frac * n ** 3
    while empty < target:
        c = rng.uniform(0, n, size=3)
        r = rng.uniform(rmin, rmax)
        # only the ball's bounding box can change
        lo = [max(int(np.floor(c[a] - r)), 0) for a in range(3)]
        hi = [min(int(np.ceil(c[a] + r)) + 1, n) for a in range(3)]
        zz, yy, xx = np.meshgrid(np.arange(lo[2], hi[2]), np.arange(lo[1], hi[1]), np.arange(lo[0], hi[0]),
                                 indexing="ij")
        sub = m[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        cut = ((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 < r * r) & (sub == 1)
        empty += int(cut.sum())
        sub[cut] = 0
    return m.ravel()


def patch_batch(npatch: int, seed0: int = 3000, np_range=(2e5, 8e6), grids=(32, 64, 96, 128, 160),
                cell: float = 1.953125, empty_frac: float = 0.2, device="cpu"):
    """Independent survey-like patches: patch i (seed seed0 + i) is a Zel'dovich
    field in a non-periodic box of N^3 cells of side `cell` Mpc/h, N drawn from
    `grids`, with about N_p particles, N_p log-uniform in np_range (the nearest
    cube of a lattice side), and a survey mask with empty_frac of its cells
    empty.  Returns a list of dicts (n, L0, particles, mask)."""
    out = []
    for i in range(npatch):
        rng = np.random.default_rng(seed0 + i)
        n = int(rng.choice(grids))
        npt = float(np.exp(rng.uniform(np.log(np_range[0]), np.log(np_range[1]))))
        side = max(int(round(npt ** (1.0 / 3.0))), 4)
        L0 = n * cell
        p = zeldovich(side, L0, seed=seed0 + i, device=device)
        p = Particles(p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m, L0, 0)
        out.append(dict(n=n, L0=L0, particles=p, mask=survey_mask(n, empty_frac, seed=seed0 + i)))
    return out


def grid_fields(n: int, L0: float, seed: int, sigma: float = 0.8, kc: float = 0.5,
                slope: float = -1.5, f: float = 0.5, H0: float = 100.0, device="cpu"):
    """Log-normal delta (> -1) and linear-theory velocity (km/s), flat x-fastest."""
    dk, (kx, ky, kz, k2) = _gaussian_k(n, L0, seed, slope, kc, device)
    g = torch.fft.irfftn(dk, s=(n, n, n)).double()
    g = g * (sigma / g.std())
    delta = torch.exp(g - 0.5 * sigma * sigma) - 1.0
    dl = torch.fft.rfftn(delta)
    vel = []
    for kj in (kx, ky, kz):
        vel.append(_fortran_flat(torch.fft.irfftn(dl * (1j * f * H0 * kj / k2), s=(n, n, n)).float()))
    return _fortran_flat(delta.float()), vel[0], vel[1], vel[2]


def amr_gas(nb: int, L0: float, seed: int, npatch: int = 12, nmin: int = 4, nmax: int = 12):
    """A MASCLET-like gas snapshot: a log-normal density contrast and velocities
    (units of c) on the base grid of nb^3 cells, and npatch level-1 patches
    (twice the resolution) of even sides nmin..nmax fine cells, at least two
    base cells from the box faces, some overlapping, with their own fields
    (the base field refined plus small-scale noise).

    Returns (base, patches): base = [delta, vx, vy, vz] float32 arrays of nb^3
    (x fastest); patches = list of dicts with n (nx, ny, nz), x (PATCHX, PATCHY,
    PATCHZ: 1-based base cell of the patch's first cells), r (PATCHRX, RY, RZ:
    the first fine cell's centre plus a quarter cell) and fields [delta, vx,
    vy, vz] float32 arrays of nx*ny*nz (x fastest), as READ_MASCLET reads them
    (reader.f90:86-405)."""
    rng = np.random.default_rng(seed)
    d, vx, vy, vz = (a.numpy() for a in grid_fields(nb, L0, seed))
    vx, vy, vz = (v / 299792.458 for v in (vx, vy, vz))
    base = [np.ascontiguousarray(a, np.float32) for a in (d, vx, vy, vz)]
    # MALLA of the level-1 grid (2 nb cells): the first fine cell's centre
    n1 = 2 * nb
    lado = np.float32(L0) - np.float32(L0) / np.float32(n1)
    a = -lado / np.float32(2.0)
    cstep = (lado / np.float32(2.0) - a) / np.float32(n1 - 1)
    dxl = (a + cstep) - a
    cube = [b.reshape(nb, nb, nb) for b in base]  # [z, y, x]
    patches = []
    for _ in range(npatch):
        nn = [2 * int(rng.integers(nmin // 2, nmax // 2 + 1)) for _ in range(3)]
        x0 = [int(rng.integers(2, nb - nn[i] // 2 - 1)) for i in range(3)]  # 1-based base cell
        fields = []
        for q in range(4):
            sl = cube[q][x0[2] - 1:x0[2] - 1 + nn[2] // 2, x0[1] - 1:x0[1] - 1 + nn[1] // 2,
                         x0[0] - 1:x0[0] - 1 + nn[0] // 2]
            up = sl.repeat(2, 0).repeat(2, 1).repeat(2, 2)
            noise = rng.normal(0.0, 0.05 if q == 0 else 2e-5, up.shape)
            f = (up + noise).astype(np.float32)
            if q == 0:
                f = np.maximum(f, np.float32(-0.99))
            fields.append(np.ascontiguousarray(f).ravel())
        r = [float(a + np.float32(2 * x0[i] - 2) * dxl + np.float32(0.25) * dxl) for i in range(3)]
        patches.append(dict(n=tuple(nn), x=tuple(x0), r=tuple(r), fields=fields))
    return base, patches


def amr_pack(patches):
    """(pint int32[6 np], preal float32[3 np], pdata float32) of a patch list, as
    avism_amr_set and the oracle take them."""
    pint = np.array([list(p["n"]) + list(p["x"]) for p in patches], np.int32).ravel()
    preal = np.array([list(p["r"]) for p in patches], np.float32).ravel()
    pdata = np.concatenate([f for p in patches for f in p["fields"]]).astype(np.float32) if patches \
        else np.zeros(1, np.float32)
    return pint, preal, pdata
