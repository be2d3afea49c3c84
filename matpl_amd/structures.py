"""Seeded synthetic atomic structures for the DP and NEP workloads.

Every generator returns plain numpy arrays; atoms are stored grouped by type
(all atoms of type 0 first, then type 1, ...), which is the order the
reference DP model assumes when it groups per-type energies
(src/model/dp_dp.py:327-341) and computes per-type statistics
(src/pre_data/dpuni_data_loader.py:464-470).

Random streams are numpy PCG64 seeded per workload.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Standard atomic masses (g/mol), used for densities and for MD.
MASS = {"H": 1.008, "Li": 6.94, "C": 12.011, "O": 15.999, "Si": 28.085, "P": 30.974,
        "S": 32.06, "Cu": 63.546, "Ge": 72.630, "Ag": 107.868, "Hf": 178.49,
        "Au": 196.967}
ATOMIC_NUMBER = {"H": 1, "Li": 3, "C": 6, "O": 8, "Si": 14, "P": 15, "S": 16, "Cu": 29,
                 "Ge": 32, "Ag": 47, "Hf": 72, "Au": 79}
AVOGADRO = 6.02214076e23
# 1 amu * (Angstrom/fs)^2 expressed in eV: 1.66053907e-27 kg * 1e10 m^2/s^2 / 1.602176634e-19 J
AMU_A2_FS2_IN_EV = 103.642696


@dataclass
class Frame:
    """One periodic frame: lattice rows are the cell vectors (Angstrom)."""
    lattice: np.ndarray          # [3, 3] float64
    pos: np.ndarray              # [N, 3] float64 Cartesian
    types: np.ndarray            # [N] int32, 0-based index into `elements`
    elements: list = field(default_factory=list)

    @property
    def natoms(self) -> int:
        return int(self.pos.shape[0])

    def frac(self) -> np.ndarray:
        return self.pos @ np.linalg.inv(self.lattice)


def _sort_by_type(lattice, pos, types, elements) -> Frame:
    order = np.argsort(types, kind="stable")
    return Frame(np.ascontiguousarray(lattice, dtype=np.float64),
                 np.ascontiguousarray(pos[order], dtype=np.float64),
                 np.ascontiguousarray(types[order], dtype=np.int32), list(elements))


def _crystal(basis_frac, basis_types, a, reps, sigma, rng, elements) -> Frame:
    nx, ny, nz = reps
    cells = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"),
                     -1).reshape(-1, 3).astype(np.float64)
    frac = (cells[:, None, :] + np.asarray(basis_frac)[None, :, :]).reshape(-1, 3)
    types = np.tile(np.asarray(basis_types, dtype=np.int32), cells.shape[0])
    lattice = np.diag([a * nx, a * ny, a * nz]).astype(np.float64)
    pos = frac * a
    pos = pos + rng.normal(0.0, sigma, size=pos.shape)
    return _sort_by_type(lattice, pos, types, elements)


FCC = [[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]


def fcc_cu(reps=(3, 3, 3), a=3.615, sigma=0.1, seed=1000) -> Frame:
    """fcc Cu, 4 atoms per conventional cell (108 atoms for 3x3x3)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return _crystal(FCC, [0] * 4, a, reps, sigma, rng, ["Cu"])


def zincblende_sic(reps=(6, 6, 6), a=4.36, sigma=0.08, seed=1001) -> Frame:
    """Zincblende SiC: fcc Si + fcc C shifted by (1/4,1/4,1/4); 8 atoms per cell."""
    rng = np.random.Generator(np.random.PCG64(seed))
    basis = FCC + [[x + .25, y + .25, z + .25] for x, y, z in FCC]
    return _crystal(basis, [0] * 4 + [1] * 4, a, reps, sigma, rng, ["Si", "C"])


def fluorite_hfo2(reps=(19, 19, 23), a=5.08, sigma=0.1, seed=1003) -> Frame:
    """Fluorite HfO2: fcc Hf + O on the 8 tetrahedral sites; 12 atoms per cell.

    19x19x23 cells give 99,636 atoms (33,212 Hf + 66,424 O).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    oxy = [[x, y, z] for x in (.25, .75) for y in (.25, .75) for z in (.25, .75)]
    return _crystal(FCC + oxy, [0] * 4 + [1] * 8, a, reps, sigma, rng, ["Hf", "O"])


def _random_frame(n, counts, elements, density, dmin, rng) -> Frame:
    """Uniform random positions in a cubic box, rejecting pairs closer than dmin."""
    L = (n / density) ** (1.0 / 3.0)
    pos = np.empty((n, 3))
    cell = max(dmin, 1e-6)
    ncell = max(1, int(L // cell))
    grid: dict = {}
    placed = 0
    tries = 0
    while placed < n:
        tries += 1
        if tries > 2000 * n:
            raise RuntimeError("random packing failed; density too high for dmin")
        p = rng.uniform(0.0, L, size=3)
        c = tuple((p / L * ncell).astype(int) % ncell)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = ((c[0] + dx) % ncell, (c[1] + dy) % ncell, (c[2] + dz) % ncell)
                    for q in grid.get(key, ()):
                        d = pos[q] - p
                        d -= L * np.round(d / L)
                        if d @ d < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pos[placed] = p
            grid.setdefault(c, []).append(placed)
            placed += 1
    types = np.repeat(np.arange(len(counts), dtype=np.int32), counts)
    return _sort_by_type(np.diag([L, L, L]), pos, types, elements)


def lgps_frames(nframes=256, nmin=64, nmax=512, rho_g_cm3=2.0, dmin=1.8, seed=1002):
    """Li10GeP2S12-stoichiometry frames with uniform random atom counts.

    Per frame: N ~ U[nmin, nmax]; Ge = max(1, round(N/25)), P = round(2N/25),
    Li = round(10N/25), S = N - Li - Ge - P. The cubic box is sized from the
    frame's own mass at rho_g_cm3.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    elements = ["Li", "Ge", "P", "S"]
    frames = []
    for _ in range(nframes):
        n = int(rng.integers(nmin, nmax + 1))
        ge = max(1, int(round(n / 25)))
        p = int(round(2 * n / 25))
        li = int(round(10 * n / 25))
        s = n - li - ge - p
        counts = [li, ge, p, s]
        mass = sum(c * MASS[e] for c, e in zip(counts, elements))     # g/mol
        vol_A3 = mass / AVOGADRO / rho_g_cm3 * 1e24
        frames.append(_random_frame(n, counts, elements, n / vol_A3, dmin, rng))
    return frames


def agausic_frames(nframes=128, nmin=32, nmax=1024, density=0.06, dmin=1.9, seed=1004):
    """4-element Ag-Au-Si-C random alloy frames, each atom's element uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    elements = ["Ag", "Au", "Si", "C"]
    frames = []
    for _ in range(nframes):
        n = int(rng.integers(nmin, nmax + 1))
        t = rng.integers(0, 4, size=n)
        counts = [int((t == k).sum()) for k in range(4)]
        frames.append(_random_frame(n, counts, elements, density, dmin, rng))
    return frames


def maxwell_velocities(frame: Frame, temperature_K=300.0, seed=2003) -> np.ndarray:
    """Seeded Maxwell-Boltzmann velocities (Angstrom/fs) with zero net momentum."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kB = 8.617333262e-5                      # eV/K
    m = np.array([MASS[frame.elements[t]] for t in frame.types]) * AMU_A2_FS2_IN_EV
    sig = np.sqrt(kB * temperature_K / m)
    v = rng.normal(size=(frame.natoms, 3)) * sig[:, None]
    p = (v * m[:, None]).sum(0)
    v -= p / m.sum()
    return v
