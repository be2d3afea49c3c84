"""LAMMPS-side DP packing, restated on the CPU (TEST INFRASTRUCTURE: the checker
for matpl_amd.neighbor.generate_neighdata; only tests may import it).

``generate_neighdata`` follows PairMATPL::generate_neighdata
(src/lmps/lammps-libtorch/MATPL/pair_matpl.cpp:522-616) entry by entry: row i =
ilist[ii], neighbors j of the full list in list order, accepted when
dx*dx + dy*dy + dz*dz <= rc^2 with d = x_j - x_i, placed at the next free slot of
bucket type_to_model[type[j]-1] as j + 1 with [sqrt(rsq), dx, dy, dz]; empty
slots 0 (the reference leaves them as the reused buffer held them).

``ghost_system`` / ``full_list`` build the LAMMPS input for a periodic frame the
way LAMMPS does: ghost atoms are the periodic images lying within ``cutoff`` of
the box (orthogonal or triclinic, via fractional distance to the faces), each
remembering its owner; the full list of every local atom holds all other atoms
(local or ghost) within ``cutoff`` (cutoff = rc + skin), in index order.
Pinned by tests/test_gpu_lmp.py: forces folded back onto owners equal the
periodic oracle's forces (an independent path: images by the minimum-image /
image-enumeration rule of NeighConst.f90).
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def generate_neighdata(x, atom_type, type_to_model, ilist, numneigh, firstneigh, neigh, ntypes, m, rcut):
    inum = len(ilist)
    S = ntypes * m
    lst = np.zeros((inum, S), np.int32)
    dR = np.zeros((inum, S, 4), np.float64)
    counts = np.zeros((inum, ntypes), np.int32)
    imagetype = np.zeros(inum, np.int32)
    rc2 = rcut * rcut
    min_rsq = math.inf
    for ii in range(inum):
        i = int(ilist[ii])
        imagetype[ii] = type_to_model[atom_type[i] - 1]
        o = int(firstneigh[i])
        for jj in range(int(numneigh[i])):
            j = int(neigh[o + jj])
            dx = float(x[j, 0]) - float(x[i, 0])
            dy = float(x[j, 1]) - float(x[i, 1])
            dz = float(x[j, 2]) - float(x[i, 2])
            rsq = dx * dx + dy * dy + dz * dz
            if rsq <= rc2:
                t = type_to_model[atom_type[j] - 1]
                k = counts[i, t]
                if k < m:
                    slot = t * m + k
                    lst[i, slot] = j + 1
                    dR[i, slot] = (math.sqrt(rsq), dx, dy, dz)
                counts[i, t] += 1
                min_rsq = min(min_rsq, rsq)
    return imagetype, lst, dR, counts, min_rsq


def ghost_system(lattice, pos, cutoff):
    """(x [nall, 3], owner [nall]): local atoms wrapped into the box (as LAMMPS keeps them), then
    every periodic image within ``cutoff`` of the box, in (image, atom) order."""
    L = np.asarray(lattice, np.float64)
    pos = np.asarray(pos, np.float64)
    inv = np.linalg.inv(L)
    frac = pos @ inv
    frac = frac - np.floor(frac)
    pos = frac @ L
    # fractional thickness of the cutoff per axis: cutoff / plane spacing
    V = abs(np.linalg.det(L))
    h = np.array([V / np.linalg.norm(np.cross(L[(d + 1) % 3], L[(d + 2) % 3])) for d in range(3)])
    w = cutoff / h
    n = np.ceil(w).astype(int)
    xs, owner = [pos], [np.arange(len(pos))]
    for s in itertools.product(*[range(-k, k + 1) for k in n]):
        if s == (0, 0, 0):
            continue
        f = frac + np.array(s, np.float64)
        keep = np.all((f > -w) & (f < 1.0 + w), axis=1)
        if keep.any():
            xs.append(f[keep] @ L)
            owner.append(np.nonzero(keep)[0])
    return np.concatenate(xs), np.concatenate(owner)


def full_list(x, nlocal, cutoff):
    """LAMMPS-style full list of the local atoms: (ilist, numneigh, firstneigh, neigh)."""
    x = np.asarray(x, np.float64)
    rc2 = cutoff * cutoff
    neigh, numneigh, first = [], [], []
    for i in range(nlocal):
        d = x - x[i]
        r2 = np.einsum("ij,ij->i", d, d)
        js = np.nonzero(r2 < rc2)[0]
        js = js[js != i]
        first.append(len(neigh))
        numneigh.append(len(js))
        neigh.extend(js.tolist())
    return (np.arange(nlocal, dtype=np.int32), np.asarray(numneigh, np.int32), np.asarray(first, np.int64),
            np.asarray(neigh, np.int32))
