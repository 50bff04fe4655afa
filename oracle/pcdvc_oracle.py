"""CPU oracle for swarm population control -- TEST INFRASTRUCTURE ONLY.

Restatement of ``PopulationControl(swarm, ...).repopulate()`` (PY/swarm/_population_control.py
:13-140) with the non-aggressive path of _PCDVC_Calculate3D / 2D
(LU/PICellerator/Utils/src/PCDVC.c:313-700, 703-1085) and the swarm's edit primitives
(LU/StgDomain/Swarm/src/SwarmClass.c: Swarm_CreateNewParticle :1656-1665, Swarm_AddParticleToCell
:1121-1136, Swarm_RemoveParticleFromCell :1016-1030, Swarm_DeleteParticle :1032-1082).  The
DVC weights and centroids come from the C oracle (``oracle.dvc_all``).

Per cell, in cell order: the cell's weights decide -- particles with w > split * 2^dim are
split (first maxSplits in table order; the copy is appended to the swarm and to the cell, placed
at the centroid of the parent's Voronoi cell mapped to global coordinates), particles with
w < delete * 2^dim are deleted by descending swarm index (at most maxDeletions): removed from the
cell table (the cell's last entry takes the hole) and the swarm's last particle is moved into
the freed index.

Pins (tests/test_pcdvc_oracle.py): the reference's doctest (PerCellGaussLayout with 4 points
per direction on the default 4x4 mesh, thresholds 0 and maxSplits 9999: 256 -> 512 particles),
a one-particle cell splitting at the element centre, and hand-worked deletion sequences.
"""
import numpy as np

from . import oracle as O


def _shape(d, xi):
    """Q1 shape functions as written in BilinearElementType.c / TrilinearElementType.c."""
    if d == 2:
        x, y = float(xi[0]), float(xi[1])
        return [0.25 * (1.0 - x) * (1.0 - y), 0.25 * (1.0 + x) * (1.0 - y),
                0.25 * (1.0 - x) * (1.0 + y), 0.25 * (1.0 + x) * (1.0 + y)]
    x, y, z = float(xi[0]), float(xi[1]), float(xi[2])
    return [0.125 * (1.0 - x) * (1.0 - y) * (1.0 - z), 0.125 * (1.0 + x) * (1.0 - y) * (1.0 - z),
            0.125 * (1.0 - x) * (1.0 + y) * (1.0 - z), 0.125 * (1.0 + x) * (1.0 + y) * (1.0 - z),
            0.125 * (1.0 - x) * (1.0 - y) * (1.0 + z), 0.125 * (1.0 + x) * (1.0 - y) * (1.0 + z),
            0.125 * (1.0 - x) * (1.0 + y) * (1.0 + z), 0.125 * (1.0 + x) * (1.0 + y) * (1.0 + z)]


def local_to_global(mesh, e, xi, verts=None):
    """FeMesh_CoordLocalToGlobal (FeMesh.c:333-365): x = sum_a N_a(xi) x_a in vertex order."""
    v = mesh.element_verts(e) if verts is None else verts[e]
    d = mesh.dim
    N = _shape(d, xi)
    x = [0.0] * d
    for a in range(2 ** d):
        for k in range(d):
            x[k] += N[a] * float(v[a, k])
    return np.array(x)


def population_control(mesh, coords, ptr, pid, cen, w, delete_threshold, split_threshold, max_deletions,
                       max_splits):
    """Returns (coords_new, src, ptr_new, pid_new); src[i] is the original index whose data the
    new particle i carries.  cen / w are the DVC centroids and weights in slot (table) order."""
    d = mesh.dim
    vol = float(2 ** d)
    maxW, minW = split_threshold * vol, delete_threshold * vol
    E = mesh.E
    tables = [list(pid[ptr[e]:ptr[e + 1]]) for e in range(E)]
    P = len(pid)
    src = list(range(P))
    newc = [None] * P
    cell_of = [0] * P
    for e in range(E):
        for i in tables[e]:
            cell_of[i] = e
    verts = mesh.coords()[mesh.element_nodes()]
    for e in range(E):
        s0, n0 = ptr[e], ptr[e + 1] - ptr[e]
        T = tables[e]
        splits = [i for i in range(n0) if w[s0 + i] > maxW][:max_splits]
        dels = [(T[i], i) for i in range(n0) if w[s0 + i] < minW]
        for i in splits:
            parent = T[i]
            src.append(src[parent])
            newc.append(local_to_global(mesh, e, cen[s0 + i], verts))
            cell_of.append(e)
            T.append(P)
            P += 1
        dels.sort(key=lambda t: -t[0])
        for idx, _ in dels[:max_deletions]:
            Tc = tables[cell_of[idx]]
            q = Tc.index(idx)
            Tc[q] = Tc[-1]
            Tc.pop()
            last = P - 1
            if idx != last:
                Tl = tables[cell_of[last]]
                Tl[Tl.index(last)] = idx
                src[idx], newc[idx], cell_of[idx] = src[last], newc[last], cell_of[last]
            src.pop()
            newc.pop()
            cell_of.pop()
            P -= 1
    src = np.array(src, dtype=np.int64)
    out = np.array(coords, dtype=np.float64)[src].copy()
    for i, c in enumerate(newc):
        if c is not None:
            out[i] = c
    nptr = np.zeros(E + 1, dtype=np.int64)
    nptr[1:] = np.cumsum([len(t) for t in tables])
    npid = np.array([i for t in tables for i in t], dtype=np.int64)
    return out, src, nptr, npid


def repopulate(mesh, coords, ptr, pid, delete_threshold=0.006, split_threshold=0.25, max_deletions=0,
               max_splits=3):
    """Coincident map + DVC of the current swarm, then the population edits."""
    xi = O.coincident_map(mesh, ptr, pid, coords)
    cen, w = O.dvc_all(mesh.dim, ptr, xi)
    return population_control(mesh, coords, ptr, pid, cen, w, delete_threshold, split_threshold, max_deletions,
                              max_splits)
