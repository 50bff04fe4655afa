"""Writes the golden fixtures under tests/golden/ from data files and documented values held
in the reference tree (run in a container that has /root/reference; the fixtures are
committed so tests never read the reference at run time).

Sources (relative to the reference tree, LU/ = src/underworld/libUnderworld/):
  dvc_3d.json / dvc_2d.json   LU/PICellerator/Weights/tests/expected/DVCWeightsSuite_testGetCentroids{,2D}.txt
                              (expected centroids and weights) with the particle coordinates the suite
                              uses (DVCWeightsSuite.c:162-195 and :372-393), resolution 10
  dvc_grid_3d.json            LU/PICellerator/Weights/tests/expected/DVCWeightsSuite_testConstructGrid.txt
  isoviscous_K.npz            LU/StgFEM/Assembly/tests/expected/testIsoviscousSiffness.matrix (PETSc binary
                              AIJ, big endian) -- 3x3 Q1 mesh on the unit square, eta = 1, Gauss quadrature
  docstrings.json             PY/swarm/layouts.py:200-214 (random layout) and
                              PY/systems/_timeintegration.py:131-147 (RK2) documented outputs
  solcx.npz                   Velic solCx closed form (LU/Underworld/SysTest/AnalyticPlugins/Velic_solCx),
                              evaluated with oracle/_ref/libsolcx.so built from those sources
  tempfield_inp_64_Ra1e4.h5   docs/examples/input/03_BlankenbachBenchmark/tempfield_inp_64_Ra1e4.h5, copied
                              as is: the steady-state temperature field (64^2 Q1, float32) that
                              docs/examples/03_BlankenbachBenchmark.ipynb loads with MeshVariable.load
"""
import ctypes
import json
import os
import re
import shutil
import sys

import numpy as np

REF = os.environ.get("UW_REFERENCE", "/root/reference")
LU = os.path.join(REF, "src/underworld/libUnderworld")
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_centroids(path, dim):
    txt = open(path).read()
    out = []
    if dim == 3:
        pat = re.compile(r"coords: \(x, y, z\) = \(([-\d.]+), ([-\d.]+), ([-\d.]+)\)\s*centroids: \(cx, cy, cz\) = "
                         r"\(([-\d.]+), ([-\d.]+), ([-\d.]+)\)\s*weight = ([-\d.]+)")
    else:
        pat = re.compile(r"coords: \(x, y\) = \(([-\d.]+), ([-\d.]+)\)\s*centroids: \(cx, cy\) = "
                         r"\(([-\d.]+), ([-\d.]+)\)\s*weight = ([-\d.]+)")
    for m in pat.finditer(txt):
        v = [float(x) for x in m.groups()]
        out.append({"coords": v[:dim], "centroid": v[dim:2 * dim], "weight": v[2 * dim]})
    return out


def suite_inputs(dim):
    src = open(os.path.join(LU, "PICellerator/Weights/tests/DVCWeightsSuite.c")).read()
    name = "pList" if dim == 3 else "pList2D"
    cols = []
    for ax in "xyz"[:dim]:
        vals = re.findall(r"data->%s\[i\]\.%s = ([-+\d.]+)\s*;\s*i\+\+;" % (name, ax), src)
        cols.append([float(v) for v in vals])
    return np.array(cols).T.tolist()


def parse_grid(path):
    txt = open(path).read()
    nb = re.findall(r"Values: \(N: (-?\d+), S: (-?\d+), E: (-?\d+), W: (-?\d+), U: (-?\d+), D: (-?\d+)\)", txt)
    co = re.findall(r"Coords: \(\s*([-\d.]+),\s*([-\d.]+),\s*([-\d.]+)\)", txt)
    return {"nb": [[int(x) for x in r] for r in nb], "coords": [[float(x) for x in r] for r in co]}


def parse_petsc(path):
    b = open(path, "rb").read()
    h = np.frombuffer(b[:16], ">i4")
    M, N, nz = int(h[1]), int(h[2]), int(h[3])
    rl = np.frombuffer(b[16:16 + 4 * M], ">i4")
    o = 16 + 4 * M
    ci = np.frombuffer(b[o:o + 4 * nz], ">i4")
    o += 4 * nz
    v = np.frombuffer(b[o:o + 8 * nz], ">f8")
    A = np.zeros((M, N))
    k = 0
    for i in range(M):
        for _ in range(rl[i]):
            A[i, ci[k]] = v[k]
            k += 1
    return A


def solcx_values():
    so = os.path.join(HERE, "..", "..", "oracle", "_ref", "libsolcx.so")
    if not os.path.exists(so):
        return None
    lib = ctypes.CDLL(so)
    f = lib._Velic_solCx
    dp = ctypes.POINTER(ctypes.c_double)
    f.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp, dp, dp]
    rng = np.random.default_rng(7)
    pts = np.concatenate([rng.random((200, 2)), np.array([[0.25, 0.5], [0.75, 0.5], [0.5, 0.25]])])
    res = {}
    for etaB in (1.0, 1.0e6):
        V, P = [], []
        for x in pts:
            pos = np.ascontiguousarray(x, dtype=np.float64)
            vel = np.zeros(2)
            pr = ctypes.c_double()
            ts = np.zeros(3)
            sr = np.zeros(3)
            f(pos.ctypes.data_as(dp), 1.0, etaB, 0.5, 1, vel.ctypes.data_as(dp), ctypes.byref(pr),
              ts.ctypes.data_as(dp), sr.ctypes.data_as(dp))
            V.append(vel.copy())
            P.append(pr.value)
        res[f"vel_etaB{etaB:g}"] = np.array(V)
        res[f"p_etaB{etaB:g}"] = np.array(P)
    res["points"] = pts
    return res


def main():
    wdir = os.path.join(LU, "PICellerator/Weights/tests/expected")
    for dim, fn, out in ((3, "DVCWeightsSuite_testGetCentroids.txt", "dvc_3d.json"),
                         (2, "DVCWeightsSuite_testGetCentroids2D.txt", "dvc_2d.json")):
        exp = parse_centroids(os.path.join(wdir, fn), dim)
        inp = suite_inputs(dim)
        assert len(exp) == 10 and len(inp) == 10, (len(exp), len(inp))
        json.dump({"resolution": 10, "inputs": inp, "expected": exp,
                   "source": f"LU/PICellerator/Weights/tests/expected/{fn}"}, open(os.path.join(HERE, out), "w"),
                  indent=1)
    g = parse_grid(os.path.join(wdir, "DVCWeightsSuite_testConstructGrid.txt"))
    json.dump(g, open(os.path.join(HERE, "dvc_grid_3d.json"), "w"))
    A = parse_petsc(os.path.join(LU, "StgFEM/Assembly/tests/expected/testIsoviscousSiffness.matrix"))
    np.savez(os.path.join(HERE, "isoviscous_K.npz"), K=A)
    json.dump({
        "random_layout_seed13_ppc4_unit_square": [[0.24261743, 0.67115852], [0.16116546, 0.70790335],
                                                  [0.73160516, 0.08792286], [0.71953113, 0.15966135]],
        "rk2_single_particle_9x9": {"start": [0.2, -0.2], "end": [0.27856742, -0.12143258], "rtol": 1e-4},
        "source": "PY/swarm/layouts.py:200-214; PY/systems/_timeintegration.py:131-147",
    }, open(os.path.join(HERE, "docstrings.json"), "w"), indent=1)
    s = solcx_values()
    if s is not None:
        np.savez(os.path.join(HERE, "solcx.npz"), **s)
    else:
        print("oracle/_ref/libsolcx.so not built; solcx.npz not written", file=sys.stderr)
    shutil.copyfile(os.path.join(REF, "docs/examples/input/03_BlankenbachBenchmark/tempfield_inp_64_Ra1e4.h5"),
                    os.path.join(HERE, "tempfield_inp_64_Ra1e4.h5"))
    print("fixtures written")


if __name__ == "__main__":
    main()
