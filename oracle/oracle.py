"""CPU oracle for the particle-in-cell Stokes path -- TEST INFRASTRUCTURE ONLY.

This module is the parity checker and the CPU baseline.  It wraps ``liboracle.so`` (the
scalar C restatement in ``uwg_oracle.c``) with numpy, and solves the assembled discrete
Stokes system with scipy's sparse direct solver.  Only ``tests/``, ``__graft_entry__.smoke``
and ``bench.py``'s ``cpu_baseline`` leg import it; the product package never does.

Reference citations are relative to the reference tree; ``LU/`` is
``src/underworld/libUnderworld/`` and ``PY/`` is ``src/underworld/``.

Pins (see also the header of ``uwg_oracle.c``):
  * random layout: PY/swarm/layouts.py:200-214 docstring values;
  * DVC: LU/PICellerator/Weights/tests/expected/DVCWeightsSuite_testGetCentroids{,2D}.txt;
  * Gauss K: LU/StgFEM/Assembly/tests/expected/testIsoviscousSiffness.matrix;
  * RK2: PY/systems/_timeintegration.py:131-147 doctest;
  * discrete solve: the bordered sparse direct solve is checked against manufactured
    solutions (f = K u*, residual ~ 1e-13) and against Velic solCx convergence rates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_search_element.restype = ctypes.c_int
        L.orc_cell_of.restype = ctypes.c_long
        L.orc_newton_local.restype = ctypes.c_int
        L.orc_dvc_cell.restype = ctypes.c_int
        L.orc_dvc_all.restype = ctypes.c_int
        L.orc_element_approxS.restype = ctypes.c_double
        L.orc_advect_rk2.restype = ctypes.c_long
        L.orc_advect_rk2_q2.restype = ctypes.c_long
        L.orc_q2_newton_local.restype = ctypes.c_int
        L.orc_q2_min_sep.restype = ctypes.c_double
        L.orc_max_dt.restype = ctypes.c_double
        L.orc_update_owners.restype = ctypes.c_long
        L.orc_bsscr_solve.restype = ctypes.c_int
        L.orc_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _l(a):
    return a.ctypes.data_as(_lp)


class Mesh:
    """Regular Q1/dQ0 box mesh description (LU/StgDomain/Mesh/src/CartesianGenerator.c)."""

    def __init__(self, res, lo, hi):
        self.dim = len(res)
        self.res = np.ascontiguousarray(res, dtype=np.int32)
        self.lo = np.ascontiguousarray(lo, dtype=np.float64)
        self.hi = np.ascontiguousarray(hi, dtype=np.float64)
        self.nnodes = tuple(int(r) + 1 for r in res)
        self.N = int(np.prod(self.nnodes))
        self.E = int(np.prod(self.res))
        self.nn = 2 ** self.dim

    def args(self):
        return (ctypes.c_int(self.dim), _i(self.res), _d(self.lo), _d(self.hi))

    def coords(self):
        out = np.zeros((self.N, self.dim))
        lib().orc_mesh_coords(*self.args(), _d(out))
        return out

    def element_nodes(self):
        """[E][2^d] node indices, element-local lexicographic order, x fastest."""
        d = self.dim
        nx = self.nnodes[0]
        ny = self.nnodes[1]
        if d == 2:
            j, i = np.meshgrid(np.arange(self.res[1]), np.arange(self.res[0]), indexing="ij")
            base = (i + nx * j).ravel()
            return np.stack([base, base + 1, base + nx, base + nx + 1], axis=1)
        k, j, i = np.meshgrid(np.arange(self.res[2]), np.arange(self.res[1]), np.arange(self.res[0]), indexing="ij")
        base = (i + nx * (j + ny * k)).ravel()
        s = nx * ny
        return np.stack([base, base + 1, base + nx, base + nx + 1,
                         base + s, base + s + 1, base + s + nx, base + s + nx + 1], axis=1)

    def element_verts(self, e):
        c = self.coords()
        return np.ascontiguousarray(c[self.element_nodes()[e]])

    def freeslip_mask(self):
        """[N][d] bool: True where the dof is a Dirichlet (normal-velocity) dof."""
        idx = np.indices(self.nnodes[::-1]).reshape(self.dim, -1)[::-1]
        m = np.zeros((self.N, self.dim), dtype=bool)
        for d in range(self.dim):
            m[:, d] = (idx[d] == 0) | (idx[d] == self.nnodes[d] - 1)
        return m


def layout_random(mesh, ppc, seed):
    P = mesh.E * ppc
    out = np.zeros((P, mesh.dim))
    lib().orc_layout_random(*mesh.args(), ctypes.c_int(ppc), ctypes.c_long(seed), _d(out))
    return out


def layout_sobol(mesh, ppc, seed=13):
    P = mesh.E * ppc
    out = np.zeros((P, mesh.dim))
    lib().orc_layout_sobol(*mesh.args(), ctypes.c_int(ppc), ctypes.c_long(seed), _d(out))
    return out


def sobol_stream(axis, n):
    out = np.zeros(n)
    lib().orc_sobol_stream(ctypes.c_int(axis), ctypes.c_long(n), _d(out))
    return out


def rng_stream(seed, n):
    out = np.zeros(n)
    lib().orc_rng_stream(ctypes.c_long(seed), ctypes.c_long(n), _d(out))
    return out


def newton_local(dim, verts, x):
    xi = np.zeros(dim)
    its = lib().orc_newton_local(ctypes.c_int(dim), _d(np.ascontiguousarray(verts, dtype=np.float64)),
                                 _d(np.ascontiguousarray(x, dtype=np.float64)), _d(xi))
    return xi, its


def cell_of(mesh, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return np.array([lib().orc_cell_of(*mesh.args(), _d(np.ascontiguousarray(p))) for p in x.reshape(-1, mesh.dim)])


def initial_cells(mesh, ppc):
    """CSR cell tables after PerCellParticleLayout: cell e holds e*ppc .. e*ppc+ppc-1."""
    ptr = np.arange(mesh.E + 1, dtype=np.int64) * ppc
    pid = np.arange(mesh.E * ppc, dtype=np.int32)
    return ptr, pid


def coincident_map(mesh, cell_ptr, cell_pid, coords):
    xi = np.zeros((len(cell_pid), mesh.dim))
    lib().orc_coincident_map(*mesh.args(), _l(np.ascontiguousarray(cell_ptr, dtype=np.int64)),
                             _i(np.ascontiguousarray(cell_pid, dtype=np.int32)),
                             _d(np.ascontiguousarray(coords, dtype=np.float64)), _d(xi))
    return xi


def dvc_cell(dim, xi, num=15):
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    n = xi.shape[0]
    numv = np.array([num] * 3, dtype=np.int32)
    out = np.zeros_like(xi)
    w = np.zeros(n)
    rc = lib().orc_dvc_cell(ctypes.c_int(dim), _i(numv), ctypes.c_int(n), _d(xi), _d(out), _d(w))
    if rc != 0:
        raise RuntimeError("DVC: under-resolved cell (no particles)")
    return out, w


def dvc_all(dim, cell_ptr, xi_slot, num=15):
    xi_slot = np.ascontiguousarray(xi_slot, dtype=np.float64)
    out = np.zeros_like(xi_slot)
    w = np.zeros(xi_slot.shape[0])
    numv = np.array([num] * 3, dtype=np.int32)
    E = len(cell_ptr) - 1
    rc = lib().orc_dvc_all(ctypes.c_int(dim), _i(numv), ctypes.c_long(E),
                           _l(np.ascontiguousarray(cell_ptr, dtype=np.int64)), _d(xi_slot), _d(out), _d(w))
    if rc != 0:
        raise RuntimeError("DVC: under-resolved cell")
    return out, w


def dvc_grid_dump(dim, num):
    numv = np.array(list(num) + [1] * (3 - len(num)), dtype=np.int32)
    n = int(np.prod(num))
    nb = np.zeros((n, 6), dtype=np.int32)
    c = np.zeros((n, 3))
    lib().orc_dvc_grid_dump(ctypes.c_int(dim), _i(numv), _i(nb), _d(c))
    return nb, c


def element_K_gauss(dim, verts, eta=1.0):
    nd = 2 ** dim * dim
    Ke = np.zeros((nd, nd))
    lib().orc_element_K_gauss(ctypes.c_int(dim), _d(np.ascontiguousarray(verts, dtype=np.float64)),
                              ctypes.c_double(eta), _d(Ke))
    return Ke


def element_K_particles(dim, verts, xi, w, eta):
    nd = 2 ** dim * dim
    Ke = np.zeros((nd, nd))
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    lib().orc_element_K_particles(ctypes.c_int(dim), _d(np.ascontiguousarray(verts, dtype=np.float64)),
                                  ctypes.c_int(xi.shape[0]), _d(xi),
                                  _d(np.ascontiguousarray(w, dtype=np.float64)),
                                  _d(np.ascontiguousarray(eta, dtype=np.float64)), _d(Ke))
    return Ke


def element_G_gauss(dim, verts):
    Ge = np.zeros(2 ** dim * dim)
    lib().orc_element_G_gauss(ctypes.c_int(dim), _d(np.ascontiguousarray(verts, dtype=np.float64)), _d(Ge))
    return Ge


def assemble_elements(mesh, cell_ptr, xi_slot, w_slot, eta_slot, force_slot):
    nd = mesh.nn * mesh.dim
    E = mesh.E
    Ke = np.zeros((E, nd, nd))
    fe = np.zeros((E, nd))
    Ge = np.zeros((E, nd))
    S = np.zeros(E)
    lib().orc_assemble_elements(*mesh.args(), _l(np.ascontiguousarray(cell_ptr, dtype=np.int64)),
                                _d(np.ascontiguousarray(xi_slot, dtype=np.float64)),
                                _d(np.ascontiguousarray(w_slot, dtype=np.float64)),
                                _d(np.ascontiguousarray(eta_slot, dtype=np.float64)),
                                _d(np.ascontiguousarray(force_slot, dtype=np.float64)),
                                _d(Ke), _d(fe), _d(Ge), _d(S))
    return Ke, fe, Ge, S


def advect_rk2(mesh, vel, coords, dt):
    c = np.ascontiguousarray(coords, dtype=np.float64).copy()
    fb = ctypes.c_long(0)
    bad = lib().orc_advect_rk2(*mesh.args(), _d(np.ascontiguousarray(vel, dtype=np.float64)),
                               ctypes.c_long(c.shape[0]), _d(c), ctypes.c_double(dt), ctypes.byref(fb))
    return c, int(bad), int(fb.value)


def interpolate(mesh, vel, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    P = x.shape[0]
    v = np.zeros_like(x)
    ok = np.zeros(P, dtype=np.int32)
    lib().orc_interpolate(*mesh.args(), _d(np.ascontiguousarray(vel, dtype=np.float64)), ctypes.c_long(P),
                          _d(x), _d(v), _i(ok))
    return v, ok.astype(bool)


def max_dt(mesh, vel):
    return lib().orc_max_dt(*mesh.args(), _d(np.ascontiguousarray(vel, dtype=np.float64)))


# --- lookups in a Q2 velocity field (vertex grid of 2 res + 1 per direction on the same element
# grid); see the Q2 section of uwg_oracle.c for the reference citations and pins
def q2_vertex_coords(mesh):
    """[(2r+1)^d][d] vertex coordinates of the Q2 mesh on ``mesh``'s element grid."""
    fine = Mesh(tuple(2 * int(r) for r in mesh.res), mesh.lo, mesh.hi)
    return fine.coords()


def q2_basis(dim, xi):
    """(N[3^d], G[d][3^d]) of the Biquadratic / Triquadratic element types at xi."""
    nn = 3 ** dim
    N = np.zeros(27)
    G = np.zeros((3, 27))
    lib().orc_q2_basis(ctypes.c_int(dim), _d(np.ascontiguousarray(xi, dtype=np.float64)), _d(N), _d(G))
    return N[:nn], G[:dim, :nn]


def q2_newton_local(dim, verts, x):
    xi = np.zeros(3)
    it = lib().orc_q2_newton_local(ctypes.c_int(dim), _d(np.ascontiguousarray(verts, dtype=np.float64)),
                                   _d(np.ascontiguousarray(x, dtype=np.float64)), _d(xi))
    return xi[:dim], int(it)


def advect_rk2_q2(mesh, vel, coords, dt):
    c = np.ascontiguousarray(coords, dtype=np.float64).copy()
    fb = ctypes.c_long(0)
    bad = lib().orc_advect_rk2_q2(*mesh.args(), _d(np.ascontiguousarray(vel, dtype=np.float64)),
                                  ctypes.c_long(c.shape[0]), _d(c), ctypes.c_double(dt), ctypes.byref(fb))
    return c, int(bad), int(fb.value)


def interpolate_q2(mesh, vel, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    P = x.shape[0]
    v = np.zeros_like(x)
    ok = np.zeros(P, dtype=np.int32)
    lib().orc_interpolate_q2(*mesh.args(), _d(np.ascontiguousarray(vel, dtype=np.float64)), ctypes.c_long(P),
                             _d(x), _d(v), _i(ok))
    return v, ok.astype(bool)


def max_dt_q2(mesh, vel):
    """0.5 * minSep(Q2 mesh) / max_vertices |v| (SwarmAdvector_MaxDt on a Q2 velocity field)."""
    ms = lib().orc_q2_min_sep(*mesh.args())
    vmax = float(np.max(np.sqrt(np.sum(np.asarray(vel, dtype=np.float64) ** 2, axis=1))))
    return 0.5 * ms / vmax


class CellTables:
    """Per-cell particle tables with the reference's swap-with-last / append edits."""

    def __init__(self, mesh, cell_ptr, cell_pid, cap=None):
        E = mesh.E
        counts = np.diff(cell_ptr).astype(np.int32)
        self.cap = int(cap or max(64, 4 * counts.max()))
        self.tbl = np.full((E, self.cap), -1, dtype=np.int32)
        for e in range(E):
            self.tbl[e, :counts[e]] = cell_pid[cell_ptr[e]:cell_ptr[e + 1]]
        self.cnt = counts.copy()
        self.owner = np.zeros(len(cell_pid), dtype=np.int32)
        for e in range(E):
            self.owner[cell_pid[cell_ptr[e]:cell_ptr[e + 1]]] = e
        self.mesh = mesh

    def update(self, coords):
        ov = ctypes.c_int(0)
        out = lib().orc_update_owners(*self.mesh.args(), ctypes.c_long(coords.shape[0]),
                                      _d(np.ascontiguousarray(coords, dtype=np.float64)),
                                      _i(self.owner), _i(self.tbl), _i(self.cnt), ctypes.c_int(self.cap),
                                      ctypes.byref(ov))
        if ov.value:
            raise RuntimeError("cell table capacity exceeded")
        return int(out)

    def csr(self):
        ptr = np.zeros(self.mesh.E + 1, dtype=np.int64)
        ptr[1:] = np.cumsum(self.cnt)
        pid = np.concatenate([self.tbl[e, :self.cnt[e]] for e in range(self.mesh.E)]).astype(np.int32)
        return ptr, pid


# ----------------------------------------------------------------------------------------
# Assembled system (A5) and the FP64 discrete solve (A10-A14) by sparse direct factorization
# ----------------------------------------------------------------------------------------
class System:
    """Assembled K (free dofs), G, f, approxS (StiffnessMatrix.c:552-744, removeBCs=True)."""

    def __init__(self, mesh, bcmask, Ke, fe, Ge, S):
        import scipy.sparse as sp
        d = mesh.dim
        nd = mesh.nn * d
        en = mesh.element_nodes()
        dofs = (en[:, :, None] * d + np.arange(d)[None, None, :]).reshape(mesh.E, nd)
        ndof = mesh.N * d
        bcflat = bcmask.reshape(-1)
        eq = -np.ones(ndof, dtype=np.int64)
        free = np.flatnonzero(~bcflat)
        eq[free] = np.arange(len(free))
        self.eq = eq
        self.free = free
        self.nfree = len(free)
        rows = np.repeat(dofs, nd, axis=1).reshape(-1)
        cols = np.tile(dofs, (1, nd)).reshape(-1)
        vals = Ke.reshape(-1)
        keep = (eq[rows] >= 0) & (eq[cols] >= 0)
        self.K = sp.csr_matrix((vals[keep], (eq[rows[keep]], eq[cols[keep]])), shape=(self.nfree, self.nfree))
        r = dofs.reshape(-1)
        e = np.repeat(np.arange(mesh.E), nd)
        keep = eq[r] >= 0
        self.G = sp.csr_matrix((Ge.reshape(-1)[keep], (eq[r[keep]], e[keep])), shape=(self.nfree, mesh.E))
        f = np.zeros(ndof)
        np.add.at(f, dofs.reshape(-1), fe.reshape(-1))
        self.f_full = f
        self.f = f[free]
        self.approxS = S.copy()
        self.mesh = mesh
        self.ndof = ndof

    def solve(self):
        """Exact discrete solution: bordered saddle system with mean(p) = 0 constraint."""
        import scipy.sparse as sp
        import scipy.sparse.linalg as spl
        E = self.mesh.E
        one = sp.csr_matrix(np.ones((1, E)))
        A = sp.bmat([[self.K, self.G, None], [self.G.T, None, one.T], [None, one, None]], format="csc")
        rhs = np.concatenate([self.f, np.zeros(E), [0.0]])
        x = spl.spsolve(A, rhs)
        u = np.zeros(self.ndof)
        u[self.free] = x[:self.nfree]
        p = x[self.nfree:self.nfree + E]
        return u.reshape(self.mesh.N, self.mesh.dim), p

    def bsscr_solve(self, outer_rtol=1e-6, inner_rtol=1e-6, coarse_max=800, max_outer=5000):
        """The reference's iterative solve restated on the CPU (cpu_baseline.c): Schur FGMRES
        with the approxS preconditioner around FGMRES + Galerkin multigrid inner solves, at most
        max_outer outer iterations.  Free-slip walls on every face.  Returns (u, p, stats dict)."""
        K = self.K.tocsr()
        G = self.G.tocsr()
        kp, kc, kv = (np.ascontiguousarray(K.indptr, np.int64), np.ascontiguousarray(K.indices, np.int32),
                      np.ascontiguousarray(K.data, np.float64))
        gp, gc, gv = (np.ascontiguousarray(G.indptr, np.int64), np.ascontiguousarray(G.indices, np.int32),
                      np.ascontiguousarray(G.data, np.float64))
        res = np.asarray(self.mesh.res, dtype=np.int32)
        E = self.mesh.E
        uf = np.zeros(self.nfree)
        p = np.zeros(E)
        st = np.zeros(4, np.int64)
        tm = np.zeros(2)
        f = np.ascontiguousarray(self.f, np.float64)
        S = np.ascontiguousarray(self.approxS, np.float64)
        rc = lib().orc_bsscr_solve(self.mesh.dim, _i(res), ctypes.c_long(self.nfree), _l(kp), _i(kc), _d(kv),
                                   ctypes.c_long(E), _l(gp), _i(gc), _d(gv), _d(S), _d(f),
                                   ctypes.c_double(outer_rtol), ctypes.c_double(inner_rtol), ctypes.c_int(coarse_max),
                                   ctypes.c_int(max_outer), _d(uf), _d(p), _l(st), _d(tm))
        if rc < 0:
            raise RuntimeError("bsscr_solve: free-DOF numbering does not match free-slip walls")
        u = np.zeros(self.ndof)
        u[self.free] = uf
        stats = dict(outer_its=int(st[0]), inner_its=int(st[1]), vcycles=int(st[2]), levels=int(st[3]),
                     setup_time=float(tm[0]), solve_time=float(tm[1]), threads=int(lib().orc_num_threads()))
        return u.reshape(self.mesh.N, self.mesh.dim), p, stats

    def residuals(self, u, p):
        """(|f - K u - G p|, |G^T u|) on free dofs (Stokes_SLE.c:204-256)."""
        uf = u.reshape(-1)[self.free]
        r1 = self.f - self.K @ uf - self.G @ p
        r2 = self.G.T @ uf
        return float(np.linalg.norm(r1)), float(np.linalg.norm(r2))


def full_pipeline(dim, res, lo, hi, ppc, seed, material_fn, eta_table, rho_table, gravity, dvc=True, solve=True):
    """Oracle run of construction + solve: layout -> material -> coincident map -> DVC ->
    element integration -> assembled system -> exact solve.  material_fn(coords) -> int array.
    Returns a dict with all intermediate arrays (slot order where applicable)."""
    mesh = Mesh(res, lo, hi)
    coords = layout_random(mesh, ppc, seed)
    mat = np.asarray(material_fn(coords), dtype=np.int32)
    ptr, pid = initial_cells(mesh, ppc)
    xi = coincident_map(mesh, ptr, pid, coords)
    if dvc:
        xi_w, w = dvc_all(dim, ptr, xi)
    else:
        xi_w, w = xi, np.full(len(pid), 2.0 ** dim / ppc)
    eta = np.asarray(eta_table, dtype=np.float64)[mat[pid]]
    rho = np.asarray(rho_table, dtype=np.float64)[mat[pid]]
    force = rho[:, None] * np.asarray(gravity, dtype=np.float64)[None, :]
    Ke, fe, Ge, S = assemble_elements(mesh, ptr, xi_w, w, eta, force)
    bc = mesh.freeslip_mask()
    sysm = System(mesh, bc, Ke, fe, Ge, S)
    u, p = sysm.solve() if solve else (None, None)
    return dict(mesh=mesh, coords=coords, mat=mat, ptr=ptr, pid=pid, xi_map=xi, xi=xi_w, w=w,
                eta=eta, force=force, Ke=Ke, fe=fe, Ge=Ge, S=S, system=sysm, u=u, p=p)
