"""CPU oracle for the Q2/dPc1 Stokes discretisation -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the quadratic-velocity / discontinuous-linear-pressure element pair of
the reference, used by ``tests/`` to check the GPU path.  The product package never imports
it.  ``LU/`` is ``src/underworld/libUnderworld/``, ``PY/`` is ``src/underworld/``.

* Q2 mesh: (2n+1)^d vertices, x fastest (C2Generator.c:147-260); element-local vertices
  lexicographic on the 3^d sub-grid, local index a = i + 3 j + 9 k.
* Q2 shape functions: tensor products of the 1-D quadratic Lagrange polynomials on
  {-1, 0, 1} (Biquadratic.c:137-160, Triquadratic.c).
* dPc1 pressure: d+1 values per element, stored element-major (Inner2DGenerator.c:258-283),
  interpolated with the reference's nodal linear basis in element-local coordinates
  (BilinearInnerElType.c SF_allNodes, TrilinearInnerElType.c SF_allNodes).
* K_e = sum_p c_p B^T diag(2,2,2,1,1,1) B, c_p = eta_p w_p detJ
  (ConstitutiveMatrixCartesian.cpp:238-429); f_e = sum_p w_p detJ f_p N_a
  (VectorAssemblyTerm_NA__Fn.cpp:143-202); G_e = -sum_g w_g detJ dN_a/dx_i N^p_r at the 5^d
  Gauss points of the Q2 mesh (GradientStiffnessMatrixTerm.c:129-221,
  PY/swarm/_integration_swarm.py:173-181); approxS_e = G_e^T diag(K_e)^-1 G_e
  (UzawaPreconditionerTerm.c:124-209).
* Solve: bordered sparse direct solve with sum(p) = 0 (the constant pressure is the null
  mode of G with free-slip walls; the dPc1 basis is a partition of unity).

Pins (tests/test_q2_oracle.py): the reference's SolCx Q2/dPc1 32^2 v_rms
0.005578279626793198 at rtol 1e-7 (docs/examples/10_Analytic_Solutions.ipynb, last cell);
partition of unity and nodality of both bases; rigid-body modes in the kernel of K_e;
G_e^T u = -int div(u) N^p for polynomial u (closed form); the particle sum of K_e against
the Gauss-rule K_e when the particles are the Gauss points; v_rms of a quadratic field.
"""
import numpy as np

_A5 = np.sqrt(5.0 - 2.0 * np.sqrt(10.0 / 7.0)) / 3.0
_B5 = np.sqrt(5.0 + 2.0 * np.sqrt(10.0 / 7.0)) / 3.0
_WA5 = (322.0 + 13.0 * np.sqrt(70.0)) / 900.0
_WB5 = (322.0 - 13.0 * np.sqrt(70.0)) / 900.0
GAUSS5_X = np.array([-_B5, -_A5, 0.0, _A5, _B5])
GAUSS5_W = np.array([_WB5, _WA5, 128.0 / 225.0, _WA5, _WB5])


def phi1(x):
    """1-D quadratic Lagrange basis on {-1, 0, 1} and its derivative, shape [..., 3]."""
    x = np.asarray(x, dtype=np.float64)
    v = np.stack([0.5 * x * (x - 1.0), 1.0 - x * x, 0.5 * x * (x + 1.0)], -1)
    dv = np.stack([x - 0.5, -2.0 * x, x + 0.5], -1)
    return v, dv


def q2_shape(dim, xi):
    """N [P][3^d] and local derivatives dN [P][d][3^d] at local points xi [P][d]."""
    xi = np.atleast_2d(xi)
    P = xi.shape[0]
    v, dv = zip(*[phi1(xi[:, d]) for d in range(dim)])
    nn = 3 ** dim
    N = np.ones((P, nn))
    dN = np.ones((P, dim, nn))
    for a in range(nn):
        idx = [(a // 3 ** d) % 3 for d in range(dim)]
        for d in range(dim):
            N[:, a] *= v[d][:, idx[d]]
            for k in range(dim):
                dN[:, k, a] *= (dv[d][:, idx[d]] if k == d else v[d][:, idx[d]])
    return N, dN


def dpc1_shape(dim, xi):
    """The reference's dPc1 basis [P][d+1] at local points (nodal at the inner vertices)."""
    xi = np.atleast_2d(xi)
    x, y = xi[:, 0], xi[:, 1]
    if dim == 2:
        return np.stack([-x - 0.5 * y + 0.25, x - 0.5 * y + 0.25, y + 0.5], -1)
    z = xi[:, 2]
    f = 1.0 / 9.0
    return np.stack([f * (-4.0 * x - 20.0 * y + 12.0 * z + 3.0), f * (16.0 * x - 64.0 * y + 60.0 * z + 6.0),
                     f * (36.0 * y - 36.0 * z), f * (-12.0 * x + 48.0 * y - 36.0 * z)], -1)


DPC1_NODES = {2: np.array([[-0.5, -0.5], [0.5, -0.5], [0.0, 0.5]]),
              3: np.array([[-0.5, -0.5, -0.5], [0.25, 0.25, 0.25], [0.5, -0.25, -0.5], [-0.25, 0.5, 0.5]])}


def gauss5(dim):
    """Tensor 5-point Gauss rule: points [5^d][d] (x fastest) and weights."""
    g = np.meshgrid(*([GAUSS5_X] * dim), indexing="ij")
    w = np.meshgrid(*([GAUSS5_W] * dim), indexing="ij")
    pts = np.stack([a.transpose().reshape(-1) for a in g], -1)
    wts = np.prod(np.stack([a.transpose().reshape(-1) for a in w], -1), -1)
    return pts, wts


class Q2Mesh:
    def __init__(self, res, lo, hi):
        self.dim = len(res)
        self.res = tuple(int(r) for r in res)
        self.lo = np.array(lo, dtype=np.float64)
        self.hi = np.array(hi, dtype=np.float64)
        self.h = (self.hi - self.lo) / np.array(self.res)
        self.nn = tuple(2 * r + 1 for r in self.res)
        self.N = int(np.prod(self.nn))
        self.E = int(np.prod(self.res))
        self.nen = 3 ** self.dim
        self.npe = self.dim + 1
        self.detJ = float(np.prod(0.5 * self.h))

    def coords(self):
        """Vertex coordinates, x_i = lo + i/(2n) (hi - lo) with the end points exact."""
        axes = []
        for d in range(self.dim):
            n = self.nn[d] - 1
            t = self.lo[d] + (np.arange(self.nn[d]) / n) * (self.hi[d] - self.lo[d])
            t[-1] = self.hi[d]
            axes.append(t)
        g = np.meshgrid(*axes[::-1], indexing="ij")
        return np.stack([a.reshape(-1) for a in g[::-1]], -1)

    def element_index(self):
        e = np.arange(self.E)
        ijk = []
        for d in range(self.dim):
            ijk.append(e % self.res[d])
            e = e // self.res[d]
        return ijk

    def element_nodes(self):
        ijk = self.element_index()
        out = np.zeros((self.E, self.nen), dtype=np.int64)
        for a in range(self.nen):
            n = 0
            stride = 1
            for d in range(self.dim):
                n = n + (2 * ijk[d] + (a // 3 ** d) % 3) * stride
                stride *= self.nn[d]
            out[:, a] = n
        return out

    def element_centres(self):
        ijk = self.element_index()
        return np.stack([self.lo[d] + (ijk[d] + 0.5) * self.h[d] for d in range(self.dim)], -1)

    def local_to_global(self, xi):
        """[E][P][d] physical points of local points xi [P][d] in every element."""
        c = self.element_centres()
        return c[:, None, :] + xi[None, :, :] * (0.5 * self.h)[None, None, :]

    def freeslip_mask(self):
        idx = np.indices(self.nn[::-1]).reshape(self.dim, -1)[::-1]
        m = np.zeros((self.N, self.dim), dtype=bool)
        for d in range(self.dim):
            m[:, d] = (idx[d] == 0) | (idx[d] == self.nn[d] - 1)
        return m


def element_K(mesh, xi, c):
    """K_e [E][nen*d][nen*d] from points xi [E][P][d] (or [P][d]) with weights c [E][P]
    (c = eta w detJ)."""
    d = mesh.dim
    E = c.shape[0]
    xi = np.broadcast_to(xi, (E,) + np.shape(xi)[-2:]) if np.ndim(xi) == 2 else xi
    P = xi.shape[1]
    _, dN = q2_shape(d, xi.reshape(-1, d))
    dN = dN.reshape(E, P, d, mesh.nen) * (2.0 / mesh.h)[None, None, :, None]   # dN/dx
    # K[a i][b j] = sum_p c (delta_ij dNa.dNb + dNa_j dNb_i)
    S = np.einsum("ep,epka,epkb->eab", c, dN, dN)
    T = np.einsum("ep,epja,epib->eaibj", c, dN, dN)
    K = T.copy()
    for i in range(d):
        K[:, :, i, :, i] += S
    return K.reshape(E, mesh.nen * d, mesh.nen * d)


def element_f(mesh, xi, cf):
    """f_e [E][nen][d] from points xi [E][P][d] with cf [E][P][d] = w detJ f(x_p)."""
    d = mesh.dim
    E = cf.shape[0]
    xi = np.broadcast_to(xi, (E,) + np.shape(xi)[-2:]) if np.ndim(xi) == 2 else xi
    N, _ = q2_shape(d, xi.reshape(-1, d))
    N = N.reshape(E, -1, mesh.nen)
    return np.einsum("epa,epi->eai", N, cf)


def element_G(mesh):
    """G_e [nen*d][d+1] (identical for every element of the regular mesh)."""
    d = mesh.dim
    pts, wts = gauss5(d)
    _, dN = q2_shape(d, pts)
    dN = dN * (2.0 / mesh.h)[None, :, None]
    Np = dpc1_shape(d, pts)
    G = -np.einsum("g,gia,gr->air", wts * mesh.detJ, dN, Np)
    return G.reshape(mesh.nen * d, d + 1)


def element_approxS(Ke, Ge):
    """approxS_e = G_e^T diag(K_e)^-1 G_e, [E][d+1][d+1]."""
    dg = np.einsum("eii->ei", Ke)
    return np.einsum("vr,ev,vs->ers", Ge, 1.0 / dg, Ge)


def gauss_points_phys(mesh):
    pts, wts = gauss5(mesh.dim)
    return pts, wts, mesh.local_to_global(pts)


def gauss_assembly(mesh, fn_eta, fn_force):
    """Gauss-swarm assembly (no voronoi swarm): returns Ke, fe."""
    pts, wts, X = gauss_points_phys(mesh)
    E, P, d = X.shape
    eta = np.asarray(fn_eta(X.reshape(-1, d)), dtype=np.float64).reshape(E, P)
    f = np.asarray(fn_force(X.reshape(-1, d)), dtype=np.float64).reshape(E, P, d)
    c = eta * (wts * mesh.detJ)[None, :]
    cf = f * (wts * mesh.detJ)[None, :, None]
    return element_K(mesh, pts, c), element_f(mesh, pts, cf)


class Q2System:
    """Assembled Q2/dPc1 Stokes system with Dirichlet dofs removed (values zero)."""

    def __init__(self, mesh, bcmask, Ke, fe, Ge):
        import scipy.sparse as sp
        d = mesh.dim
        nd = mesh.nen * d
        en = mesh.element_nodes()
        dofs = (en[:, :, None] * d + np.arange(d)[None, None, :]).reshape(mesh.E, nd)
        ndof = mesh.N * d
        bcflat = np.asarray(bcmask, dtype=bool).reshape(-1)
        eq = -np.ones(ndof, dtype=np.int64)
        free = np.flatnonzero(~bcflat)
        eq[free] = np.arange(len(free))
        self.eq, self.free, self.nfree = eq, free, len(free)
        rows = np.repeat(dofs, nd, axis=1).reshape(-1)
        cols = np.tile(dofs, (1, nd)).reshape(-1)
        keep = (eq[rows] >= 0) & (eq[cols] >= 0)
        self.K = sp.csr_matrix((Ke.reshape(-1)[keep], (eq[rows[keep]], eq[cols[keep]])),
                               shape=(self.nfree, self.nfree))
        npr = mesh.npe
        prow = np.repeat(dofs, npr, axis=1).reshape(-1)
        pcol = np.tile((np.arange(mesh.E)[:, None] * npr + np.arange(npr)[None, :]), (1, nd)).reshape(-1)
        Gfull = np.broadcast_to(Ge, (mesh.E,) + Ge.shape).reshape(-1)
        keep = eq[prow] >= 0
        self.G = sp.csr_matrix((Gfull[keep], (eq[prow[keep]], pcol[keep])), shape=(self.nfree, mesh.E * npr))
        f = np.zeros(ndof)
        np.add.at(f, dofs.reshape(-1), fe.reshape(-1))
        self.f_full = f
        self.f = f[free]
        self.mesh = mesh
        self.ndof = ndof

    def solve(self):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spl
        npt = self.mesh.E * self.mesh.npe
        one = sp.csr_matrix(np.ones((1, npt)))
        A = sp.bmat([[self.K, self.G, None], [self.G.T, None, one.T], [None, one, None]], format="csc")
        rhs = np.concatenate([self.f, np.zeros(npt), [0.0]])
        x = spl.spsolve(A, rhs)
        u = np.zeros(self.ndof)
        u[self.free] = x[:self.nfree]
        return u.reshape(self.mesh.N, self.mesh.dim), x[self.nfree:self.nfree + npt]

    def residuals(self, u, p):
        uf = np.asarray(u).reshape(-1)[self.free]
        r1 = self.f - self.K @ uf - self.G @ p
        r2 = self.G.T @ uf
        return float(np.linalg.norm(r1)), float(np.linalg.norm(r2))


def velocity_rms(mesh, u):
    """sqrt(int v.v / int 1) with the 5^d Gauss rule (exact for Q2 fields)."""
    pts, wts = gauss5(mesh.dim)
    N, _ = q2_shape(mesh.dim, pts)
    en = mesh.element_nodes()
    v = np.einsum("ga,eai->egi", N, np.asarray(u)[en])
    v2 = np.einsum("g,egi,egi->", wts * mesh.detJ, v, v)
    vol = float(np.prod(mesh.hi - mesh.lo))
    return float(np.sqrt(v2 / vol))


def solcx_functions(eta_a=1.0, eta_b=1.0e5, x_c=0.75, n_z=1):
    """SolCx viscosity step and body force (LU/Underworld/Function/src/AnalyticSolCx.hpp:39-47)."""
    def eta(x):
        return np.where(x[:, 0] < x_c, eta_a, eta_b)

    def force(x):
        return np.stack([np.zeros(x.shape[0]), np.cos(np.pi * x[:, 0]) * np.sin(np.pi * n_z * x[:, 1])], -1)
    return eta, force


def solve_gauss_problem(res, fn_eta, fn_force, lo=None, hi=None):
    d = len(res)
    mesh = Q2Mesh(res, lo if lo is not None else (0.0,) * d, hi if hi is not None else (1.0,) * d)
    Ke, fe = gauss_assembly(mesh, fn_eta, fn_force)
    Ge = element_G(mesh)
    sysm = Q2System(mesh, mesh.freeslip_mask(), Ke, fe, Ge)
    u, p = sysm.solve()
    return dict(mesh=mesh, Ke=Ke, fe=fe, Ge=Ge, system=sysm, u=u, p=p)
