"""CPU oracle for SUPG advection-diffusion -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the reference's `AdvectionDiffusion(method="SUPG")` time step on a Q1
mesh, used by ``tests/`` to check the GPU path.  ``LU/`` is ``src/underworld/libUnderworld/``;
``AD/`` is ``LU/Underworld/Utils/AdvectionDiffusion/src/``.

* Residual (AD/Residual.cpp:223-343): per element an upwind diffusivity
  (AD/Residual.cpp:430-513: element-mean diffusivity over the integration points, floored at
  1e-20; velocity at the element centre; per axis Peclet number v h / (2 kappa), upwind
  parameter xi(Pe), sum of xi v h, times 1/sqrt(15)); per integration point the SUPG test
  functions N + (kappa_up / |u|^2) u.grad N, the total derivative phidot + u.grad phi - H,
  and r_i -= w detJ (SUPGN_i (Dphi/Dt) + kappa grad N_i . grad phi).
* Upwind parameter (AD/Residual.cpp:557-606): exact coth(a) - 1/a (default), doubly
  asymptotic, critical.
* Lumped mass (AD/LumpedMassMatrixForceTerm.c:88-129): M_i = sum_p w detJ N_i sum_j N_j.
* Multicorrector (AD/Multicorrector.c:158-260, Brooks & Hughes 1982 sec. 4.2, gamma 0.5,
  2 corrector passes): the predictor acts on the solution vectors (free dofs); the nodal
  fields, which the residual reads, are refreshed from the vectors only at the end of each
  corrector pass, so the first pass sees the fields as they were at the start of the step.
* Time step (AD/Timestep.c:36-110): 0.5 min(min_nodes,axes |h_d / v_d|, h_min^2 / kappa_max).
Integration points: the 3^d Gauss rule of the Q1 mesh (PY/swarm/_integration_swarm.py:173-181).

Pins (tests/test_supg_oracle.py): closed forms of the upwind functions; the lumped mass sums
to the domain volume; a linear field under uniform velocity moves by the trapezoidal-rule
amounts the multicorrector defines (interior nodes); pure diffusion with a source relaxes to
the nodally exact 1-D parabola; the Galerkin residual of a steady state vanishes.
"""
import numpy as np

ISQRT15 = 0.25819888974716112567
SUPG_MIN_DIFFUSIVITY = 1.0e-20
_GA = np.sqrt(0.6)


def gauss3(dim):
    x = np.array([-_GA, 0.0, _GA])
    w = np.array([5.0, 8.0, 5.0]) / 9.0
    g = np.meshgrid(*([x] * dim), indexing="ij")
    ww = np.meshgrid(*([w] * dim), indexing="ij")
    pts = np.stack([a.transpose().reshape(-1) for a in g], -1)
    wts = np.prod(np.stack([a.transpose().reshape(-1) for a in ww], -1), -1)
    return pts, wts


def q1_shape(dim, xi):
    xi = np.atleast_2d(xi)
    nn = 2 ** dim
    N = np.ones((xi.shape[0], nn))
    dN = np.ones((xi.shape[0], dim, nn))
    for a in range(nn):
        s = [1.0 if (a >> d) & 1 else -1.0 for d in range(dim)]
        for d in range(dim):
            N[:, a] *= 0.5 * (1.0 + s[d] * xi[:, d])
            for k in range(dim):
                dN[:, k, a] *= (0.5 * s[d]) if k == d else 0.5 * (1.0 + s[d] * xi[:, d])
    return N, dN


def upwind_xi(pe, kind=0):
    pe = float(pe)
    if kind == 0:
        if abs(pe) < 1.0e-8:
            return 0.33333333333333 * pe
        if pe < -20.0:
            return -1.0 - 1.0 / pe
        if pe > 20.0:
            return 1.0 - 1.0 / pe
        return np.cosh(pe) / np.sinh(pe) - 1.0 / pe
    if kind == 1:
        if pe <= -3.0:
            return -1.0
        if pe <= 3.0:
            return 0.33333333333333 * pe
        return 1.0
    if pe <= -1.0:
        return -1.0 - 1.0 / pe
    if pe <= 1.0:
        return 0.0
    return 1.0 - 1.0 / pe


class Q1Mesh:
    def __init__(self, res, lo, hi):
        self.dim = len(res)
        self.res = tuple(int(r) for r in res)
        self.lo = np.array(lo, dtype=np.float64)
        self.hi = np.array(hi, dtype=np.float64)
        self.nn = tuple(r + 1 for r in self.res)
        self.N = int(np.prod(self.nn))
        self.E = int(np.prod(self.res))
        self.axes = []
        for d in range(self.dim):
            t = self.lo[d] + (np.arange(self.nn[d]) / self.res[d]) * (self.hi[d] - self.lo[d])
            t[-1] = self.hi[d]
            self.axes.append(t)

    def coords(self):
        g = np.meshgrid(*self.axes[::-1], indexing="ij")
        return np.stack([a.reshape(-1) for a in g[::-1]], -1)

    def element_nodes(self):
        e = np.arange(self.E)
        ijk = []
        for d in range(self.dim):
            ijk.append(e % self.res[d])
            e = e // self.res[d]
        out = np.zeros((self.E, 2 ** self.dim), dtype=np.int64)
        for a in range(2 ** self.dim):
            n, stride = 0, 1
            for d in range(self.dim):
                n = n + (ijk[d] + ((a >> d) & 1)) * stride
                stride *= self.nn[d]
            out[:, a] = n
        return out

    def element_h(self):
        """per element and axis: vertex spacing [E][d] (greatest minus least vertex)."""
        X = self.coords()[self.element_nodes()]
        return np.abs(X[:, -1, :] - X[:, 0, :])

    def min_separation(self):
        """(min separation, per-axis separations of the first element attaining it)
        (Mesh_HexType_GetMinimumSeparation + _Mesh_Algorithms_GetMinimumSeparation)."""
        best, per = np.inf, None
        en = self.element_nodes()
        X = self.coords()
        for e in range(self.E):
            v = X[en[e]]
            sep = [v[1, 0] - v[0, 0]] + ([v[2, 1] - v[0, 1]] if self.dim >= 2 else []) + \
                  ([v[4, 2] - v[0, 2]] if self.dim == 3 else [])
            cur = min(sep)
            if cur < best:
                best, per = cur, sep
        return best, np.array(per)

    def gauss_points(self):
        pts, wts = gauss3(self.dim)
        X = self.coords()[self.element_nodes()]
        N, _ = q1_shape(self.dim, pts)
        return pts, wts, np.einsum("ga,ead->egd", N, X)


def residual(mesh, phi, phidot, vel, kappa_g, source_g, upwind=0):
    """Assembled residual on all nodes (Dirichlet rows included; the caller drops them)."""
    d = mesh.dim
    pts, wts, _ = mesh.gauss_points()
    N, dN = q1_shape(d, pts)
    h = mesh.element_h()
    en = mesh.element_nodes()
    detJ = np.prod(h / 2.0, axis=1)
    R = np.zeros(mesh.N)
    Nc, _ = q1_shape(d, np.zeros((1, d)))
    for e in range(mesh.E):
        nodes = en[e]
        kap = kappa_g[e]
        avg = 0.0
        for q in range(len(wts)):
            avg += kap[q]
        avg /= float(len(wts))
        if avg < SUPG_MIN_DIFFUSIVITY:
            avg = SUPG_MIN_DIFFUSIVITY
        vc = Nc[0] @ vel[nodes]
        up = 0.0
        for k in range(d):
            pe = vc[k] * h[e, k] / (2.0 * avg)
            up += upwind_xi(pe, upwind) * vc[k] * h[e, k]
        up *= ISQRT15
        GNx = dN * (2.0 / h[e])[None, :, None]          # [g][d][a]
        r = np.zeros(len(nodes))
        for q in range(len(wts)):
            v = N[q] @ vel[nodes]
            udotu = float(v @ v)
            if abs(up) < SUPG_MIN_DIFFUSIVITY:
                supg = N[q].copy()
            else:
                supg = N[q] + (up / udotu) * (v @ GNx[q])
            ph = N[q] @ phi[nodes]
            gphi = GNx[q] @ phi[nodes]
            pdot = N[q] @ phidot[nodes]
            total = pdot + float(v @ gphi) - source_g[e, q]
            fac = wts[q] * detJ[e]
            diff = kap[q] * (gphi @ GNx[q])
            r -= fac * (supg * total + diff)
            del ph
        np.add.at(R, nodes, r)
    return R


def lumped_mass(mesh):
    d = mesh.dim
    pts, wts = gauss3(d)
    N, _ = q1_shape(d, pts)
    h = mesh.element_h()
    detJ = np.prod(h / 2.0, axis=1)
    Me = np.einsum("g,e,ga,gb->ea", wts, detJ, N, N)
    M = np.zeros(mesh.N)
    np.add.at(M, mesh.element_nodes(), Me)
    return M


class SUPG:
    def __init__(self, mesh, bcmask, gamma=0.5, iterations=2, upwind=0):
        self.mesh = mesh
        self.free = ~np.asarray(bcmask, dtype=bool)
        self.gamma = gamma
        self.iterations = iterations
        self.upwind = upwind
        self.M = lumped_mass(mesh)

    def integrate(self, phi, phidot, vel, kappa_g, source_g, dt):
        """One step; returns the new nodal (phi, phidot)."""
        f = self.free
        phi_n, pdot_n = phi.copy(), phidot.copy()        # nodal fields
        phi_v, pdot_v = phi.copy(), phidot.copy()        # solution vectors (free entries used)
        phi_v[f] += dt * (1.0 - self.gamma) * pdot_v[f]  # predictor
        pdot_v[f] = 0.0
        for _ in range(self.iterations):
            R = residual(self.mesh, phi_n, pdot_n, vel, kappa_g, source_g, self.upwind)
            dp = R[f] / self.M[f]
            phi_v[f] += dt * self.gamma * dp
            pdot_v[f] += dp
            phi_n[f] = phi_v[f]
            pdot_n[f] = pdot_v[f]
        return phi_n, pdot_n

    def max_dt(self, vel, kappa_g):
        sep, per = self.mesh.min_separation()
        adv = np.inf
        for k in range(self.mesh.dim):
            v = vel[:, k]
            nz = v != 0.0
            if nz.any():
                adv = min(adv, float(np.min(np.abs(per[k] / v[nz]))))
        adv = 0.5 * adv
        dif = 0.5 * sep * sep / float(np.max(kappa_g))
        return min(adv, dif)
