"""Pins of the oracle's Schur preconditioner approxS (UzawaPreconditionerTerm.c:124-209:
approxS_e = sum_v G_e[v]^2 / K_e[v][v], the element-local diagonal of K_e) by closed forms.

Q1/dQ0 on a rectangular element (box h_1 x ... x h_d, volume V, detJ = V / 2^d) with Gauss
quadrature, which integrates these polynomial integrands exactly:
  |G_e[a, i]| = int dN_a/dx_i dV = V / (2^(d-1) h_i)                      (every node a)
  K_e[ai][ai] = sum_j c_ij int eta (dN_a/dx_j)^2 dV,  c_ij = 2 if i = j else 1
  with int eta (dN_a/dx_j)^2 dV = detJ (4 / h_j^2) / 4^d prod_k J_k,
  J_k = int (1 + s_k xi)^2 w_k(xi) dxi  (k != j),  J_j = int w_j(xi) dxi,
  where eta = eta_0 prod_k w_k(xi_k) and s_k = +-1 is node a's corner sign.
For uniform eta this gives 2 h^2 / eta on a square and 27 h^3 / (8 eta) on a cube.  A
transposed G, a dropped off-axis term of D, a missing factor 2 on the diagonal of D or a
wrong detJ each change these values.

Q2/dPc1 (2D): one element block approxS_e = G_e^T diag(K_e)^-1 G_e [3][3], computed here by
exact rational integration of the polynomial integrands (no quadrature), for a rectangular
element with eta = 1, against q2_oracle's 5-point Gauss assembly.
"""
from fractions import Fraction as Fr

import numpy as np

from oracle import oracle as O
from oracle import q2_oracle as Q

_GA = np.sqrt(0.6)


def _gauss_slots(m):
    x1 = np.array([-_GA, 0.0, _GA])
    w1 = np.array([5, 8, 5]) / 9.0
    d = m.dim
    grids = np.meshgrid(*([x1] * d), indexing="ij")
    gx = np.stack([g.ravel() for g in grids[::-1]], 1)
    wg = np.meshgrid(*([w1] * d), indexing="ij")
    gw = np.prod(np.stack([g.ravel() for g in wg[::-1]], 0), 0)
    Q_ = len(gw)
    return np.arange(m.E + 1, dtype=np.int64) * Q_, np.tile(gx, (m.E, 1)), np.tile(gw, m.E), Q_


def _closed_form_q1(h, eta0, beta_x=0.0):
    """approxS of one Q1 element with eta = eta0 (1 + beta_x xi_x), by the formulas above."""
    d = len(h)
    V = float(np.prod(h))
    detJ = V / 2 ** d
    Gmag = [V / (2 ** (d - 1) * h[i]) for i in range(d)]

    def J(k, s, deriv):
        beta = beta_x if k == 0 else 0.0
        if deriv:  # int (1 + beta xi) dxi
            return 2.0
        # int (1 + s xi)^2 (1 + beta xi) dxi = 8/3 + 4 s beta / 3
        return 8.0 / 3.0 + 4.0 * s * beta / 3.0

    total = 0.0
    for a in range(2 ** d):
        s = [1.0 if (a >> k) & 1 else -1.0 for k in range(d)]
        integ = []
        for j in range(d):
            v = eta0 * detJ * (4.0 / h[j] ** 2) / 4 ** d
            for k in range(d):
                v *= J(k, s[k], k == j)
            integ.append(v)
        for i in range(d):
            kd = sum((2.0 if j == i else 1.0) * integ[j] for j in range(d))
            total += Gmag[i] ** 2 / kd
    return total


def test_q1_approxS_closed_form_square_and_cube():
    h = 0.125
    assert abs(_closed_form_q1([h, h], 1.0) - 2 * h * h) < 1e-15
    assert abs(_closed_form_q1([h, h, h], 3.0) - 27 * h ** 3 / (8 * 3.0)) < 1e-15


def _oracle_approxS(res, hi, eta_fn):
    m = O.Mesh(res, (0.0,) * len(res), hi)
    ptr, xi, w, Qn = _gauss_slots(m)
    eta = eta_fn(m, xi, Qn)
    Ke, fe, Ge, S = O.assemble_elements(m, ptr, xi, w, eta, np.zeros((len(w), len(res))))
    return m, S


def test_q1_approxS_rectangles_nonuniform_eta():
    """2D and 3D rectangular (non-square) elements, viscosity varying from element to element,
    and in one case linearly inside each element."""
    rng = np.random.default_rng(3)
    for res, hi in (((3, 2), (0.9, 0.5)), ((3, 2, 2), (0.6, 0.5, 0.3))):
        d = len(res)
        E = int(np.prod(res))
        eta_e = 10.0 ** rng.uniform(-2, 3, E)
        h = [hi[k] / res[k] for k in range(d)]
        for beta in (0.0, 0.5):
            def eta_fn(m, xi, Qn):
                return np.repeat(eta_e, Qn) * (1.0 + beta * xi[:, 0])
            m, S = _oracle_approxS(res, hi, eta_fn)
            ref = np.array([_closed_form_q1(h, eta_e[e], beta) for e in range(E)])
            np.testing.assert_allclose(S, ref, rtol=1e-13, err_msg=f"dim {d} beta {beta}")


# ---------------------------------------------------------------------------------------
# Q2/dPc1: exact rational integration
# ---------------------------------------------------------------------------------------
def _pmul(a, b):
    out = {}
    for ea, ca in a.items():
        for eb, cb in b.items():
            e = (ea[0] + eb[0], ea[1] + eb[1])
            out[e] = out.get(e, Fr(0)) + ca * cb
    return out


def _pint(p):
    """int over [-1,1]^2 of a polynomial {(i, j): c} in (xi, eta)."""
    def m1(k):
        return Fr(0) if k % 2 else Fr(2, k + 1)
    return sum((c * m1(i) * m1(j) for (i, j), c in p.items()), Fr(0))


def _pscale(p, s):
    return {e: c * s for e, c in p.items()}


def test_q2_dpc1_approxS_block_exact():
    hx, hy = Fr(1, 2), Fr(1, 4)
    detJ = hx * hy / 4
    # 1-D quadratic Lagrange basis on {-1, 0, 1} and derivatives, as {power: coefficient}
    phi = [{2: Fr(1, 2), 1: Fr(-1, 2)}, {0: Fr(1), 2: Fr(-1)}, {2: Fr(1, 2), 1: Fr(1, 2)}]
    dphi = [{1: Fr(1), 0: Fr(-1, 2)}, {1: Fr(-2)}, {1: Fr(1), 0: Fr(1, 2)}]

    def tens(px, py):
        return {(i, j): cx * cy for i, cx in px.items() for j, cy in py.items()}

    # dPc1 basis (BilinearInnerElType.c): -xi - eta/2 + 1/4, xi - eta/2 + 1/4, eta + 1/2
    Np = [{(1, 0): Fr(-1), (0, 1): Fr(-1, 2), (0, 0): Fr(1, 4)},
          {(1, 0): Fr(1), (0, 1): Fr(-1, 2), (0, 0): Fr(1, 4)},
          {(0, 1): Fr(1), (0, 0): Fr(1, 2)}]
    G = np.zeros((18, 3), dtype=object)
    Kd = np.zeros(18, dtype=object)
    for a in range(9):
        ax, ay = a % 3, a // 3
        dNx = _pscale(tens(dphi[ax], phi[ay]), 2 / hx)  # dN_a/dx
        dNy = _pscale(tens(phi[ax], dphi[ay]), 2 / hy)  # dN_a/dy
        Ixx = _pint(_pmul(dNx, dNx)) * detJ
        Iyy = _pint(_pmul(dNy, dNy)) * detJ
        Kd[2 * a] = 2 * Ixx + Iyy
        Kd[2 * a + 1] = Ixx + 2 * Iyy
        for r in range(3):
            G[2 * a, r] = -_pint(_pmul(dNx, Np[r])) * detJ
            G[2 * a + 1, r] = -_pint(_pmul(dNy, Np[r])) * detJ
    exact = np.array([[float(sum(G[v, r] * G[v, s] / Kd[v] for v in range(18))) for s in range(3)]
                      for r in range(3)])
    mesh = Q.Q2Mesh((1, 1), (0.0, 0.0), (0.5, 0.25))
    pts, wts = Q.gauss5(2)
    Ke = Q.element_K(mesh, pts, (wts * mesh.detJ)[None, :])
    Ge = Q.element_G(mesh)
    S = Q.element_approxS(Ke, Ge)[0]
    np.testing.assert_allclose(S, exact, rtol=1e-13, atol=1e-15 * np.abs(exact).max())
    # the block is symmetric positive definite
    assert np.all(np.linalg.eigvalsh(exact) > 0)
