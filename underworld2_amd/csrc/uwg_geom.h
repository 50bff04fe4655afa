// Geometry primitives of the regular Q1 (and Q2) mesh used by the bit-exact swarm kernels.
// Every routine keeps the reference's IEEE operation order; the translation unit that
// includes this header is compiled with -ffp-contract=off so no multiply-add is fused.
// Citations: LU/ = src/underworld/libUnderworld/.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace uwg {

#define UWG_HD __host__ __device__ __forceinline__

// CartesianGenerator_CalcGeom, LU/StgDomain/Mesh/src/CartesianGenerator.c:2001-2039
UWG_HD double node_coord_1d(int i, int nnodes, double lo, double hi) {
  if (i <= 0) return lo;
  if (i >= nnodes - 1) return hi;
  double steps = hi - lo;
  return (lo + 0.0) + ((double)i / (double)(nnodes - 1)) * (steps - 0.0 * 0.0);
}

// Element (i, j, k) of cell index e (x fastest).  Cell indices are below 2^31 (owners are
// int), so the decomposition uses 32-bit unsigned division.
template <int DIM>
UWG_HD void element_ijk(const int* res, long e, int* ijk) {
  const unsigned u = (unsigned)e, rx = (unsigned)res[0], ry = (unsigned)res[1];
  const unsigned q = u / rx;
  ijk[0] = (int)(u - q * rx);
  ijk[1] = (int)(q % ry);
  ijk[2] = DIM == 3 ? (int)(q / ry) : 0;
}

// Element-local vertex coordinates of element (i, j, k), local order x fastest
// (TrilinearElementType.c:216-233)
template <int DIM>
UWG_HD void element_verts_ijk(const int* res, const double* lo, const double* hi, const int* ijk, double* v) {
  const int i = ijk[0], j = ijk[1], k = ijk[2];
  double X[2] = {node_coord_1d((int)i, res[0] + 1, lo[0], hi[0]), node_coord_1d((int)i + 1, res[0] + 1, lo[0], hi[0])};
  double Y[2] = {node_coord_1d((int)j, res[1] + 1, lo[1], hi[1]), node_coord_1d((int)j + 1, res[1] + 1, lo[1], hi[1])};
  if (DIM == 2) {
    for (int a = 0; a < 4; a++) {
      v[a * 2 + 0] = X[a & 1];
      v[a * 2 + 1] = Y[(a >> 1) & 1];
    }
  } else {
    double Z[2] = {node_coord_1d((int)k, res[2] + 1, lo[2], hi[2]), node_coord_1d((int)k + 1, res[2] + 1, lo[2], hi[2])};
    for (int a = 0; a < 8; a++) {
      v[a * 3 + 0] = X[a & 1];
      v[a * 3 + 1] = Y[(a >> 1) & 1];
      v[a * 3 + 2] = Z[(a >> 2) & 1];
    }
  }
}

template <int DIM>
UWG_HD void element_verts(const int* res, const double* lo, const double* hi, long e, double* v) {
  int ijk[3];
  element_ijk<DIM>(res, e, ijk);
  element_verts_ijk<DIM>(res, lo, hi, ijk, v);
}

template <int DIM>
UWG_HD long element_base_node_ijk(const int* res, const int* ijk) {
  const long nx = res[0] + 1, ny = res[1] + 1;
  return ijk[0] + nx * (ijk[1] + ny * (long)ijk[2]);
}

template <int DIM>
UWG_HD long element_base_node(const int* res, long e) {
  int ijk[3];
  element_ijk<DIM>(res, e, ijk);
  return element_base_node_ijk<DIM>(res, ijk);
}

// Shape functions exactly as written in Bilinear/TrilinearElementType.c
template <int DIM>
UWG_HD void shape(const double* xi, double* N) {
  if (DIM == 2) {
    double x = xi[0], y = xi[1];
    N[0] = 0.25 * (1.0 - x) * (1.0 - y);
    N[1] = 0.25 * (1.0 + x) * (1.0 - y);
    N[2] = 0.25 * (1.0 - x) * (1.0 + y);
    N[3] = 0.25 * (1.0 + x) * (1.0 + y);
  } else {
    double x = xi[0], y = xi[1], z = xi[2];
    N[0] = 0.125 * (1.0 - x) * (1.0 - y) * (1.0 - z);
    N[2] = 0.125 * (1.0 - x) * (1.0 + y) * (1.0 - z);
    N[3] = 0.125 * (1.0 + x) * (1.0 + y) * (1.0 - z);
    N[1] = 0.125 * (1.0 + x) * (1.0 - y) * (1.0 - z);
    N[4] = 0.125 * (1.0 - x) * (1.0 - y) * (1.0 + z);
    N[6] = 0.125 * (1.0 - x) * (1.0 + y) * (1.0 + z);
    N[7] = 0.125 * (1.0 + x) * (1.0 + y) * (1.0 + z);
    N[5] = 0.125 * (1.0 + x) * (1.0 - y) * (1.0 + z);
  }
}

template <int DIM>
UWG_HD void shape_local_derivs(const double* xi, double (*G)[8]) {
  if (DIM == 2) {
    double x = xi[0], y = xi[1];
    G[0][0] = -0.25 * (1.0 - y);
    G[0][1] = 0.25 * (1.0 - y);
    G[0][3] = 0.25 * (1.0 + y);
    G[0][2] = -0.25 * (1.0 + y);
    G[1][0] = -0.25 * (1.0 - x);
    G[1][1] = -0.25 * (1.0 + x);
    G[1][3] = 0.25 * (1.0 + x);
    G[1][2] = 0.25 * (1.0 - x);
  } else {
    double x = xi[0], y = xi[1], z = xi[2];
    G[0][0] = -0.125 * (1.0 - y) * (1.0 - z);
    G[0][2] = -0.125 * (1.0 + y) * (1.0 - z);
    G[0][3] = 0.125 * (1.0 + y) * (1.0 - z);
    G[0][1] = 0.125 * (1.0 - y) * (1.0 - z);
    G[0][4] = -0.125 * (1.0 - y) * (1.0 + z);
    G[0][6] = -0.125 * (1.0 + y) * (1.0 + z);
    G[0][7] = 0.125 * (1.0 + y) * (1.0 + z);
    G[0][5] = 0.125 * (1.0 - y) * (1.0 + z);
    G[1][0] = -0.125 * (1.0 - x) * (1.0 - z);
    G[1][2] = 0.125 * (1.0 - x) * (1.0 - z);
    G[1][3] = 0.125 * (1.0 + x) * (1.0 - z);
    G[1][1] = -0.125 * (1.0 + x) * (1.0 - z);
    G[1][4] = -0.125 * (1.0 - x) * (1.0 + z);
    G[1][6] = 0.125 * (1.0 - x) * (1.0 + z);
    G[1][7] = 0.125 * (1.0 + x) * (1.0 + z);
    G[1][5] = -0.125 * (1.0 + x) * (1.0 + z);
    G[2][0] = -0.125 * (1.0 - x) * (1.0 - y);
    G[2][2] = -0.125 * (1.0 - x) * (1.0 + y);
    G[2][3] = -0.125 * (1.0 + x) * (1.0 + y);
    G[2][1] = -0.125 * (1.0 + x) * (1.0 - y);
    G[2][4] = 0.125 * (1.0 - x) * (1.0 - y);
    G[2][6] = 0.125 * (1.0 - x) * (1.0 + y);
    G[2][7] = 0.125 * (1.0 + x) * (1.0 + y);
    G[2][5] = 0.125 * (1.0 + x) * (1.0 - y);
  }
}

// Cramer's rule, LU/StgDomain/Geometry/src/TensorMath.c:959-1020 (row-major J)
template <int DIM>
UWG_HD void cramer(const double* J, double* s, const double* r) {
  if (DIM == 3) {
    double det = +J[0] * (J[4] * J[8] - J[7] * J[5]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                 J[2] * (J[3] * J[7] - J[4] * J[6]);
    s[0] = (r[0] * (J[4] * J[8] - J[7] * J[5]) - J[1] * (r[1] * J[8] - J[5] * r[2]) +
            J[2] * (r[1] * J[7] - J[4] * r[2])) / det;
    s[1] = (J[0] * (r[1] * J[8] - r[2] * J[5]) - r[0] * (J[3] * J[8] - J[5] * J[6]) +
            J[2] * (J[3] * r[2] - r[1] * J[6])) / det;
    s[2] = (J[0] * (J[4] * r[2] - J[7] * r[1]) - J[1] * (J[3] * r[2] - r[1] * J[6]) +
            r[0] * (J[3] * J[7] - J[4] * J[6])) / det;
  } else {
    double det = J[0] * J[3] - J[2] * J[1];
    s[0] = (r[0] * J[3] - J[1] * r[1]) / det;
    s[1] = (r[1] * J[0] - J[2] * r[0]) / det;
  }
}

// Newton global->local, LU/StgFEM/Discretisation/src/ElementType.c:177-301
template <int DIM>
UWG_HD void newton_local(const double* v, const double* x, double* xi) {
  constexpr int NN = DIM == 3 ? 8 : 4;
  for (int d = 0; d < DIM; d++) xi[d] = 0.0;
  for (int it = 0; it < 100; it++) {
    double J[9], rhs[3], inc[3], N[8], G[3][8];
    for (int q = 0; q < 9; q++) J[q] = 0.0;
    rhs[0] = rhs[1] = rhs[2] = 0.0;
    shape<DIM>(xi, N);
    shape_local_derivs<DIM>(xi, G);
    for (int n = 0; n < NN; n++) {
      const double* c = v + n * DIM;
      double sf = N[n];
      J[0 * DIM + 0] += G[0][n] * c[0];
      J[0 * DIM + 1] += G[1][n] * c[0];
      J[1 * DIM + 0] += G[0][n] * c[1];
      J[1 * DIM + 1] += G[1][n] * c[1];
      rhs[0] -= sf * c[0];
      rhs[1] -= sf * c[1];
      if (DIM == 3) {
        J[2] += G[2][n] * c[0];
        J[5] += G[2][n] * c[1];
        J[6] += G[0][n] * c[2];
        J[7] += G[1][n] * c[2];
        J[8] += G[2][n] * c[2];
        rhs[2] -= sf * c[2];
      }
    }
    rhs[0] += x[0];
    rhs[1] += x[1];
    if (DIM == 3) rhs[2] += x[2];
    cramer<DIM>(J, inc, rhs);
    xi[0] += inc[0];
    xi[1] += inc[1];
    if (DIM == 3) xi[2] += inc[2];
    double m = fabs(inc[0]);
    if (m < fabs(inc[1])) m = fabs(inc[1]);
    if (DIM == 3 && m < fabs(inc[2])) m = fabs(inc[2]);
    if (m < 0.0001) return;
  }
  xi[0] = 1.1;
}

// ---------------------------------------------------------------------------------------
// Q2 (Biquadratic / Triquadratic) element geometry for lookups in a quadratic velocity field.
// Element-local vertex a = ax + 3 ay (+ 9 az) on the element's 3^d sub-grid of the vertex grid
// (2 res + 1 vertices per direction), the C2Generator incidence order.
// ---------------------------------------------------------------------------------------
template <int DIM>
UWG_HD void q2_element_verts_ijk(const int* res, const double* lo, const double* hi, const int* ijk, double* v) {
  const long i = ijk[0], j = ijk[1], k = ijk[2];
  double X[3], Y[3], Z[3];
  for (int t = 0; t < 3; t++) {
    X[t] = node_coord_1d((int)(2 * i + t), 2 * res[0] + 1, lo[0], hi[0]);
    Y[t] = node_coord_1d((int)(2 * j + t), 2 * res[1] + 1, lo[1], hi[1]);
    Z[t] = DIM == 3 ? node_coord_1d((int)(2 * k + t), 2 * res[2] + 1, lo[2], hi[2]) : 0.0;
  }
  for (int a = 0; a < (DIM == 3 ? 27 : 9); a++) {
    v[a * DIM + 0] = X[a % 3];
    v[a * DIM + 1] = Y[(a / 3) % 3];
    if (DIM == 3) v[a * DIM + 2] = Z[a / 9];
  }
}

template <int DIM>
UWG_HD void q2_element_verts(const int* res, const double* lo, const double* hi, long e, double* v) {
  int ijk[3];
  element_ijk<DIM>(res, e, ijk);
  q2_element_verts_ijk<DIM>(res, lo, hi, ijk, v);
}

// Shape functions with the factorisation and evaluation order of Biquadratic_EvalBasis
// (LU/StgFEM/Discretisation/src/Biquadratic.c:138-158) and Triquadratic_EvalBasis
// (Triquadratic.c:193-243): N_a = m * f_x(ax) * f_y(ay) [* f_z(az)], the 1D factors
// (t-1, 1-t^2, t+1) and a prefactor m = (0.5^k) times the coordinates of the non-middle
// directions (k of them), absent when every direction is middle.
template <int DIM>
UWG_HD void q2_basis(const double* xi, double* N) {
  if (DIM == 2) {
    const double x = xi[0], y = xi[1];
    const double A[3] = {x - 1.0, 1.0 - x * x, x + 1.0}, B[3] = {y - 1.0, 1.0 - y * y, y + 1.0};
    const double m0 = 0.5 * x, m1 = 0.5 * y, m2 = 0.25 * x * y;
    for (int a = 0; a < 9; a++) {
      const int ax = a % 3, ay = a / 3;
      const bool cx = ax == 1, cy = ay == 1;
      if (cx && cy)
        N[a] = A[ax] * B[ay];
      else
        N[a] = (cx ? m1 : (cy ? m0 : m2)) * A[ax] * B[ay];
    }
  } else {
    const double x = xi[0], y = xi[1], z = xi[2];
    const double A[3] = {x - 1.0, 1.0 - x * x, x + 1.0}, B[3] = {y - 1.0, 1.0 - y * y, y + 1.0},
                 C[3] = {z - 1.0, 1.0 - z * z, z + 1.0};
    const double m0 = 0.5 * x, m1 = 0.5 * y, m2 = 0.5 * z, m3 = 0.25 * x * y, m4 = 0.25 * x * z, m5 = 0.25 * y * z,
                 m6 = 0.125 * x * y * z;
    for (int a = 0; a < 27; a++) {
      const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
      const int mid = (ax == 1) | ((ay == 1) << 1) | ((az == 1) << 2);
      // prefactor by the set of middle directions: {} m6, {x} m5, {y} m4, {z} m3, {x,y} m2,
      // {x,z} m1, {y,z} m0, {x,y,z} none
      const double M[8] = {m6, m5, m4, m2, m3, m1, m0, 1.0};
      if (mid == 7)
        N[a] = A[ax] * B[ay] * C[az];
      else
        N[a] = M[mid] * A[ax] * B[ay] * C[az];
    }
  }
}

// Local derivatives G[d][a] = dN_a/dxi_d with the factorisation and order of
// Biquadratic_EvalLocalDerivs (Biquadratic.c:160-192) and Triquadratic_EvalLocalDerivs
// (Triquadratic.c:245-374).
template <int DIM>
UWG_HD void q2_derivs(const double* xi, double (*G)[27]) {
  if (DIM == 2) {
    const double x = xi[0], y = xi[1];
    const double a0 = x - 1.0, b0 = y - 1.0, a1 = x + 1.0, b1 = y + 1.0;
    const double a2 = 2.0 * x - 1.0, b2 = 2.0 * y - 1.0, a3 = 2.0 * x + 1.0, b3 = 2.0 * y + 1.0;
    const double a4 = 1.0 - x * x, b4 = 1.0 - y * y;
    const double m0 = 0.25 * x, m1 = 0.25 * y, m2 = -x * y;
    // corners
    G[0][0] = m1 * a2 * b0;
    G[0][2] = m1 * a3 * b0;
    G[0][6] = m1 * a2 * b1;
    G[0][8] = m1 * a3 * b1;
    G[1][0] = m0 * a0 * b2;
    G[1][2] = m0 * a1 * b2;
    G[1][6] = m0 * a0 * b3;
    G[1][8] = m0 * a1 * b3;
    // sides
    G[0][1] = m2 * b0;
    G[0][7] = m2 * b1;
    G[0][3] = 0.5 * a2 * b4;
    G[0][5] = 0.5 * a3 * b4;
    G[1][1] = 0.5 * a4 * b2;
    G[1][7] = 0.5 * a4 * b3;
    G[1][3] = m2 * a0;
    G[1][5] = m2 * a1;
    // centre
    G[0][4] = -2.0 * x * b4;
    G[1][4] = -2.0 * y * a4;
  } else {
    const double x = xi[0], y = xi[1], z = xi[2];
    const double A[3] = {x - 1.0, 1.0 - x * x, x + 1.0}, B[3] = {y - 1.0, 1.0 - y * y, y + 1.0},
                 C[3] = {z - 1.0, 1.0 - z * z, z + 1.0};
    const double DA[3] = {x - 0.5, -2.0 * x, x + 0.5}, DB[3] = {y - 0.5, -2.0 * y, y + 0.5},
                 DC[3] = {z - 0.5, -2.0 * z, z + 0.5};
    const double m0 = 0.5 * x, m1 = 0.5 * y, m2 = 0.5 * z, m3 = 0.25 * x * y, m4 = 0.25 * x * z, m5 = 0.25 * y * z;
    for (int a = 0; a < 27; a++) {
      const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
      const bool cx = ax == 1, cy = ay == 1, cz = az == 1;
      // d/dxi: prefactor from (y, z): neither middle m5, y middle m2, z middle m1, both none
      if (cy && cz)
        G[0][a] = DA[ax] * B[ay] * C[az];
      else
        G[0][a] = DA[ax] * (cy ? m2 : (cz ? m1 : m5)) * B[ay] * C[az];
      // d/deta: prefactor from (x, z): neither m4, x middle m2, z middle m0
      if (cx && cz)
        G[1][a] = DB[ay] * A[ax] * C[az];
      else
        G[1][a] = DB[ay] * (cx ? m2 : (cz ? m0 : m4)) * A[ax] * C[az];
      // d/dzeta: prefactor from (x, y): neither m3, x middle m1, y middle m0
      if (cx && cy)
        G[2][a] = DC[az] * A[ax] * B[ay];
      else
        G[2][a] = DC[az] * (cx ? m1 : (cy ? m0 : m3)) * A[ax] * B[ay];
    }
  }
}

// Newton global->local (ElementType.c:177-301) on a Q2 element: the same iteration as
// newton_local with the 9 / 27 quadratic shape functions, nodes summed in incidence order.
template <int DIM>
UWG_HD void q2_newton_local(const double* v, const double* x, double* xi) {
  constexpr int NN = DIM == 3 ? 27 : 9;
  for (int d = 0; d < DIM; d++) xi[d] = 0.0;
  for (int it = 0; it < 100; it++) {
    double J[9], rhs[3], inc[3], N[27], G[3][27];
    for (int q = 0; q < 9; q++) J[q] = 0.0;
    rhs[0] = rhs[1] = rhs[2] = 0.0;
    q2_basis<DIM>(xi, N);
    q2_derivs<DIM>(xi, G);
    for (int n = 0; n < NN; n++) {
      const double* c = v + n * DIM;
      const double sf = N[n];
      J[0 * DIM + 0] += G[0][n] * c[0];
      J[0 * DIM + 1] += G[1][n] * c[0];
      J[1 * DIM + 0] += G[0][n] * c[1];
      J[1 * DIM + 1] += G[1][n] * c[1];
      rhs[0] -= sf * c[0];
      rhs[1] -= sf * c[1];
      if (DIM == 3) {
        J[2] += G[2][n] * c[0];
        J[5] += G[2][n] * c[1];
        J[6] += G[0][n] * c[2];
        J[7] += G[1][n] * c[2];
        J[8] += G[2][n] * c[2];
        rhs[2] -= sf * c[2];
      }
    }
    rhs[0] += x[0];
    rhs[1] += x[1];
    if (DIM == 3) rhs[2] += x[2];
    cramer<DIM>(J, inc, rhs);
    xi[0] += inc[0];
    xi[1] += inc[1];
    if (DIM == 3) xi[2] += inc[2];
    double m = fabs(inc[0]);
    if (m < fabs(inc[1])) m = fabs(inc[1]);
    if (DIM == 3 && m < fabs(inc[2])) m = fabs(inc[2]);
    if (m < 0.0001) return;
  }
  xi[0] = 1.1;
}

// Num_Approx, LU/StGermain/Base/Foundation/src/Numerics.c:15-21
UWG_HD bool num_approx(double var, double val) {
  return var >= val - 1e-12 && var <= val + 1e-12;
}

// Mesh_RegularAlgorithms_SearchElements, LU/StgDomain/Mesh/src/Mesh_RegularAlgorithms.c:174-206
// sep = (max-min)/nel as computed once in Mesh_RegularAlgorithms_Update (:153-154).
// The final clamp only guards memory; the reference has no element past nel-1.
// The element's (i, j, k) instead of its index: search_element_ijk.  out >= 0 here, so the
// fractional part modf would return is out - trunc(out), exactly.
template <int DIM>
UWG_HD bool search_element_ijk(const int* res, const double* lo, const double* hi, const double* x, int* ind) {
  ind[0] = ind[1] = ind[2] = 0;
  for (int d = 0; d < DIM; d++) {
    double sep = (hi[d] - lo[d]) / res[d];
    if (num_approx(x[d] - hi[d], 0.0)) {
      ind[d] = res[d] - 1;
    } else if (x[d] < lo[d] || x[d] > hi[d]) {
      return false;
    } else {
      const double out = (x[d] - lo[d]) / sep;
      const double integer = trunc(out);
      const double frac = out - integer;
      long id = (long)(unsigned)integer;
      if (id > 0 && num_approx(frac, 0.0)) id--;
      if (id > res[d] - 1) id = res[d] - 1;
      ind[d] = (int)id;
    }
  }
  return true;
}

template <int DIM>
UWG_HD bool search_element(const int* res, const double* lo, const double* hi, const double* x, long* el) {
  int ind[3];
  if (!search_element_ijk<DIM>(res, lo, hi, x, ind)) return false;
  *el = ind[0] + (long)res[0] * (ind[1] + (long)res[1] * ind[2]);
  return true;
}

// _ElementCellLayout_CellOf, LU/StgDomain/Swarm/src/ElementCellLayout.c:291-320
template <int DIM>
UWG_HD long cell_of(const int* res, const double* lo, const double* hi, const double* x, long E) {
  for (int d = 0; d < DIM; d++)
    if (x[d] < lo[d] || x[d] > hi[d]) return E;
  long el;
  if (!search_element<DIM>(res, lo, hi, x, &el)) return E;
  return el;
}

}  // namespace uwg
