// fn_kernels.hip -- per-particle evaluation of the viscosity / body-force function graph.
//
// The Python layer compiles an underworld.function graph (PY/function/_function.py operators,
// PY/function/branching.py map/conditional, PY/function/math.py; C++ closures in
// LU/Underworld/Function/src: Binary.cpp, Relational.cpp, Map.cpp, Conditional.cpp,
// FeVariableFn.cpp, SwarmVariableFn.cpp) into a short register program.  One thread per
// particle runs it: inputs are the particle's coordinate, nodal mesh fields interpolated at
// the particle (search + Newton + shape functions, FeVariable.c:389-403 / 1007-1037), swarm
// variables uploaded to the device, and the material index.  Values are vectors of up to 3
// components with numpy broadcasting of single components, as in the host evaluator.
// Compiled with -ffp-contract=off: the same IEEE operations as the host evaluator.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "uwg_geom.h"
#include "uwg_internal.h"

namespace uwg {

namespace {
struct FnGeom {
  int res[3];
  double lo[3], hi[3];
};
FnGeom fn_pack(const Geom& g) {
  FnGeom a;
  for (int d = 0; d < 3; d++) {
    a.res[d] = g.res[d];
    a.lo[d] = g.lo[d];
    a.hi[d] = g.hi[d];
  }
  return a;
}
inline unsigned fn_blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
}  // namespace

enum FnOp {
  FN_CONST = 0,   // dst = consts[a .. a+n)
  FN_COORD = 1,   // dst = x_p
  FN_FIELD = 2,   // dst = mesh field a interpolated at x_p (n components)
  FN_SWVAR = 3,   // dst = swarm variable a at p (n components)
  FN_MAT = 4,     // dst = material index of p
  FN_ADD = 10, FN_SUB = 11, FN_MUL = 12, FN_DIV = 13, FN_POW = 14,
  FN_LT = 15, FN_LE = 16, FN_GT = 17, FN_GE = 18, FN_AND = 19, FN_OR = 20,
  FN_NEG = 30, FN_EXP = 31, FN_LOG = 32, FN_SQRT = 33, FN_ABS = 34, FN_SIN = 35, FN_COS = 36, FN_NOT = 37,
  FN_DOT = 40,    // dst = sum_c a_c b_c
  FN_INDEX = 41,  // dst = a[b]
  FN_APPEND = 42, // dst = (a, b) concatenated
  FN_EQK = 51,    // dst = (long)a[0] == (long)consts[b]
  FN_SELECT = 52, // if (b[0] != 0) dst = a      (dst keeps its value otherwise)
  FN_CHECK = 53,  // error unless a[0] != 0     (map key missing / no true clause)
  FN_MOVE = 54,   // dst = a
};

struct FnIns {
  int op, dst, a, b, n;
};

constexpr int FN_MAXREG = 48;

template <int DIM>
__device__ bool field_at(const FnGeom& g, const double* f, int nc, const double* x, double* v) {
  int ijk[3];
  if (!search_element_ijk<DIM>(g.res, g.lo, g.hi, x, ijk)) return false;
  double verts[24], xi[3], N[8];
  element_verts_ijk<DIM>(g.res, g.lo, g.hi, ijk, verts);
  newton_local<DIM>(verts, x, xi);
  shape<DIM>(xi, N);
  long b = element_base_node_ijk<DIM>(g.res, ijk);
  long nx = g.res[0] + 1, nxy = nx * (g.res[1] + 1);
  for (int c = 0; c < nc; c++) v[c] = 0.0;
  for (int a = 0; a < (DIM == 3 ? 8 : 4); a++) {
    long n = b + (a & 1) + ((a >> 1) & 1) * nx + (DIM == 3 ? ((a >> 2) & 1) * nxy : 0);
    for (int c = 0; c < nc; c++) v[c] += f[n * nc + c] * N[a];
  }
  return true;
}

struct FnFieldTab {
  const double* d[8];
  int nc[8];
};
struct FnVarTab {
  const double* d[8];
  int nc[8];
};

template <int DIM>
__global__ void k_fn_eval(FnGeom g, long P, const double* __restrict__ coords, const int* __restrict__ mat,
                          const FnIns* __restrict__ prog, int nins, const double* __restrict__ consts, FnFieldTab F,
                          FnVarTab V, int out_eta, int out_force, double* __restrict__ eta, double* __restrict__ force,
                          int* __restrict__ err) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double R[FN_MAXREG][3];
  int RN[FN_MAXREG];
  double x[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < DIM; d++) x[d] = coords[p * DIM + d];
  for (int i = 0; i < nins; i++) {
    const FnIns I = prog[i];
    double* D = R[I.dst];
    const double* A = R[I.a < FN_MAXREG && I.a >= 0 ? I.a : 0];
    const double* B = R[I.b < FN_MAXREG && I.b >= 0 ? I.b : 0];
    const int na = (I.a >= 0 && I.a < FN_MAXREG) ? RN[I.a] : 1;
    const int nb = (I.b >= 0 && I.b < FN_MAXREG) ? RN[I.b] : 1;
    switch (I.op) {
      case FN_CONST:
        for (int c = 0; c < I.n; c++) D[c] = consts[I.a + c];
        break;
      case FN_COORD:
        for (int c = 0; c < I.n; c++) D[c] = x[c];
        break;
      case FN_FIELD:
        if (!field_at<DIM>(g, F.d[I.a], F.nc[I.a], x, D)) err[0] = 1;
        break;
      case FN_SWVAR:
        for (int c = 0; c < I.n; c++) D[c] = V.d[I.a][p * V.nc[I.a] + c];
        break;
      case FN_MAT:
        D[0] = (double)mat[p];
        break;
      case FN_ADD: case FN_SUB: case FN_MUL: case FN_DIV: case FN_POW:
      case FN_LT: case FN_LE: case FN_GT: case FN_GE: case FN_AND: case FN_OR:
        for (int c = 0; c < I.n; c++) {
          const double u = A[na == 1 ? 0 : c], w = B[nb == 1 ? 0 : c];
          double r;
          switch (I.op) {
            case FN_ADD: r = u + w; break;
            case FN_SUB: r = u - w; break;
            case FN_MUL: r = u * w; break;
            case FN_DIV: r = u / w; break;
            case FN_POW: r = pow(u, w); break;
            case FN_LT: r = u < w ? 1.0 : 0.0; break;
            case FN_LE: r = u <= w ? 1.0 : 0.0; break;
            case FN_GT: r = u > w ? 1.0 : 0.0; break;
            case FN_GE: r = u >= w ? 1.0 : 0.0; break;
            case FN_AND: r = (u != 0.0 && w != 0.0) ? 1.0 : 0.0; break;
            default: r = (u != 0.0 || w != 0.0) ? 1.0 : 0.0; break;
          }
          D[c] = r;
        }
        break;
      case FN_NEG: case FN_EXP: case FN_LOG: case FN_SQRT: case FN_ABS: case FN_SIN: case FN_COS: case FN_NOT:
        for (int c = 0; c < I.n; c++) {
          const double u = A[c];
          double r;
          switch (I.op) {
            case FN_NEG: r = -u; break;
            case FN_EXP: r = exp(u); break;
            case FN_LOG: r = log(u); break;
            case FN_SQRT: r = sqrt(u); break;
            case FN_ABS: r = fabs(u); break;
            case FN_SIN: r = sin(u); break;
            case FN_COS: r = cos(u); break;
            default: r = u != 0.0 ? 0.0 : 1.0; break;
          }
          D[c] = r;
        }
        break;
      case FN_DOT: {
        double s = 0.0;
        for (int c = 0; c < na; c++) s += A[c] * B[nb == 1 ? 0 : c];
        D[0] = s;
        break;
      }
      case FN_INDEX:
        D[0] = A[I.b];
        break;
      case FN_APPEND: {
        double t[3];
        for (int c = 0; c < na; c++) t[c] = A[c];
        for (int c = 0; c < nb && na + c < 3; c++) t[na + c] = B[c];
        for (int c = 0; c < I.n; c++) D[c] = t[c];
        break;
      }
      case FN_EQK:
        D[0] = ((long long)A[0] == (long long)consts[I.b]) ? 1.0 : 0.0;
        break;
      case FN_SELECT:
        if (B[0] != 0.0)
          for (int c = 0; c < I.n; c++) D[c] = A[c];
        break;
      case FN_CHECK:
        if (A[0] == 0.0) err[0] = 2;
        break;
      case FN_MOVE:
        for (int c = 0; c < I.n; c++) D[c] = A[c];
        break;
      default:
        err[0] = 3;
        break;
    }
    if (I.op != FN_CHECK) RN[I.dst] = I.n;
  }
  if (out_eta >= 0) {
    const double e = R[out_eta][0];
    eta[p] = e;
    if (!(e > 0.0)) err[1] = 1;
  }
  if (out_force >= 0)
    for (int d = 0; d < DIM; d++) force[p * DIM + d] = R[out_force][RN[out_force] == 1 ? 0 : d];
}

void launch_fn_eval(const Geom& g, long P, const double* coords, const int* mat, const void* prog, int nins,
                    const double* consts, const double* const* fields, const int* field_nc, int nfields,
                    const double* const* vars, const int* var_nc, int nvars, int out_eta, int out_force, double* eta,
                    double* force, int* err, hipStream_t s) {
  if (P == 0) return;
  FnFieldTab F{};
  FnVarTab V{};
  for (int i = 0; i < nfields && i < 8; i++) {
    F.d[i] = fields[i];
    F.nc[i] = field_nc[i];
  }
  for (int i = 0; i < nvars && i < 8; i++) {
    V.d[i] = vars[i];
    V.nc[i] = var_nc[i];
  }
  if (g.dim == 3)
    hipLaunchKernelGGL(k_fn_eval<3>, fn_blocks(P, 128), 128, 0, s, fn_pack(g), P, coords, mat, (const FnIns*)prog, nins,
                       consts, F, V, out_eta, out_force, eta, force, err);
  else
    hipLaunchKernelGGL(k_fn_eval<2>, fn_blocks(P, 128), 128, 0, s, fn_pack(g), P, coords, mat, (const FnIns*)prog, nins,
                       consts, F, V, out_eta, out_force, eta, force, err);
  UWG_HIP(hipGetLastError());
}

// interpolation of any nodal field (n components) at points: FeVariable_InterpolateValueAt
template <int DIM>
__global__ void k_field_interp(FnGeom g, long P, const double* f, int nc, const double* x, double* v, int* ok) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double xx[3], vv[9];
  for (int d = 0; d < DIM; d++) xx[d] = x[p * DIM + d];
  bool good = field_at<DIM>(g, f, nc, xx, vv);
  ok[p] = good ? 1 : 0;
  for (int c = 0; c < nc; c++) v[p * nc + c] = good ? vv[c] : 0.0;
}

void launch_field_interpolate(const Geom& g, long P, const double* f, int nc, const double* x, double* v, int* ok,
                              hipStream_t s) {
  if (P == 0) return;
  if (g.dim == 3)
    hipLaunchKernelGGL(k_field_interp<3>, fn_blocks(P, 256), 256, 0, s, fn_pack(g), P, f, nc, x, v, ok);
  else
    hipLaunchKernelGGL(k_field_interp<2>, fn_blocks(P, 256), 256, 0, s, fn_pack(g), P, f, nc, x, v, ok);
  UWG_HIP(hipGetLastError());
}

}  // namespace uwg
