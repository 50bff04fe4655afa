// Finite-element kernels of the Stokes solve: particle -> element quadrature moments,
// Galerkin coarse moments, matrix-free K / G / G^T applies, diagonal, transfer
// operators and vector kernels.
//
// Quadrature representation.  On an affine Q1 element every entry of
//   K_e = sum_p eta_p w_p detJ  B(xi_p)^T D0 B(xi_p),  D0 = diag(2,2,2,1,1,1)
// (_ConstitutiveMatrixCartesian_AssembleElement, LU/Underworld/Rheology/src/
// ConstitutiveMatrixCartesian.cpp:238-429) is a polynomial of degree <= 2 per axis in xi,
// so K_e depends on the particles only through the 3^d moments
//   M_k = sum_p eta_p w_p detJ xi^kx eta^ky zeta^kz .
// We store, per element, the 3^d weights W_g of the Gauss points that reproduce these
// moments exactly (W = (V^-1 (x) V^-1 (x) V^-1) M with V the 3x3 Vandermonde of the
// Gauss abscissae).  K_e = sum_g W_g B(g)^T D0 B(g) then holds to roundoff, for any
// particle distribution (W may be negative).  The same identity gives the Galerkin coarse
// operator P^T K P of trilinear prolongation exactly: the children's Gauss points act as
// the coarse element's quadrature points.
#include "uwg_internal.h"

namespace uwg {

static inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

#define GA 0.77459666924148337703585307995647992216658434105832  // sqrt(3/5)
#define GA2 0.6  // GA^2

// 1D: values of the Gauss abscissae
__host__ __device__ constexpr double gpt(int i) { return i == 0 ? -GA : (i == 1 ? 0.0 : GA); }

// 1D inverse Vandermonde: (M0, M1, M2) -> (W0, W1, W2)
__host__ __device__ inline void vinv1(double m0, double m1, double m2, double* w) {
  double a = m2 / GA2, b = m1 / GA;
  w[0] = 0.5 * (a - b);
  w[1] = m0 - a;
  w[2] = 0.5 * (a + b);
}

// tensor transform of moments M[kz][ky][kx] (27 or 9) into Gauss weights, in place
template <int DIM>
__device__ inline void moments_to_weights(double* M) {
  constexpr int Q1 = 3;
  // along x
  for (int r = 0; r < (DIM == 3 ? 9 : 3); r++) {
    double w[3];
    vinv1(M[r * 3 + 0], M[r * 3 + 1], M[r * 3 + 2], w);
    M[r * 3 + 0] = w[0]; M[r * 3 + 1] = w[1]; M[r * 3 + 2] = w[2];
  }
  // along y
  for (int kz = 0; kz < (DIM == 3 ? 3 : 1); kz++)
    for (int kx = 0; kx < Q1; kx++) {
      double w[3];
      int b = kz * 9 + kx;
      vinv1(M[b], M[b + 3], M[b + 6], w);
      M[b] = w[0]; M[b + 3] = w[1]; M[b + 6] = w[2];
    }
  if (DIM == 3)
    for (int r = 0; r < 9; r++) {
      double w[3];
      vinv1(M[r], M[r + 9], M[r + 18], w);
      M[r] = w[0]; M[r + 9] = w[1]; M[r + 18] = w[2];
    }
}

// derivative dN_a/dxi_j of the Q1 shape function of local node a at reference point X
template <int DIM>
__host__ __device__ inline void dN_ref(int a, const double* X, double* d) {
  double s[3] = {(a & 1) ? 1.0 : -1.0, (a & 2) ? 1.0 : -1.0, (a & 4) ? 1.0 : -1.0};
  if (DIM == 2) {
    d[0] = 0.25 * s[0] * (1.0 + s[1] * X[1]);
    d[1] = 0.25 * s[1] * (1.0 + s[0] * X[0]);
  } else {
    d[0] = 0.125 * s[0] * (1.0 + s[1] * X[1]) * (1.0 + s[2] * X[2]);
    d[1] = 0.125 * s[1] * (1.0 + s[0] * X[0]) * (1.0 + s[2] * X[2]);
    d[2] = 0.125 * s[2] * (1.0 + s[0] * X[0]) * (1.0 + s[1] * X[1]);
  }
}

// element stiffness diagonal K[(a,i),(a,i)] = sum_g W_g ((dN_a/dx_i)^2 + sum_j (dN_a/dx_j)^2)
template <int DIM>
__device__ inline double kdiag_entry(const double* Wg, int a, int i, const double* sc) {
  double acc = 0.0;
  for (int g = 0; g < (DIM == 3 ? 27 : 9); g++) {
    double X[3] = {gpt(g % 3), gpt((g / 3) % 3), gpt(g / 9)};
    double d[3];
    dN_ref<DIM>(a, X, d);
    double s2 = 0.0;
    for (int j = 0; j < DIM; j++) { d[j] *= sc[j]; s2 += d[j] * d[j]; }
    acc += Wg[g] * (d[i] * d[i] + s2);
  }
  return acc;
}

// Element stiffness diagonal from the raw moments (before the conversion to Gauss weights):
// dN_a/dxi_j = 2^-d s_j prod_{k != j} (1 + s_k xi_k) (s = the node's corner signs), so
//   sum_p om_p (dN_a/dxi_j)^2 = 4^-d sum over the other axes of c_p(s_k) c_q(s_l) M[...]
// with c_0 = 1, c_1 = 2 s, c_2 = 1 and exponent 0 on axis j; K[(a,i),(a,i)] =
// sc_i^2 T_i + sum_j sc_j^2 T_j (sc = 2/h).  Equal to kdiag_entry of the converted weights: the
// integrand has degree <= 2 per axis, which the Gauss weights reproduce exactly.
template <int DIM>
__device__ __forceinline__ void kdiag_from_moments(const double* M, int a, const double* sc, double* kd) {
  const double sgn[3] = {(a & 1) ? 1.0 : -1.0, (a & 2) ? 1.0 : -1.0, (a & 4) ? 1.0 : -1.0};
  double T[3];
  if (DIM == 3) {
    const int str[3] = {1, 3, 9};
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const int k = j == 0 ? 1 : 0, l = j == 2 ? 1 : 2;
      const double ck[3] = {1.0, 2.0 * sgn[k], 1.0}, cl[3] = {1.0, 2.0 * sgn[l], 1.0};
      double t = 0.0;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        double u = 0.0;
#pragma unroll
        for (int p = 0; p < 3; p++) u += ck[p] * M[p * str[k] + q * str[l]];
        t += cl[q] * u;
      }
      T[j] = (1.0 / 64.0) * sc[j] * sc[j] * t;
    }
    const double sum = T[0] + T[1] + T[2];
#pragma unroll
    for (int i = 0; i < 3; i++) kd[i] = T[i] + sum;
  } else {
    const int str[2] = {1, 3};
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int k = 1 - j;
      const double ck[3] = {1.0, 2.0 * sgn[k], 1.0};
      double t = 0.0;
#pragma unroll
      for (int p = 0; p < 3; p++) t += ck[p] * M[p * str[k]];
      T[j] = (1.0 / 16.0) * sc[j] * sc[j] * t;
    }
    const double sum = T[0] + T[1];
#pragma unroll
    for (int i = 0; i < 2; i++) kd[i] = T[i] + sum;
  }
}

// ---------------------------------------------------------------------------------------
// A6/A7/A9 particle -> element moments, element force vector, approxS diagonal.
// Eight lanes per element: lane l takes the element's slots l, l + 8, ... (consecutive lanes
// read consecutive slots, and the dependent loads -- particle id, material, table -- of four
// slots per lane are issued together), then a three-step butterfly over the eight lanes sums the 3^d
// moments and the force terms (every lane ends with the same sums; fixed order, so the result
// is deterministic).  Each lane converts the moments and writes every eighth output value.
// Function values by particle id: material tables (mat != null) or per-particle arrays.
//   W[g*Es+e] (Es: padded stride), fe[e*NN*DIM + a*DIM + i], approxS[e]
// f_e: VectorAssemblyTerm_NA__Fn.cpp:143-202; approxS: UzawaPreconditionerTerm.c:124-209
// ---------------------------------------------------------------------------------------
struct MomArgs {
  int dim, res[3];
  double h[3], detJ;
  long E, Es;
  int nmat;
  double Ge[24];  // constant element gradient matrix, passed by value
};

constexpr int kMomLanes = 8;
constexpr int kMomUnroll = 4;

template <int DIM>
__global__ __launch_bounds__(256) void k_moments(MomArgs A, const int* __restrict__ cell_ptr,
                                                 const int* __restrict__ cell_pid, const double* __restrict__ xi_slot,
                                                 const double* __restrict__ w_slot, const int* __restrict__ mat,
                                                 const double* __restrict__ eta_tab,
                                                 const double* __restrict__ force_tab, const double* __restrict__ eta_p,
                                                 const double* __restrict__ force_p, double* __restrict__ W,
                                                 double* __restrict__ fe, double* __restrict__ approxS) {
  constexpr int NN = DIM == 3 ? 8 : 4, Q = DIM == 3 ? 27 : 9, NF = NN * DIM;
  const int lane = threadIdx.x & (kMomLanes - 1);
  const long e = ((long)blockIdx.x * blockDim.x + threadIdx.x) / kMomLanes;
  const bool live = e < A.E;
  double M[Q], F[NF];
#pragma unroll
  for (int q = 0; q < Q; q++) M[q] = 0.0;
#pragma unroll
  for (int q = 0; q < NF; q++) F[q] = 0.0;
  const int s0 = live ? cell_ptr[e] : 0, s1 = live ? cell_ptr[e + 1] : 0;
  // kMomUnroll slots per lane and pass: the dependent loads (particle id, then its material,
  // then the table entries) of the batch are issued together
  for (int sb = s0 + lane; sb < s1; sb += kMomLanes * kMomUnroll) {
    int pid[kMomUnroll];
    bool ok[kMomUnroll];
    double xi[kMomUnroll][3], w[kMomUnroll];
#pragma unroll
    for (int u = 0; u < kMomUnroll; u++) {
      const int s = sb + u * kMomLanes;
      ok[u] = s < s1;
      const int sc = ok[u] ? s : sb;
      pid[u] = cell_pid[sc];
      for (int d = 0; d < DIM; d++) xi[u][d] = xi_slot[(long)sc * DIM + d];
      w[u] = w_slot[sc] * A.detJ;
    }
    double eta[kMomUnroll], f[kMomUnroll][3];
#pragma unroll
    for (int u = 0; u < kMomUnroll; u++) {
      const int p = pid[u];
      if (mat) {
        int m = mat[p];
        if (m < 0 || m >= A.nmat) {  // index outside the table: poison the element
          m = 0;
          eta[u] = __builtin_nan("");
        } else {
          eta[u] = eta_tab[m];
        }
        for (int d = 0; d < DIM; d++) f[u][d] = force_tab[m * DIM + d];
      } else {
        eta[u] = eta_p[p];
        for (int d = 0; d < DIM; d++) f[u][d] = force_p[(long)p * DIM + d];
      }
    }
#pragma unroll
    for (int u = 0; u < kMomUnroll; u++) {
      if (!ok[u]) continue;
      const double om = eta[u] * w[u];
      const double px[3] = {1.0, xi[u][0], xi[u][0] * xi[u][0]}, py[3] = {1.0, xi[u][1], xi[u][1] * xi[u][1]};
      if (DIM == 3) {
        const double pz[3] = {1.0, xi[u][2], xi[u][2] * xi[u][2]};
#pragma unroll
        for (int kz = 0; kz < 3; kz++)
#pragma unroll
          for (int ky = 0; ky < 3; ky++) {
            const double t = om * pz[kz] * py[ky];
#pragma unroll
            for (int kx = 0; kx < 3; kx++) M[kz * 9 + ky * 3 + kx] += t * px[kx];
          }
      } else {
#pragma unroll
        for (int ky = 0; ky < 3; ky++) {
          const double t = om * py[ky];
#pragma unroll
          for (int kx = 0; kx < 3; kx++) M[ky * 3 + kx] += t * px[kx];
        }
      }
      // f_e[a][i] += w detJ f_i N_a(xi)
      const double lx[2] = {0.5 * (1.0 - xi[u][0]), 0.5 * (1.0 + xi[u][0])};
      const double ly[2] = {0.5 * (1.0 - xi[u][1]), 0.5 * (1.0 + xi[u][1])};
      double lz[2] = {1.0, 1.0};
      if (DIM == 3) {
        lz[0] = 0.5 * (1.0 - xi[u][2]);
        lz[1] = 0.5 * (1.0 + xi[u][2]);
      }
      double wf[3];
      for (int d = 0; d < DIM; d++) wf[d] = w[u] * f[u][d];
#pragma unroll
      for (int a = 0; a < NN; a++) {
        const double Na = lx[a & 1] * ly[(a >> 1) & 1] * lz[(a >> 2) & 1];
#pragma unroll
        for (int d = 0; d < DIM; d++) F[a * DIM + d] += wf[d] * Na;
      }
    }
  }
  // butterfly over the element's eight lanes (all lanes take part; dead elements add zeros)
#pragma unroll
  for (int m = 1; m < kMomLanes; m <<= 1) {
#pragma unroll
    for (int q = 0; q < Q; q++) M[q] += __shfl_xor(M[q], m, kMomLanes);
#pragma unroll
    for (int q = 0; q < NF; q++) F[q] += __shfl_xor(F[q], m, kMomLanes);
  }
  if (!live) return;
  // approxS_e = sum_v Ge[v]^2 / Ke[v][v], Ke's diagonal from the raw moments
  const double sc[3] = {2.0 / A.h[0], 2.0 / A.h[1], DIM == 3 ? 2.0 / A.h[2] : 1.0};
  if (lane == 0) {
    double S = 0.0;
#pragma unroll
    for (int a = 0; a < NN; a++) {
      double kd[3];
      kdiag_from_moments<DIM>(M, a, sc, kd);
#pragma unroll
      for (int i = 0; i < DIM; i++) {
        const double g = A.Ge[a * DIM + i];
        S += g * (g / kd[i]);
      }
    }
    approxS[e] = S;
  }
  moments_to_weights<DIM>(M);
#pragma unroll
  for (int q = 0; q < Q; q++)
    if ((q & (kMomLanes - 1)) == lane) W[(long)q * A.Es + e] = M[q];
#pragma unroll
  for (int q = 0; q < NF; q++)
    if ((q & (kMomLanes - 1)) == lane) fe[e * NF + q] = F[q];
}

static MomArgs mom_args(const Geom& g) {
  MomArgs A;
  A.dim = g.dim;
  A.detJ = 1.0;
  for (int d = 0; d < 3; d++) {
    A.res[d] = g.res[d];
    A.h[d] = d < g.dim ? (g.hi[d] - g.lo[d]) / g.res[d] : 1.0;
    if (d < g.dim) A.detJ *= 0.5 * A.h[d];
  }
  A.E = g.E;
  A.Es = g.Es;
  A.nmat = 0;
  return A;
}

void launch_moments(const Geom& g, const int* cell_ptr, const int* cell_pid, const double* xi_slot,
                    const double* w_slot, const int* mat, const double* eta_tab, const double* force_tab,
                    const double* eta_p, const double* force_p, double* W, double* fe, double* approxS,
                    const double* Ge, int nmat, hipStream_t s) {
  MomArgs A = mom_args(g);
  A.nmat = nmat;
  for (int q = 0; q < 24; q++) A.Ge[q] = q < (1 << g.dim) * g.dim ? Ge[q] : 0.0;
  const unsigned nb = nblocks(g.E * kMomLanes, 256);
  if (g.dim == 3)
    hipLaunchKernelGGL(k_moments<3>, nb, 256, 0, s, A, cell_ptr, cell_pid, xi_slot, w_slot, mat, eta_tab, force_tab,
                       eta_p, force_p, W, fe, approxS);
  else
    hipLaunchKernelGGL(k_moments<2>, nb, 256, 0, s, A, cell_ptr, cell_pid, xi_slot, w_slot, mat, eta_tab, force_tab,
                       eta_p, force_p, W, fe, approxS);
  UWG_HIP(hipGetLastError());
}

// Gauss-swarm assembly (no voronoi swarm, PY/swarm/_integration_swarm.py:173-181): the weights
// are eta(g) w_g detJ at the 3^dim Gauss points, f_e = sum_g w_g detJ f(g) N_a(g), approxS as
// above.  eta_g [E][Q], force_g [E][Q][DIM] (null: zero force), point index qx + 3 qy + 9 qz.
template <int DIM>
__global__ void k_gauss_assemble(MomArgs A, const double* __restrict__ eta_g, const double* __restrict__ force_g,
                                 double* __restrict__ W, double* __restrict__ fe, double* __restrict__ approxS) {
  constexpr int NN = DIM == 3 ? 8 : 4, Q = DIM == 3 ? 27 : 9;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= A.E) return;
  const double w1[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  double M[27], F[NN * DIM];
  for (int q = 0; q < NN * DIM; q++) F[q] = 0.0;
  for (int q = 0; q < Q; q++) {
    const double wq = w1[q % 3] * w1[(q / 3) % 3] * (DIM == 3 ? w1[q / 9] : 1.0);
    M[q] = eta_g[e * Q + q] * wq * A.detJ;
    if (force_g) {
      const double X[3] = {gpt(q % 3), gpt((q / 3) % 3), gpt(q / 9)};
      for (int a = 0; a < NN; a++) {
        double Na = 1.0;
        for (int d = 0; d < DIM; d++) Na *= 0.5 * (1.0 + (((a >> d) & 1) ? 1.0 : -1.0) * X[d]);
        for (int d = 0; d < DIM; d++) F[a * DIM + d] += wq * A.detJ * force_g[(e * Q + q) * DIM + d] * Na;
      }
    }
  }
  for (int q = 0; q < Q; q++) W[(long)q * A.Es + e] = M[q];
  for (int q = 0; q < NN * DIM; q++) fe[e * NN * DIM + q] = F[q];
  double sc[3] = {2.0 / A.h[0], 2.0 / A.h[1], DIM == 3 ? 2.0 / A.h[2] : 1.0};
  double S = 0.0;
  for (int a = 0; a < NN; a++)
    for (int i = 0; i < DIM; i++) {
      const double kd = kdiag_entry<DIM>(M, a, i, sc);
      const double g = A.Ge[a * DIM + i];
      S += g * (g / kd);
    }
  approxS[e] = S;
}

void launch_gauss_assemble(const Geom& g, const double* eta_g, const double* force_g, double* W, double* fe,
                           double* approxS, const double* Ge, hipStream_t s) {
  MomArgs A = mom_args(g);
  for (int q = 0; q < 24; q++) A.Ge[q] = q < (1 << g.dim) * g.dim ? Ge[q] : 0.0;
  if (g.dim == 3)
    hipLaunchKernelGGL(k_gauss_assemble<3>, nblocks(g.E, 128), 128, 0, s, A, eta_g, force_g, W, fe, approxS);
  else
    hipLaunchKernelGGL(k_gauss_assemble<2>, nblocks(g.E, 128), 128, 0, s, A, eta_g, force_g, W, fe, approxS);
  UWG_HIP(hipGetLastError());
}

// f[n][i] = sum over the elements around node n of fe (fixed element order); 0 on BC dofs
template <int DIM>
__global__ void k_gather_fe(int rx, int ry, int rz, long N, const double* __restrict__ fe,
                            const uint8_t* __restrict__ bc, double* __restrict__ f) {
  constexpr int NN = DIM == 3 ? 8 : 4;
  long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  long nx = rx + 1, ny = ry + 1;
  int i = (int)(n % nx), j = (int)((n / nx) % ny), k = (int)(n / (nx * ny));
  double acc[3] = {0.0, 0.0, 0.0};
  for (int oz = 0; oz < (DIM == 3 ? 2 : 1); oz++)
    for (int oy = 0; oy < 2; oy++)
      for (int ox = 0; ox < 2; ox++) {
        int ei = i - 1 + ox, ej = j - 1 + oy, ek = DIM == 3 ? k - 1 + oz : 0;
        if (ei < 0 || ei >= rx || ej < 0 || ej >= ry || ek < 0 || ek >= rz) continue;
        long e = ei + (long)rx * (ej + (long)ry * ek);
        int a = (1 - ox) + 2 * (1 - oy) + (DIM == 3 ? 4 * (1 - oz) : 0);
        for (int d = 0; d < DIM; d++) acc[d] += fe[e * NN * DIM + a * DIM + d];
      }
  uint8_t m = bc[n];
  for (int d = 0; d < DIM; d++) f[n * DIM + d] = ((m >> d) & 1) ? 0.0 : acc[d];
}

void launch_node_gather_fe(const Geom& g, const double* fe, const uint8_t* bc, double* f, hipStream_t s) {
  if (g.dim == 3)
    hipLaunchKernelGGL(k_gather_fe<3>, nblocks(g.N, 256), 256, 0, s, g.res[0], g.res[1], g.res[2], g.N, fe, bc, f);
  else
    hipLaunchKernelGGL(k_gather_fe<2>, nblocks(g.N, 256), 256, 0, s, g.res[0], g.res[1], 1, g.N, fe, bc, f);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// Galerkin coarse weights: children's Gauss points mapped into the coarse element
// (xi_c = (xi_child + s)/2, s = -1/+1) give the coarse moments exactly.
// ---------------------------------------------------------------------------------------
template <int DIM>
__global__ void k_restrict_W(int crx, int cry, int crz, long Ec, long Ecs, long Efs, const double* __restrict__ Wf,
                             double* __restrict__ Wc) {
  constexpr int Q = DIM == 3 ? 27 : 9;
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Ec) return;
  int ci = (int)(e % crx), cj = (int)((e / crx) % cry), ck = (int)(e / ((long)crx * cry));
  int frx = 2 * crx, fry = 2 * cry;
  double M[27];
  for (int q = 0; q < Q; q++) M[q] = 0.0;
  for (int oz = 0; oz < (DIM == 3 ? 2 : 1); oz++)
    for (int oy = 0; oy < 2; oy++)
      for (int ox = 0; ox < 2; ox++) {
        long fe = (2 * ci + ox) + (long)frx * ((2 * cj + oy) + (long)fry * (DIM == 3 ? (2 * ck + oz) : 0));
        double sx = ox ? 1.0 : -1.0, sy = oy ? 1.0 : -1.0, sz = oz ? 1.0 : -1.0;
        for (int g = 0; g < Q; g++) {
          double w = Wf[(long)g * Efs + fe];
          double X = 0.5 * (gpt(g % 3) + sx), Y = 0.5 * (gpt((g / 3) % 3) + sy);
          double px[3] = {1.0, X, X * X}, py[3] = {1.0, Y, Y * Y};
          if (DIM == 3) {
            double Z = 0.5 * (gpt(g / 9) + sz);
            double pz[3] = {1.0, Z, Z * Z};
            for (int kz = 0; kz < 3; kz++)
              for (int ky = 0; ky < 3; ky++)
                for (int kx = 0; kx < 3; kx++) M[kz * 9 + ky * 3 + kx] += w * pz[kz] * py[ky] * px[kx];
          } else {
            for (int ky = 0; ky < 3; ky++)
              for (int kx = 0; kx < 3; kx++) M[ky * 3 + kx] += w * py[ky] * px[kx];
          }
        }
      }
  moments_to_weights<DIM>(M);
  for (int q = 0; q < Q; q++) Wc[(long)q * Ecs + e] = M[q];
}

void launch_restrict_W(const Geom& fine, const Geom& coarse, const double* Wf, double* Wc, hipStream_t s) {
  if (coarse.dim == 3)
    hipLaunchKernelGGL(k_restrict_W<3>, nblocks(coarse.E, 128), 128, 0, s, coarse.res[0], coarse.res[1],
                       coarse.res[2], coarse.E, coarse.Es, fine.Es, Wf, Wc);
  else
    hipLaunchKernelGGL(k_restrict_W<2>, nblocks(coarse.E, 128), 128, 0, s, coarse.res[0], coarse.res[1], 1,
                       coarse.E, coarse.Es, fine.Es, Wf, Wc);
  UWG_HIP(hipGetLastError());
}

// inverse diagonal of K (identity rows on Dirichlet dofs)
template <int DIM>
__global__ void k_diag(int rx, int ry, int rz, long N, long Es, double hx, double hy, double hz,
                       const double* __restrict__ W, const uint8_t* __restrict__ bc, double* __restrict__ dinv) {
  constexpr int Q = DIM == 3 ? 27 : 9;
  long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  long nx = rx + 1, ny = ry + 1;
  int i = (int)(n % nx), j = (int)((n / nx) % ny), k = (int)(n / (nx * ny));
  double sc[3] = {2.0 / hx, 2.0 / hy, 2.0 / hz};
  double acc[3] = {0.0, 0.0, 0.0};
  for (int oz = 0; oz < (DIM == 3 ? 2 : 1); oz++)
    for (int oy = 0; oy < 2; oy++)
      for (int ox = 0; ox < 2; ox++) {
        int ei = i - 1 + ox, ej = j - 1 + oy, ek = DIM == 3 ? k - 1 + oz : 0;
        if (ei < 0 || ei >= rx || ej < 0 || ej >= ry || ek < 0 || ek >= rz) continue;
        long e = ei + (long)rx * (ej + (long)ry * ek);
        int a = (1 - ox) + 2 * (1 - oy) + (DIM == 3 ? 4 * (1 - oz) : 0);
        double Wg[27];
        for (int g = 0; g < Q; g++) Wg[g] = W[(long)g * Es + e];
        for (int d = 0; d < DIM; d++) acc[d] += kdiag_entry<DIM>(Wg, a, d, sc);
      }
  uint8_t m = bc[n];
  for (int d = 0; d < DIM; d++) dinv[n * DIM + d] = ((m >> d) & 1) ? 1.0 : 1.0 / acc[d];
}

void launch_diag(const Geom& g, const double* W, const uint8_t* bc, double* dinv, hipStream_t s) {
  double h[3];
  for (int d = 0; d < 3; d++) h[d] = d < g.dim ? (g.hi[d] - g.lo[d]) / g.res[d] : 1.0;
  if (g.dim == 3)
    hipLaunchKernelGGL(k_diag<3>, nblocks(g.N, 128), 128, 0, s, g.res[0], g.res[1], g.res[2], g.N, g.Es, h[0], h[1],
                       h[2], W, bc, dinv);
  else
    hipLaunchKernelGGL(k_diag<2>, nblocks(g.N, 128), 128, 0, s, g.res[0], g.res[1], 1, g.N, g.Es, h[0], h[1], h[2],
                       W, bc, dinv);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// Matrix-free K apply.  Element kernel in monomial form: u = c0 + cx X + cy Y + cz Z +
// cxy XY + cxz XZ + cyz YZ + cxyz XYZ per component; gradients at the 3^d Gauss points
// are 3-term sums; stresses sigma = W (G + G^T); the transposed contraction gives the
// dual coefficients and r_a = (1/8) H^T dc.
//
// 3D tiling: a 16x16 block of element columns streams up a z-chunk.  Each thread owns one
// element column; node (lx,ly) with 1 <= lx,ly <= 15 sums the four surrounding
// elements' corner contributions from LDS in a fixed order (deterministic, no atomics).
// Tiles overlap by one element so every node has exactly one owner.
// ---------------------------------------------------------------------------------------
template <typename T>
struct KArgs {
  int rx, ry, rz, KZ;  // KZ: node levels per workgroup (3D), element rows per block (2D)
  T sx, sy, sz;
  long Ws, nx, nxy;  // Ws: stride of the per-point weight arrays
  int ntx, nty;      // 3D: tiles of 16 x 16 element columns (overlapping by one)
};

template <typename T, typename WF>
__device__ __forceinline__ void elem3(WF wq, const T (&u)[8][3], T (&r)[8][3], T sx, T sy, T sz) {
  const T q8 = (T)0.125;
  T cX[3], cY[3], cZ[3], cXY[3], cXZ[3], cYZ[3], cXYZ[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    T a0 = u[0][i], a1 = u[1][i], a2 = u[2][i], a3 = u[3][i], a4 = u[4][i], a5 = u[5][i], a6 = u[6][i], a7 = u[7][i];
    // x differences / sums
    T d01 = a1 - a0, d23 = a3 - a2, d45 = a5 - a4, d67 = a7 - a6;
    T s01 = a1 + a0, s23 = a3 + a2, s45 = a5 + a4, s67 = a7 + a6;
    cX[i] = q8 * (d01 + d23 + d45 + d67);
    cXY[i] = q8 * ((d23 - d01) + (d67 - d45));
    cXZ[i] = q8 * ((d45 - d01) + (d67 - d23));
    cXYZ[i] = q8 * ((d67 - d45) - (d23 - d01));
    cY[i] = q8 * ((s23 - s01) + (s67 - s45));
    cZ[i] = q8 * ((s45 - s01) + (s67 - s23));
    cYZ[i] = q8 * ((s67 - s45) - (s23 - s01));
  }
  T dX[3] = {0, 0, 0}, dY[3] = {0, 0, 0}, dZ[3] = {0, 0, 0}, dXY[3] = {0, 0, 0}, dXZ[3] = {0, 0, 0},
    dYZ[3] = {0, 0, 0}, dXYZ[3] = {0, 0, 0};
  // Partial sum factorisation.  du/dx is constant along x, so within a plane of constant Z its
  // values at the 9 points are 3 "lines" indexed by Y; likewise du/dy by X.  The dual
  // contraction factorises the same way: the x-stress is summed along each x-line before it
  // is folded into the dual coefficients.  du/dz and the z-stress are formed point by point
  // and folded at once, which keeps only one plane's lines live (register pressure sets the
  // occupancy of this kernel).
#pragma unroll
  for (int gz = 0; gz < 3; gz++) {
    const T Z = (T)gpt(gz);
    T GX[3][3], GY[3][3], PX[3][3], PY[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const T K = (T)gpt(k);  // Y for x-lines, X for y-lines
        GX[i][k] = sx * (cX[i] + cXY[i] * K + cXZ[i] * Z + cXYZ[i] * (K * Z));
        GY[i][k] = sy * (cY[i] + cXY[i] * K + cYZ[i] * Z + cXYZ[i] * (K * Z));
        PX[i][k] = (T)0;
        PY[i][k] = (T)0;
      }
#pragma unroll
    for (int gy = 0; gy < 3; gy++)
#pragma unroll
      for (int gx = 0; gx < 3; gx++) {
        const T X = (T)gpt(gx), Y = (T)gpt(gy);
        const T w = wq(gx + 3 * gy + 9 * gz);
        const T w2 = w + w;
        const T G00 = GX[0][gy], G10 = GX[1][gy], G20 = GX[2][gy];
        const T G01 = GY[0][gx], G11 = GY[1][gx], G21 = GY[2][gx];
        T G2[3];
#pragma unroll
        for (int i = 0; i < 3; i++) G2[i] = sz * (cZ[i] + cXZ[i] * X + cYZ[i] * Y + cXYZ[i] * (X * Y));
        const T S01 = w * (G01 + G10), S02 = w * (G2[0] + G20), S12 = w * (G2[1] + G21);
        PX[0][gy] += w2 * G00;
        PX[1][gy] += S01;
        PX[2][gy] += S02;
        PY[0][gx] += S01;
        PY[1][gx] += w2 * G11;
        PY[2][gx] += S12;
        const T pz[3] = {S02 * sz, S12 * sz, (w2 * G2[2]) * sz};
#pragma unroll
        for (int i = 0; i < 3; i++) {
          dZ[i] += pz[i];
          dXZ[i] += pz[i] * X;
          dYZ[i] += pz[i] * Y;
          dXYZ[i] += pz[i] * (X * Y);
        }
      }
    // fold the x- and y-lines of this plane into the dual coefficients
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const T K = (T)gpt(k);
        const T px = PX[i][k] * sx, py = PY[i][k] * sy;
        dX[i] += px;
        dXY[i] += px * K;
        dXZ[i] += px * Z;
        dXYZ[i] += px * (K * Z);
        dY[i] += py;
        dXY[i] += py * K;
        dYZ[i] += py * Z;
        dXYZ[i] += py * (K * Z);
      }
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    // r_a = (1/8) (s_a dX + t_a dY + r_a dZ + s t dXY + s r dXZ + t r dYZ + s t r dXYZ)
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const T s = (a & 1) ? 1 : -1, t = (a & 2) ? 1 : -1, rr = (a & 4) ? 1 : -1;
      r[a][i] = q8 * (s * dX[i] + t * dY[i] + rr * dZ[i] + (s * t) * dXY[i] + (s * rr) * dXZ[i] + (t * rr) * dYZ[i] +
                      (s * t * rr) * dXYZ[i]);
    }
  }
}

// Line-factorised form of elem3 (same operator, about half the arithmetic).  Each gradient
// component is constant along one axis: du/dx along x, du/dy along y, du/dz along z.  Its
// values at the 27 points are therefore 9 "lines" per component (indexed by the other two
// abscissae), and the dual contraction collects each stress component along the same lines
// before folding it into the dual coefficients.  The z-lines are formed once per element
// and accumulated over the three planes (27 values + 27 accumulators held), the y-lines once
// per plane, the x-lines once per row.  Abscissae enter only as signs s in {-1, 0, 1}
// (X = s g, g = sqrt(3/5)): line values and folds are additions of pre-scaled coefficients,
// the powers of g and the 1/8 and 2/h factors are applied once to the coefficients and once
// to the dual coefficients, and the normal-stress factor 2 is folded into the line of the
// component's own direction.
template <typename T, typename WF>
__device__ __forceinline__ void elem3_lines(WF wq, const T (&u)[8][3], T (&r)[8][3], T sx, T sy, T sz) {
  const T g = (T)GA, g2 = (T)GA2;
  const T hx = (T)0.125 * sx, hy = (T)0.125 * sy, hz = (T)0.125 * sz;
  // x-line of component i at (Y = g sy, Z = g sz): (ax + sz axz) + sy (bx + sz bxz)
  T ax[3], axz[3], bx[3], bxz[3];
  T ay[3], ayz[3], by[3], byz[3];
  T GZ[3][3][3];  // [i][gy][gx]
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const T a0 = u[0][i], a1 = u[1][i], a2 = u[2][i], a3 = u[3][i], a4 = u[4][i], a5 = u[5][i], a6 = u[6][i],
            a7 = u[7][i];
    const T d01 = a1 - a0, d23 = a3 - a2, d45 = a5 - a4, d67 = a7 - a6;
    const T s01 = a1 + a0, s23 = a3 + a2, s45 = a5 + a4, s67 = a7 + a6;
    const T e1 = d23 - d01, e2 = d67 - d45, f1 = d01 + d23, f2 = d45 + d67;
    const T k1 = s23 - s01, k2 = s67 - s45, m1 = s01 + s23, m2 = s45 + s67;
    // sum_a of (node sign products) x u_a: u = (c + cX X + cY Y + cZ Z + cXY XY + ...) / 8
    const T cX = f1 + f2, cXY = e1 + e2, cXZ = f2 - f1, cXYZ = e2 - e1;
    const T cY = k1 + k2, cYZ = k2 - k1, cZ = m2 - m1;
    const T fx = (i == 0 ? (T)2 : (T)1) * hx, fy = (i == 1 ? (T)2 : (T)1) * hy, fz = (i == 2 ? (T)2 : (T)1) * hz;
    ax[i] = fx * cX;
    axz[i] = (fx * g) * cXZ;
    bx[i] = (fx * g) * cXY;
    bxz[i] = (fx * g2) * cXYZ;
    ay[i] = fy * cY;
    ayz[i] = (fy * g) * cYZ;
    by[i] = (fy * g) * cXY;
    byz[i] = (fy * g2) * cXYZ;
    const T A = fz * cZ, B = (fz * g) * cXZ, C = (fz * g) * cYZ, D = (fz * g2) * cXYZ;
#pragma unroll
    for (int gy = 0; gy < 3; gy++) {
      const T a = gy == 0 ? A - C : (gy == 1 ? A : A + C);
      const T b = gy == 0 ? B - D : (gy == 1 ? B : B + D);
      GZ[i][gy][0] = a - b;
      GZ[i][gy][1] = a;
      GZ[i][gy][2] = a + b;
    }
  }
  T PZ[3][3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int gy = 0; gy < 3; gy++)
#pragma unroll
      for (int gx = 0; gx < 3; gx++) PZ[i][gy][gx] = (T)0;
  // x-, y-part dual accumulators (before the h and g factors)
  T eX[3], eXYx[3], eXZx[3], eXYZx[3], eY[3], eXYy[3], eYZy[3], eXYZy[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    eX[i] = eXYx[i] = eXZx[i] = eXYZx[i] = (T)0;
    eY[i] = eXYy[i] = eYZy[i] = eXYZy[i] = (T)0;
  }
#pragma unroll
  for (int gz = 0; gz < 3; gz++) {
    T GY[3][3], PY[3][3];  // [i][gx]
    T axp[3], bxp[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const T aY = gz == 0 ? ay[i] - ayz[i] : (gz == 1 ? ay[i] : ay[i] + ayz[i]);
      const T bY = gz == 0 ? by[i] - byz[i] : (gz == 1 ? by[i] : by[i] + byz[i]);
      GY[i][0] = aY - bY;
      GY[i][1] = aY;
      GY[i][2] = aY + bY;
      PY[i][0] = PY[i][1] = PY[i][2] = (T)0;
      axp[i] = gz == 0 ? ax[i] - axz[i] : (gz == 1 ? ax[i] : ax[i] + axz[i]);
      bxp[i] = gz == 0 ? bx[i] - bxz[i] : (gz == 1 ? bx[i] : bx[i] + bxz[i]);
    }
#pragma unroll
    for (int gy = 0; gy < 3; gy++) {
      T GX[3], PX[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        GX[i] = gy == 0 ? axp[i] - bxp[i] : (gy == 1 ? axp[i] : axp[i] + bxp[i]);
        PX[i] = (T)0;
      }
#pragma unroll
      for (int gx = 0; gx < 3; gx++) {
        const T w = wq(gx + 3 * gy + 9 * gz);
        const T t01 = GY[0][gx] + GX[1];
        const T t02 = GZ[0][gy][gx] + GX[2];
        const T t12 = GZ[1][gy][gx] + GY[2][gx];
        PX[0] += w * GX[0];
        PX[1] += w * t01;
        PX[2] += w * t02;
        PY[0][gx] += w * t01;
        PY[1][gx] += w * GY[1][gx];
        PY[2][gx] += w * t12;
        PZ[0][gy][gx] += w * t02;
        PZ[1][gy][gx] += w * t12;
        PZ[2][gy][gx] += w * GZ[2][gy][gx];
      }
      // fold the x-line (gy, gz): weights 1, Y, Z, YZ
#pragma unroll
      for (int i = 0; i < 3; i++) {
        eX[i] += PX[i];
        if (gy == 0) eXYx[i] -= PX[i];
        if (gy == 2) eXYx[i] += PX[i];
        if (gz == 0) eXZx[i] -= PX[i];
        if (gz == 2) eXZx[i] += PX[i];
        if (gy != 1 && gz != 1) {
          if (gy == gz)
            eXYZx[i] += PX[i];
          else
            eXYZx[i] -= PX[i];
        }
      }
    }
    // fold the y-lines of plane gz: weights 1, X, Z, XZ
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const T s = PY[i][0] + PY[i][1] + PY[i][2], dd = PY[i][2] - PY[i][0];
      eY[i] += s;
      eXYy[i] += dd;
      if (gz == 0) {
        eYZy[i] -= s;
        eXYZy[i] -= dd;
      }
      if (gz == 2) {
        eYZy[i] += s;
        eXYZy[i] += dd;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    // fold the z-lines: weights 1, X, Y, XY
    T q0[3], q1[3];
#pragma unroll
    for (int gy = 0; gy < 3; gy++) {
      q0[gy] = PZ[i][gy][0] + PZ[i][gy][1] + PZ[i][gy][2];
      q1[gy] = PZ[i][gy][2] - PZ[i][gy][0];
    }
    const T eZ = q0[0] + q0[1] + q0[2], eYZz = q0[2] - q0[0];
    const T eXZz = q1[0] + q1[1] + q1[2], eXYZz = q1[2] - q1[0];
    const T dX = hx * eX[i], dY = hy * eY[i], dZ = hz * eZ;
    const T dXY = g * (hx * eXYx[i] + hy * eXYy[i]);
    const T dXZ = g * (hx * eXZx[i] + hz * eXZz);
    const T dYZ = g * (hy * eYZy[i] + hz * eYZz);
    const T dXYZ = g2 * (hx * eXYZx[i] + hy * eXYZy[i] + hz * eXYZz);
    // r_a = s dX + t dY + r dZ + st dXY + sr dXZ + tr dYZ + str dXYZ, a = (s, t, r) bits
    const T pp = dX + dXY, pm = dX - dXY;  // s-coefficient of F for t = +1 / -1
    const T F0 = -pm - dY, F1 = pm - dY, F2 = dY - pp, F3 = pp + dY;  // F(s,t), a = s + 2t bits
    const T hp = dZ + dYZ, hm = dZ - dYZ, kp = dXZ + dXYZ, km = dXZ - dXYZ;
    const T H0 = hm - km, H1 = hm + km, H2 = hp - kp, H3 = hp + kp;  // H(s,t)
    r[0][i] = F0 - H0;
    r[1][i] = F1 - H1;
    r[2][i] = F2 - H2;
    r[3][i] = F3 - H3;
    r[4][i] = F0 + H0;
    r[5][i] = F1 + H1;
    r[6][i] = F2 + H2;
    r[7][i] = F3 + H3;
  }
}

// Loads and stores addressed as a uniform base plus a 32-bit byte offset: the address is then
// formed by the memory instruction (scalar base + vector offset) instead of by 64-bit vector
// arithmetic per access.  The launcher checks that the arrays of the 3D apply (the weights of
// all 27 planes included) stay below 4 GiB.
template <typename T>
__device__ __forceinline__ T ld_off(const T* __restrict__ base, uint32_t boff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}
template <typename T>
__device__ __forceinline__ void st_off(T* __restrict__ base, uint32_t boff, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}

// The element weights of the FP32 apply are read through a buffer resource: the plane offset
// q * Ws * 4 is a scalar operand of the load and the element offset a 32-bit vector operand,
// so the weight loads of a level need no vector address arithmetic.  Descriptor word 3 = 0x00020000
// (32-bit data format, the raw-buffer setting of the gfx9 family); num_records bounds the
// buffer in bytes (loads past it return 0).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t weight_rsrc(const float* W, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(W), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t ld_buf_u32(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}

// Node values of one level of an element column tile (corners c = 0..3 of element (ei, ej)).
// Every load is issued unconditionally (out-of-range corners read node 0 and are zeroed when the
// level is finished), so no load waits on another: the mask byte does not gate the value loads.
template <typename T>
struct LevelLoad {
  T v[4][3];
  uint8_t m[4];
  bool ok[4];
  __device__ __forceinline__ void issue(const KArgs<T>& A, const T* __restrict__ x, const uint8_t* __restrict__ bc,
                                        int ei, int ej, int lev) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int ni = ei + (c & 1), nj = ej + (c >> 1);
      ok[c] = ni >= 0 && ni <= A.rx && nj >= 0 && nj <= A.ry && lev >= 0 && lev <= A.rz;
      const uint32_t n = ok[c] ? (uint32_t)(ni + A.nx * nj + A.nxy * lev) : 0u;
      m[c] = ld_off(bc, n);
#pragma unroll
      for (int i = 0; i < 3; i++) v[c][i] = ld_off(x, (n * 3u + i) * (uint32_t)sizeof(T));
    }
  }
  // u[off + c] = node values with Dirichlet components and absent nodes zeroed
  __device__ __forceinline__ void finish(T (&u)[8][3], int off) const {
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int i = 0; i < 3; i++) u[off + c][i] = (ok[c] && !((m[c] >> i) & 1)) ? v[c][i] : (T)0;
  }
};

template <typename T>
__device__ __forceinline__ void load_level3(const KArgs<T>& A, const T* __restrict__ x, const uint8_t* __restrict__ bc,
                                            int ei, int ej, int lev, T (&u)[8][3], int off) {
  LevelLoad<T> L;
  L.issue(A, x, bc, ei, ej, lev);
  L.finish(u, off);
}

// What the apply kernel writes for an owned node (the operator value is kx = (K x)_n, or x_n on
// a Dirichlet row):
//   APPLY_Y        y = kx
//   APPLY_RESIDUAL y = b - kx
//   APPLY_CHEB     r = b - kx; v = a d + c dinv r; d = v; y = x + v   (a Chebyshev step)
//   APPLY_CHEB0    the same with d taken as zero (the first step from a nonzero guess)
// The Chebyshev forms write the new iterate into y, a buffer distinct from x (other tiles
// still read x), so the smoother alternates two buffers.
enum ApplyMode { APPLY_Y = 0, APPLY_RESIDUAL = 1, APPLY_CHEB = 2, APPLY_CHEB0 = 3 };
template <typename T>
struct SmoothArgs {
  const T* b = nullptr;
  const T* dinv = nullptr;
  T* d = nullptr;
  T a = 0, c = 0;
};
// Operands of the owned node's update, loaded at the top of a level so that their latency
// overlaps the element work.
template <typename T, int MODE, int NC = 3>
struct NodeOperands {
  uint8_t m;
  T x[NC], b[NC], dinv[NC], d[NC];
  __device__ __forceinline__ void load(const SmoothArgs<T>& S, const T* __restrict__ xs, const uint8_t* __restrict__ bc,
                                       uint32_t n) {
    m = ld_off(bc, n);
#pragma unroll
    for (int i = 0; i < NC; i++) {
      const uint32_t o = (n * (uint32_t)NC + i) * (uint32_t)sizeof(T);
      x[i] = ld_off(xs, o);
      if (MODE != APPLY_Y) b[i] = ld_off(S.b, o);
      if (MODE == APPLY_CHEB || MODE == APPLY_CHEB0) dinv[i] = ld_off(S.dinv, o);
      if (MODE == APPLY_CHEB) d[i] = ld_off((const T*)S.d, o);
    }
  }
  // kx: the operator row sum of component i (replaced by x on a Dirichlet row)
  __device__ __forceinline__ void store(const SmoothArgs<T>& S, T* __restrict__ y, uint32_t n, int i, T kx) const {
    const uint32_t o = (n * (uint32_t)NC + i) * (uint32_t)sizeof(T);
    if ((m >> i) & 1) kx = x[i];
    if (MODE == APPLY_Y) {
      st_off(y, o, kx);
    } else {
      const T r = b[i] - kx;
      if (MODE == APPLY_RESIDUAL) {
        st_off(y, o, r);
      } else {
        const T v = MODE == APPLY_CHEB ? S.a * d[i] + S.c * (dinv[i] * r) : S.c * (dinv[i] * r);
        st_off(S.d, o, v);
        st_off(y, o, x[i] + v);
      }
    }
  }
};

// Workgroups are dealt to the 8 XCDs round-robin by linear id; renumber them so that each XCD
// receives a contiguous run of logical ids.
__device__ __forceinline__ int xcd_logical_id() {
  const int total = gridDim.x;
  const int id = blockIdx.x;
  const int xcd = id & 7, loc = id >> 3, per = total >> 3, rem = total & 7;
  return xcd < rem ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
}

// The workgroup's chunk: KZ node levels [k0, k1) of tile (bx, by).  Logical ids run x-fastest
// through the tiles of one z-chunk, so an XCD's concurrently running workgroups hold
// x-neighbouring tiles of the same levels, and the cache lines straddling two tiles' columns
// (each tile row covers 16 weights, half of a 128-byte line) are fetched into that XCD's L2
// once for both.  Ordering the work tile-major instead (long z-runs per tile) fetched the
// weights about twice over.
template <typename T, typename Body>
__device__ __forceinline__ void for_each_segment(const KArgs<T>& A, Body body) {
  const int t = xcd_logical_id();
  const int bx = t % A.ntx, by = (t / A.ntx) % A.nty, zc = t / (A.ntx * A.nty);
  const int k0 = zc * A.KZ;
  body(bx, by, k0, min(k0 + A.KZ, A.rz + 1));
}

// Sum of the four element-column contributions (this level's below / above halves) at the
// owned node of thread tid, component i.
template <typename T>
__device__ __forceinline__ T node_sum(const T (*sb)[256], const T (*st)[256], int tid, int i) {
  return st[9 + i][tid - 17] + st[6 + i][tid - 16] + st[3 + i][tid - 1] + st[i][tid] + sb[9 + i][tid - 17] +
         sb[6 + i][tid - 16] + sb[3 + i][tid - 1] + sb[i][tid];
}

// FP64 instance: held to 256 VGPRs (two waves per SIMD).  The element weights are read as the
// element kernel consumes them; the node levels and the owned node's operands are issued at the
// top of each level.
template <typename T, int MODE>
__global__ __launch_bounds__(256, 2) void k_apply3(KArgs<T> A, const T* __restrict__ W, const T* __restrict__ x,
                                                   T* __restrict__ y, const uint8_t* __restrict__ bc,
                                                   SmoothArgs<T> S) {
  __shared__ T sb[12][256];
  __shared__ T st[12][256];
  const int tid = threadIdx.x, lx = tid & 15, ly = tid >> 4;
  for_each_segment(A, [&](int bx, int by, int k0, int k1) {
    const int ei = bx * 15 - 1 + lx, ej = by * 15 - 1 + ly;
    const bool colE = ei >= 0 && ei < A.rx && ej >= 0 && ej < A.ry;
    const bool ownN = lx >= 1 && ly >= 1 && ei <= A.rx && ej <= A.ry;
    T u[8][3];
    load_level3(A, x, bc, ei, ej, k0 - 1, u, 0);
    for (int ek = k0 - 1; ek < k1; ++ek) {
      load_level3(A, x, bc, ei, ej, ek + 1, u, 4);
      const bool doN = ownN && ek >= k0;
      const uint32_t nN = doN ? (uint32_t)(ei + A.nx * ej + A.nxy * ek) : 0u;
      NodeOperands<T, MODE> nd;
      nd.load(S, x, bc, nN);
      T r[8][3];
      if (colE && ek >= 0 && ek < A.rz) {
        const uint32_t eo = (uint32_t)(ei + A.rx * (ej + A.ry * ek)) * (uint32_t)sizeof(T);
        elem3<T>([&](int q) { return ld_off(W + (long)q * A.Ws, eo); }, u, r, A.sx, A.sy, A.sz);
      } else {
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
          for (int i = 0; i < 3; i++) r[a][i] = (T)0;
      }
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int i = 0; i < 3; i++) sb[c * 3 + i][tid] = r[c][i];
      __syncthreads();
      if (doN) {
#pragma unroll
        for (int i = 0; i < 3; i++) nd.store(S, y, nN, i, node_sum<T>(sb, st, tid, i));
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int i = 0; i < 3; i++) st[c * 3 + i][tid] = r[4 + c][i];
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int i = 0; i < 3; i++) u[c][i] = u[4 + c][i];
    }
  });
}

// FP32 instance, software-pipelined: while level ek is computed, the loads of the next level
// are in flight -- node level ek + 2 and the weights of element level ek + 1, held in
// registers until the top of the next level, where the weights move to this lane's own LDS
// column (read back as the line-factorised element kernel consumes them; no barrier is needed,
// each lane reads only what it wrote).  Memory and arithmetic of a workgroup then overlap
// instead of alternating.  Two waves per SIMD (up to 256 VGPRs).
// Weight formats: BF = false reads FP32 weights [27][Ws]; BF = true reads bfloat16 weights
// packed in pairs, word p of element e = bf16(W_2p) | bf16(W_2p+1) << 16 in [14][Ws] (the high
// half of word 13 is zero): 14 loads and half the bytes per element, the arithmetic stays FP32
// (a bf16 value widens exactly to FP32 by a 16-bit shift).
template <int MODE, bool BF>
__global__ __launch_bounds__(256, 2) void k_apply3f(KArgs<float> A, const void* __restrict__ W,
                                                    const float* __restrict__ x, float* __restrict__ y,
                                                    const uint8_t* __restrict__ bc, SmoothArgs<float> S) {
  constexpr int NW = BF ? 14 : 27;  // 32-bit weight words per element
  __shared__ float sb[12][256];
  __shared__ float st[12][256];
  __shared__ uint32_t Wl[NW][256];
  const int tid = threadIdx.x, lx = tid & 15, ly = tid >> 4;
  const __amdgpu_buffer_rsrc_t wr = weight_rsrc((const float*)W, (uint32_t)(NW * A.Ws * 4));
  const uint32_t plane = (uint32_t)(A.Ws * 4);
  for_each_segment(A, [&](int bx, int by, int k0, int k1) {
    const int ei = bx * 15 - 1 + lx, ej = by * 15 - 1 + ly;
    const bool colE = ei >= 0 && ei < A.rx && ej >= 0 && ej < A.ry;
    const bool ownN = lx >= 1 && ly >= 1 && ei <= A.rx && ej <= A.ry;
    // byte offset of element level k in a weight plane, or 0 (a valid address) outside the
    // column's elements
    const auto elem_off = [&](int k) -> uint32_t {
      return (colE && k >= 0 && k < A.rz) ? (uint32_t)(ei + A.rx * (ej + A.ry * k)) * 4u : 0u;
    };
    float u[8][3];
    load_level3(A, x, bc, ei, ej, k0 - 1, u, 0);
    LevelLoad<float> xn;
    xn.issue(A, x, bc, ei, ej, k0);
    uint32_t wn[NW];
    {
      const uint32_t eo = elem_off(k0 - 1);
#pragma unroll
      for (int q = 0; q < NW; q++) wn[q] = ld_buf_u32(wr, eo, q * plane);
    }
    for (int ek = k0 - 1; ek < k1; ++ek) {
      xn.finish(u, 4);
#pragma unroll
      for (int q = 0; q < NW; q++) Wl[q][tid] = wn[q];
      // the owned node's operands first: the memory counter is waited on in issue order, so the
      // epilogue's wait for them does not also wait for the prefetch issued after them
      const bool doN = ownN && ek >= k0;
      const uint32_t nN = doN ? (uint32_t)(ei + A.nx * ej + A.nxy * ek) : 0u;
      NodeOperands<float, MODE> nd;
      nd.load(S, x, bc, nN);
      // prefetch for level ek + 1
      xn.issue(A, x, bc, ei, ej, ek + 2);
      {
        const uint32_t eo = elem_off(ek + 1);
#pragma unroll
        for (int q = 0; q < NW; q++) wn[q] = ld_buf_u32(wr, eo, q * plane);
      }
      float r[8][3];
      if (colE && ek >= 0 && ek < A.rz) {
        elem3_lines<float>(
            [&](int q) {
              if constexpr (BF) {
                const uint32_t v = Wl[q >> 1][tid];
                return __uint_as_float((q & 1) ? (v & 0xffff0000u) : (v << 16));
              } else {
                return __uint_as_float(Wl[q][tid]);
              }
            },
            u, r, A.sx, A.sy, A.sz);
      } else {
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
          for (int i = 0; i < 3; i++) r[a][i] = 0.0f;
      }
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int i = 0; i < 3; i++) sb[c * 3 + i][tid] = r[c][i];
      __syncthreads();
      if (doN) {
#pragma unroll
        for (int i = 0; i < 3; i++) nd.store(S, y, nN, i, node_sum<float>(sb, st, tid, i));
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int i = 0; i < 3; i++) st[c * 3 + i][tid] = r[4 + c][i];
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int i = 0; i < 3; i++) u[c][i] = u[4 + c][i];
    }
  });
}

// bfloat16 rounding (round to nearest, ties to even) of a finite float
__device__ __forceinline__ uint32_t bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// FP64 weights [27][Es] -> bfloat16 pairs [14][Es] (via FP32: the FP64 -> FP32 -> bf16 double
// rounding stays within one bf16 half-unit plus 2^-24 relative)
__global__ void k_pack_w3_bf16(long Es, const double* __restrict__ W, uint32_t* __restrict__ Wb) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= Es) return;
#pragma unroll
  for (int p = 0; p < 14; p++) {
    const uint32_t lo = bf16_rne((float)W[(2 * p) * Es + e]);
    const uint32_t hi = 2 * p + 1 < 27 ? bf16_rne((float)W[(2 * p + 1) * Es + e]) : 0u;
    Wb[p * Es + e] = lo | (hi << 16);
  }
}

void launch_pack_w3_bf16(long Es, const double* W, uint32_t* Wb, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_w3_bf16, nblocks(Es, 256), 256, 0, s, Es, W, Wb);
  UWG_HIP(hipGetLastError());
}

// 2D: 64 element columns per block stream along y
template <typename T>
__device__ __forceinline__ void elem2(const T* __restrict__ W, long E, long e, const T (&u)[4][2], T (&r)[4][2], T sx,
                                      T sy) {
  const T q4 = (T)0.25;
  T cX[2], cY[2], cXY[2];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    T a0 = u[0][i], a1 = u[1][i], a2 = u[2][i], a3 = u[3][i];
    cX[i] = q4 * ((a1 - a0) + (a3 - a2));
    cY[i] = q4 * ((a2 - a0) + (a3 - a1));
    cXY[i] = q4 * ((a3 - a2) - (a1 - a0));
  }
  T dX[2] = {0, 0}, dY[2] = {0, 0}, dXY[2] = {0, 0};
#pragma unroll
  for (int g = 0; g < 9; g++) {
    const T X = (T)gpt(g % 3), Y = (T)gpt(g / 3);
    const T w = W[(long)g * E + e];
    T G[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      G[i][0] = sx * (cX[i] + cXY[i] * Y);
      G[i][1] = sy * (cY[i] + cXY[i] * X);
    }
    const T w2 = w + w;
    T S00 = w2 * G[0][0], S11 = w2 * G[1][1], S01 = w * (G[0][1] + G[1][0]);
    T S[2][2] = {{S00, S01}, {S01, S11}};
#pragma unroll
    for (int i = 0; i < 2; i++) {
      T px = S[i][0] * sx, py = S[i][1] * sy;
      dX[i] += px;
      dY[i] += py;
      dXY[i] += px * Y + py * X;
    }
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const T s = (a & 1) ? 1 : -1, t = (a & 2) ? 1 : -1;
      r[a][i] = q4 * (s * dX[i] + t * dY[i] + (s * t) * dXY[i]);
    }
}

template <typename T>
__device__ __forceinline__ void load_level2(const KArgs<T>& A, const T* __restrict__ x, const uint8_t* __restrict__ bc,
                                            int ei, int lev, T (&u)[4][2], int off) {
  // unconditional loads (absent nodes read node 0), selected afterwards: no load waits on
  // the mask byte
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int ni = ei + c;
    const bool ok = ni >= 0 && ni <= A.rx && lev >= 0 && lev <= A.ry;
    const long n = ok ? ni + A.nx * lev : 0;
    const uint8_t m = bc[n];
    T v[2];
#pragma unroll
    for (int i = 0; i < 2; i++) v[i] = x[n * 2 + i];
#pragma unroll
    for (int i = 0; i < 2; i++) u[off + c][i] = (ok && !((m >> i) & 1)) ? v[i] : (T)0;
  }
}

template <typename T, int MODE>
__global__ __launch_bounds__(64) void k_apply2(KArgs<T> A, const T* __restrict__ W, const T* __restrict__ x,
                                               T* __restrict__ y, const uint8_t* __restrict__ bc, SmoothArgs<T> S) {
  __shared__ T sb[4][64];
  __shared__ T st[4][64];
  const int tid = threadIdx.x;
  const int ei = blockIdx.x * 63 - 1 + tid;
  const int k0 = blockIdx.y * A.KZ;
  const int k1 = min(k0 + A.KZ, A.ry + 1);
  const bool colE = ei >= 0 && ei < A.rx;
  const bool ownN = tid >= 1 && ei <= A.rx;
  T u[4][2];
  load_level2(A, x, bc, ei, k0 - 1, u, 0);
  for (int ek = k0 - 1; ek < k1; ++ek) {
    // the owned node's operands and the next node level are loaded before the element work
    const bool doN = ownN && ek >= k0;
    const uint32_t nN = doN ? (uint32_t)(ei + A.nx * ek) : 0u;
    NodeOperands<T, MODE, 2> nd;
    nd.load(S, x, bc, nN);
    load_level2(A, x, bc, ei, ek + 1, u, 2);
    T r[4][2];
    if (colE && ek >= 0 && ek < A.ry) {
      long e = ei + (long)A.rx * ek;
      elem2<T>(W, A.Ws, e, u, r, A.sx, A.sy);
    } else {
#pragma unroll
      for (int a = 0; a < 4; a++) r[a][0] = r[a][1] = (T)0;
    }
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int i = 0; i < 2; i++) sb[c * 2 + i][tid] = r[c][i];
    __syncthreads();
    if (doN) {
#pragma unroll
      for (int i = 0; i < 2; i++)
        nd.store(S, y, nN, i, st[2 + i][tid - 1] + st[i][tid] + sb[2 + i][tid - 1] + sb[i][tid]);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int i = 0; i < 2; i++) st[c * 2 + i][tid] = r[2 + c][i];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int i = 0; i < 2; i++) u[c][i] = u[2 + c][i];
  }
}

static int g_kz3 = 0, g_kz2 = 32;  // g_kz3 = 0: automatic 3D chunk
void set_apply_chunk(int kz3, int kz2) {
  if (kz3 > 0) g_kz3 = kz3;
  if (kz3 < 0) g_kz3 = 0;
  if (kz2 > 0) g_kz2 = kz2;
}

// compute units of the current device (cached per device)
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  UWG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int n = 0;
    UWG_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    cus[dev] = n > 0 ? n : 256;
  }
  return cus[dev];
}

// z-chunk of the 3D apply: a workgroup streams KZ + 1 element levels (one of them the halo
// level below its chunk), and two workgroups are resident per CU, so a launch takes about
// ceil(workgroups / (2 CUs)) rounds x (KZ + 1) level times.  The chunk minimising that product
// avoids a nearly empty last round (measured at 128^3: 85 us at 11 or 22 levels, 111 us at 16,
// whose 729 workgroups need a second round of 217).  Small grids get one level per workgroup.
static int auto_kz3(long tiles, int rz) {
  const long slots = 2L * device_cus();
  int best = 1;
  long best_cost = -1;
  for (int kz = 1; kz <= rz + 1; kz++) {
    const long nwg = tiles * ((rz + kz) / kz);
    const long cost = ((nwg + slots - 1) / slots) * (kz + 1);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = kz;
    }
  }
  return best;
}

template <typename T, int MODE>
static void launch_apply_mode(const Geom& g, const T* W, const T* x, T* y, const uint8_t* bc, const SmoothArgs<T>& S,
                              hipStream_t s, const uint32_t* Wbf) {
  KArgs<T> A{};
  A.rx = g.res[0];
  A.ry = g.res[1];
  A.rz = g.res[2];
  A.sx = (T)(2.0 / ((g.hi[0] - g.lo[0]) / g.res[0]));
  A.sy = (T)(2.0 / ((g.hi[1] - g.lo[1]) / g.res[1]));
  A.sz = g.dim == 3 ? (T)(2.0 / ((g.hi[2] - g.lo[2]) / g.res[2])) : (T)1;
  A.Ws = g.Es;
  A.nx = g.res[0] + 1;
  A.nxy = A.nx * (g.res[1] + 1);
  if (g.dim == 3) {
    if (27.0 * g.Es * sizeof(T) >= 4294967296.0 || 3.0 * g.N * sizeof(T) >= 4294967296.0)
      throw std::runtime_error("3D K apply: weight planes and node vectors must stay below 4 GiB");
    A.ntx = (g.res[0] + 1 + 14) / 15;
    A.nty = (g.res[1] + 1 + 14) / 15;
    const long tiles = (long)A.ntx * A.nty;
    A.KZ = g_kz3 > 0 ? std::min(g_kz3, g.res[2] + 1) : auto_kz3(tiles, g.res[2]);
    const int nwg = (int)(tiles * ((g.res[2] + A.KZ) / A.KZ));
    if constexpr (sizeof(T) == 4) {
      if (Wbf)
        hipLaunchKernelGGL((k_apply3f<MODE, true>), dim3(nwg), 256, 0, s, A, (const void*)Wbf, x, y, bc, S);
      else
        hipLaunchKernelGGL((k_apply3f<MODE, false>), dim3(nwg), 256, 0, s, A, (const void*)W, x, y, bc, S);
    } else
      hipLaunchKernelGGL((k_apply3<T, MODE>), dim3(nwg), 256, 0, s, A, W, x, y, bc, S);
  } else {
    if (2.0 * g.N * sizeof(T) >= 4294967296.0)
      throw std::runtime_error("2D K apply: node vectors must stay below 4 GiB");
    A.KZ = g_kz2;
    dim3 grid((g.res[0] + 1 + 62) / 63, (g.res[1] + 1 + A.KZ - 1) / A.KZ);
    hipLaunchKernelGGL((k_apply2<T, MODE>), grid, 64, 0, s, A, W, x, y, bc, S);
  }
  UWG_HIP(hipGetLastError());
}

template <typename T>
void launch_apply_K(const Geom& g, const T* W, const T* x, T* y, const uint8_t* bc, hipStream_t s,
                    const uint32_t* Wbf) {
  launch_apply_mode<T, APPLY_Y>(g, W, x, y, bc, SmoothArgs<T>(), s, Wbf);
}
template void launch_apply_K<double>(const Geom&, const double*, const double*, double*, const uint8_t*, hipStream_t,
                                     const uint32_t*);
template void launch_apply_K<float>(const Geom&, const float*, const float*, float*, const uint8_t*, hipStream_t,
                                    const uint32_t*);

// r = b - K x
template <typename T>
void launch_residual_K(const Geom& g, const T* W, const T* x, const T* b, T* r, const uint8_t* bc, hipStream_t s,
                       const uint32_t* Wbf) {
  SmoothArgs<T> S;
  S.b = b;
  launch_apply_mode<T, APPLY_RESIDUAL>(g, W, x, r, bc, S, s, Wbf);
}
template void launch_residual_K<double>(const Geom&, const double*, const double*, const double*, double*,
                                        const uint8_t*, hipStream_t, const uint32_t*);
template void launch_residual_K<float>(const Geom&, const float*, const float*, const float*, float*, const uint8_t*,
                                       hipStream_t, const uint32_t*);

// one Chebyshev step: r = b - K x; d = a d + c dinv r (a d dropped when first); xn = x + d
template <typename T>
void launch_cheb_step_K(const Geom& g, const T* W, const T* x, T* xn, const T* b, const T* dinv, T* d, T a, T c,
                        bool first, const uint8_t* bc, hipStream_t s, const uint32_t* Wbf) {
  SmoothArgs<T> S;
  S.b = b;
  S.dinv = dinv;
  S.d = d;
  S.a = a;
  S.c = c;
  if (first)
    launch_apply_mode<T, APPLY_CHEB0>(g, W, x, xn, bc, S, s, Wbf);
  else
    launch_apply_mode<T, APPLY_CHEB>(g, W, x, xn, bc, S, s, Wbf);
}
template void launch_cheb_step_K<double>(const Geom&, const double*, const double*, double*, const double*,
                                         const double*, double*, double, double, bool, const uint8_t*, hipStream_t,
                                         const uint32_t*);
template void launch_cheb_step_K<float>(const Geom&, const float*, const float*, float*, const float*, const float*,
                                        float*, float, float, bool, const uint8_t*, hipStream_t, const uint32_t*);

// ---------------------------------------------------------------------------------------
// G p (node gather) and G^T u (element gather); G_e constant on a regular mesh
// (_GradientStiffnessMatrixTerm_AssembleElement, GradientStiffnessMatrixTerm.c:129-221)
// ---------------------------------------------------------------------------------------

struct GeConst {
  double v[24];
};

template <typename T, int DIM>
__global__ void k_apply_G(int rx, int ry, int rz, long N, GeConst Ge, const T* __restrict__ p,
                          const uint8_t* __restrict__ bc, T* __restrict__ y) {
  long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  long nx = rx + 1, ny = ry + 1;
  int i = (int)(n % nx), j = (int)((n / nx) % ny), k = (int)(n / (nx * ny));
  T acc[3] = {0, 0, 0};
  for (int oz = 0; oz < (DIM == 3 ? 2 : 1); oz++)
    for (int oy = 0; oy < 2; oy++)
      for (int ox = 0; ox < 2; ox++) {
        int ei = i - 1 + ox, ej = j - 1 + oy, ek = DIM == 3 ? k - 1 + oz : 0;
        if (ei < 0 || ei >= rx || ej < 0 || ej >= ry || ek < 0 || ek >= rz) continue;
        long e = ei + (long)rx * (ej + (long)ry * ek);
        int a = (1 - ox) + 2 * (1 - oy) + (DIM == 3 ? 4 * (1 - oz) : 0);
        T pe = p[e];
        for (int d = 0; d < DIM; d++) acc[d] += (T)Ge.v[a * DIM + d] * pe;
      }
  uint8_t m = bc[n];
  for (int d = 0; d < DIM; d++) y[n * DIM + d] = ((m >> d) & 1) ? (T)0 : acc[d];
}

template <typename T, int DIM>
__global__ void k_apply_GT(int rx, int ry, int rz, long E, GeConst Ge, const T* __restrict__ u,
                           const uint8_t* __restrict__ bc, T* __restrict__ q) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  long nx = rx + 1, ny = ry + 1;
  int i = (int)(e % rx), j = (int)((e / rx) % ry), k = DIM == 3 ? (int)(e / ((long)rx * ry)) : 0;
  long base = i + nx * (j + ny * k);
  constexpr int NN = DIM == 3 ? 8 : 4;
  // every load first (the Dirichlet mask selects afterwards, so no load waits on another)
  uint8_t m[NN];
  T uv[NN][DIM];
#pragma unroll
  for (int a = 0; a < NN; a++) {
    const long n = base + (a & 1) + ((a >> 1) & 1) * nx + (DIM == 3 ? ((a >> 2) & 1) * nx * ny : 0);
    m[a] = bc[n];
#pragma unroll
    for (int d = 0; d < DIM; d++) uv[a][d] = u[n * DIM + d];
  }
  T acc = 0;
#pragma unroll
  for (int a = 0; a < NN; a++)
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const T t = (T)Ge.v[a * DIM + d] * uv[a][d];
      acc = ((m[a] >> d) & 1) ? acc : acc + t;
    }
  q[e] = acc;
}

template <typename T>
void launch_apply_G(const Geom& g, const double* Ge, const T* p, const uint8_t* bc, T* y, hipStream_t s) {
  GeConst G;
  for (int q = 0; q < 24; q++) G.v[q] = q < (1 << g.dim) * g.dim ? Ge[q] : 0.0;
  if (g.dim == 3)
    hipLaunchKernelGGL((k_apply_G<T, 3>), nblocks(g.N, 256), 256, 0, s, g.res[0], g.res[1], g.res[2], g.N, G, p, bc, y);
  else
    hipLaunchKernelGGL((k_apply_G<T, 2>), nblocks(g.N, 256), 256, 0, s, g.res[0], g.res[1], 1, g.N, G, p, bc, y);
  UWG_HIP(hipGetLastError());
}

template <typename T>
void launch_apply_GT(const Geom& g, const double* Ge, const T* u, const uint8_t* bc, T* q, hipStream_t s) {
  GeConst G;
  for (int k = 0; k < 24; k++) G.v[k] = k < (1 << g.dim) * g.dim ? Ge[k] : 0.0;
  if (g.dim == 3)
    hipLaunchKernelGGL((k_apply_GT<T, 3>), nblocks(g.E, 256), 256, 0, s, g.res[0], g.res[1], g.res[2], g.E, G, u, bc, q);
  else
    hipLaunchKernelGGL((k_apply_GT<T, 2>), nblocks(g.E, 256), 256, 0, s, g.res[0], g.res[1], 1, g.E, G, u, bc, q);
  UWG_HIP(hipGetLastError());
}
template void launch_apply_G<double>(const Geom&, const double*, const double*, const uint8_t*, double*, hipStream_t);
template void launch_apply_G<float>(const Geom&, const double*, const float*, const uint8_t*, float*, hipStream_t);
template void launch_apply_GT<double>(const Geom&, const double*, const double*, const uint8_t*, double*, hipStream_t);
template void launch_apply_GT<float>(const Geom&, const double*, const float*, const uint8_t*, float*, hipStream_t);

// ---------------------------------------------------------------------------------------
// Transfer operators: trilinear prolongation (a fine node averages its 1, 2, 4 or 8
// coarse parents, SROpGenerator.c:748-1052) and its transpose.  Dirichlet dofs are
// excluded on both levels.
// ---------------------------------------------------------------------------------------
template <typename T, int DIM>
__global__ void k_prolong_add(int fnx, int fny, int fnz, int cnx, int cny, long Nf, const T* __restrict__ xc,
                              T* __restrict__ xf, const uint8_t* __restrict__ bcf) {
  long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= Nf) return;
  int I = (int)(n % fnx), J = (int)((n / fnx) % fny), K = (int)(n / ((long)fnx * fny));
  T acc[3] = {0, 0, 0};
  int ni = (I & 1) ? 2 : 1, nj = (J & 1) ? 2 : 1, nk = DIM == 3 ? ((K & 1) ? 2 : 1) : 1;
  T wgt = (T)1;
  if (I & 1) wgt *= (T)0.5;
  if (J & 1) wgt *= (T)0.5;
  if (DIM == 3 && (K & 1)) wgt *= (T)0.5;
  for (int c = 0; c < nk; c++)
    for (int b = 0; b < nj; b++)
      for (int a = 0; a < ni; a++) {
        int ci = (I >> 1) + a, cj = (J >> 1) + b, ck = DIM == 3 ? (K >> 1) + c : 0;
        long cn = ci + (long)cnx * (cj + (long)cny * ck);
        for (int d = 0; d < DIM; d++) acc[d] += xc[cn * DIM + d];
      }
  // the mask and the old values load independently of each other; a Dirichlet dof is
  // rewritten with its own value
  const uint8_t m = bcf[n];
  T old[3];
  for (int d = 0; d < DIM; d++) old[d] = xf[n * DIM + d];
  for (int d = 0; d < DIM; d++) xf[n * DIM + d] = ((m >> d) & 1) ? old[d] : old[d] + wgt * acc[d];
}

template <typename T, int DIM>
__global__ void k_restrict(int fnx, int fny, int fnz, int cnx, int cny, int cnz, long Nc, const T* __restrict__ rf,
                           T* __restrict__ rc, const uint8_t* __restrict__ bcc, const uint8_t* __restrict__ bcf) {
  long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= Nc) return;
  int i = (int)(n % cnx), j = (int)((n / cnx) % cny), k = (int)(n / ((long)cnx * cny));
  T acc[3] = {0, 0, 0};
  // every fine value and mask is loaded unconditionally (out-of-grid neighbours read node 0
  // and are skipped by the select), so no load waits on a mask byte
#pragma unroll
  for (int oz = (DIM == 3 ? -1 : 0); oz <= (DIM == 3 ? 1 : 0); oz++) {
    const int K = 2 * k + oz;
#pragma unroll
    for (int oy = -1; oy <= 1; oy++) {
      const int J = 2 * j + oy;
#pragma unroll
      for (int ox = -1; ox <= 1; ox++) {
        const int I = 2 * i + ox;
        const bool in = K >= 0 && K < fnz && J >= 0 && J < fny && I >= 0 && I < fnx;
        T w = (T)1;
        if (ox) w *= (T)0.5;
        if (oy) w *= (T)0.5;
        if (oz) w *= (T)0.5;
        const long fn = in ? I + (long)fnx * (J + (long)fny * K) : 0;
        const uint8_t mf = bcf[fn];
        T v[3];
        for (int d = 0; d < DIM; d++) v[d] = rf[fn * DIM + d];
        for (int d = 0; d < DIM; d++) acc[d] = (in && !((mf >> d) & 1)) ? acc[d] + w * v[d] : acc[d];
      }
    }
  }
  uint8_t m = bcc[n];
  for (int d = 0; d < DIM; d++) rc[n * DIM + d] = ((m >> d) & 1) ? (T)0 : acc[d];
}

template <typename T>
void launch_prolong_add(const Geom& f, const Geom& c, const T* xc, T* xf, const uint8_t* bcf, hipStream_t s) {
  if (f.dim == 3)
    hipLaunchKernelGGL((k_prolong_add<T, 3>), nblocks(f.N, 256), 256, 0, s, f.nn[0], f.nn[1], f.nn[2], c.nn[0], c.nn[1],
                       f.N, xc, xf, bcf);
  else
    hipLaunchKernelGGL((k_prolong_add<T, 2>), nblocks(f.N, 256), 256, 0, s, f.nn[0], f.nn[1], 1, c.nn[0], c.nn[1], f.N,
                       xc, xf, bcf);
  UWG_HIP(hipGetLastError());
}
template void launch_prolong_add<double>(const Geom&, const Geom&, const double*, double*, const uint8_t*, hipStream_t);
template void launch_prolong_add<float>(const Geom&, const Geom&, const float*, float*, const uint8_t*, hipStream_t);

// restriction needs both masks; the fine mask is passed through a global set by the host
template <typename T>
void launch_restrict2(const Geom& f, const Geom& c, const T* rf, T* rc, const uint8_t* bcc, const uint8_t* bcf,
                      hipStream_t s) {
  if (f.dim == 3)
    hipLaunchKernelGGL((k_restrict<T, 3>), nblocks(c.N, 256), 256, 0, s, f.nn[0], f.nn[1], f.nn[2], c.nn[0], c.nn[1],
                       c.nn[2], c.N, rf, rc, bcc, bcf);
  else
    hipLaunchKernelGGL((k_restrict<T, 2>), nblocks(c.N, 256), 256, 0, s, f.nn[0], f.nn[1], 1, c.nn[0], c.nn[1], 1, c.N,
                       rf, rc, bcc, bcf);
  UWG_HIP(hipGetLastError());
}
template void launch_restrict2<double>(const Geom&, const Geom&, const double*, double*, const uint8_t*, const uint8_t*,
                                       hipStream_t);
template void launch_restrict2<float>(const Geom&, const Geom&, const float*, float*, const uint8_t*, const uint8_t*,
                                      hipStream_t);

// Dirichlet mask for walls: wallbits bit (2*d + side) set -> component d fixed on that wall
__global__ void k_bc_walls(int nx, int ny, int nz, int dim, long N, int wallbits, uint8_t* bc) {
  long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  int idx[3] = {(int)(n % nx), (int)((n / nx) % ny), (int)(n / ((long)nx * ny))};
  int nn[3] = {nx, ny, nz};
  uint8_t m = 0;
  for (int d = 0; d < dim; d++) {
    bool lo = idx[d] == 0, hi = idx[d] == nn[d] - 1;
    if ((lo && (wallbits >> (2 * d) & 1)) || (hi && (wallbits >> (2 * d + 1) & 1))) m |= (uint8_t)(1 << d);
  }
  bc[n] = m;
}
void launch_bc_mask_from_walls(const Geom& g, uint8_t* bc, int wallbits, hipStream_t s) {
  hipLaunchKernelGGL(k_bc_walls, nblocks(g.N, 256), 256, 0, s, g.nn[0], g.nn[1], g.nn[2], g.dim, g.N, wallbits, bc);
  UWG_HIP(hipGetLastError());
}

struct Grid3i {
  int nx, ny, nz;
};

// coarse Dirichlet mask by injection: a coarse dof is constrained when its coincident fine
// dof is (nested grids, fine node 2i)
__global__ void k_inject_bc(Grid3i f, Grid3i c, long Nc, const uint8_t* __restrict__ fbc, uint8_t* __restrict__ cbc) {
  long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= Nc) return;
  long i = n % c.nx, j = (n / c.nx) % c.ny, k = n / ((long)c.nx * c.ny);
  cbc[n] = fbc[2 * i + (long)f.nx * (2 * j + (long)f.ny * 2 * k)];
}

void launch_inject_bc(const Geom& f, const Geom& c, const uint8_t* fbc, uint8_t* cbc, hipStream_t s) {
  Grid3i gf{f.nn[0], f.nn[1], f.dim == 3 ? f.nn[2] : 1}, gc{c.nn[0], c.nn[1], c.dim == 3 ? c.nn[2] : 1};
  hipLaunchKernelGGL(k_inject_bc, nblocks(c.N, 256), 256, 0, s, gf, gc, c.N, fbc, cbc);
  UWG_HIP(hipGetLastError());
}

// Dirichlet helpers.  ubc = v on constrained dofs, 0 elsewhere; *flag = 1 if any constrained
// value is nonzero (plain stores of the same value: no ordering needed).
__global__ void k_bc_split(long N, int dim, const uint8_t* __restrict__ bc, const double* __restrict__ v,
                           double* __restrict__ ubc, int* __restrict__ flag) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint8_t m = bc[i];
  for (int d = 0; d < dim; d++) {
    const long k = i * dim + d;
    const bool c = (m >> d) & 1;
    const double x = v[k];
    ubc[k] = c ? x : 0.0;
    if (c && x != 0.0) flag[0] = 1;
  }
}
__global__ void k_zero_bc(long N, int dim, const uint8_t* __restrict__ bc, double* __restrict__ v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint8_t m = bc[i];
  for (int d = 0; d < dim; d++)
    if ((m >> d) & 1) v[i * dim + d] = 0.0;
}
template <typename T>
__global__ void k_neg_recip(long n, const double* __restrict__ s, T* __restrict__ out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = s[i] != 0.0 ? (T)(-1.0 / s[i]) : (T)0;
}

void launch_bc_split(const Geom& g, const uint8_t* bc, const double* v, double* ubc, int* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_bc_split, nblocks(g.N, 256), 256, 0, s, g.N, g.dim, bc, v, ubc, flag);
  UWG_HIP(hipGetLastError());
}
void launch_zero_bc(const Geom& g, const uint8_t* bc, double* v, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_bc, nblocks(g.N, 256), 256, 0, s, g.N, g.dim, bc, v);
  UWG_HIP(hipGetLastError());
}
template <typename T>
void launch_neg_recip(long n, const double* sv, T* out, hipStream_t s) {
  hipLaunchKernelGGL(k_neg_recip<T>, nblocks(n, 256), 256, 0, s, n, sv, out);
  UWG_HIP(hipGetLastError());
}
template void launch_neg_recip<double>(long, const double*, double*, hipStream_t);
template void launch_neg_recip<float>(long, const double*, float*, hipStream_t);

// deterministic positive start vector for power iterations: 0.25 + hash(i) / 2^16 in [0.25, 1.25)
__global__ void k_fill_hash(long n, double* __restrict__ x) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    x[i] = (double)(h & 0xffffu) / 65536.0 + 0.25;
  }
}

void launch_fill_hash(long n, double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_hash, nblocks(n < 2097152 ? n : 2097152, 256), 256, 0, s, n, x);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(long n, T v, T* x) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}
template <typename T>
__global__ void k_copy(long n, const T* x, T* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = x[i];
}
template <typename T>
__global__ void k_axpy(long n, T a, const T* x, T* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] += a * x[i];
}
template <typename T>
__global__ void k_axpby(long n, T a, const T* x, T b, T* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
template <typename T>
__global__ void k_scale(long n, T a, const T* x, T* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = a * x[i];
}
template <typename T>
__global__ void k_xmy(long n, const T* d, const T* x, T* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = d[i] * x[i];
}
template <typename T>
__global__ void k_residual(long n, const T* b, T* r) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) r[i] = b[i] - r[i];
}
// first Chebyshev step from a zero guess: d = c dinv b, x = d
template <typename T>
__global__ void k_cheb_first(long n, T c, const T* dinv, const T* b, T* d, T* x) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const T v = c * (dinv[i] * b[i]);
    d[i] = v;
    x[i] = v;
  }
}
template <typename T>
__global__ void k_cheb(long n, T a, T b, const T* dinv, const T* r, T* d, T* x) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    T v = a * d[i] + b * (dinv[i] * r[i]);
    d[i] = v;
    x[i] += v;
  }
}
__global__ void k_d2f(long n, double s, const double* x, float* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (float)(s * x[i]);
}
__global__ void k_f2d_axpy(long n, double s, const float* x, double* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] += s * (double)x[i];
}
__global__ void k_f2d(long n, double s, const float* x, double* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = s * (double)x[i];
}

static unsigned vgrid(long n) {
  long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

template <typename T> void vk_fill(long n, T v, T* x, hipStream_t s) {
  hipLaunchKernelGGL(k_fill<T>, vgrid(n), 256, 0, s, n, v, x);
}
template <typename T> void vk_copy(long n, const T* x, T* y, hipStream_t s) {
  hipLaunchKernelGGL(k_copy<T>, vgrid(n), 256, 0, s, n, x, y);
}
template <typename T> void vk_axpy(long n, T a, const T* x, T* y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy<T>, vgrid(n), 256, 0, s, n, a, x, y);
}
template <typename T> void vk_axpby(long n, T a, const T* x, T b, T* y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby<T>, vgrid(n), 256, 0, s, n, a, x, b, y);
}
template <typename T> void vk_scale(long n, T a, const T* x, T* y, hipStream_t s) {
  hipLaunchKernelGGL(k_scale<T>, vgrid(n), 256, 0, s, n, a, x, y);
}
template <typename T> void vk_xmy_scale(long n, const T* d, const T* x, T* y, hipStream_t s) {
  hipLaunchKernelGGL(k_xmy<T>, vgrid(n), 256, 0, s, n, d, x, y);
}
template <typename T> void vk_residual(long n, const T* b, T* r, hipStream_t s) {
  hipLaunchKernelGGL(k_residual<T>, vgrid(n), 256, 0, s, n, b, r);
}
template <typename T>
void vk_cheb_update(long n, T a, T b, const T* dinv, const T* r, T* d, T* x, hipStream_t s) {
  hipLaunchKernelGGL(k_cheb<T>, vgrid(n), 256, 0, s, n, a, b, dinv, r, d, x);
}
template <typename T>
void vk_cheb_first(long n, T c, const T* dinv, const T* b, T* d, T* x, hipStream_t s) {
  hipLaunchKernelGGL(k_cheb_first<T>, vgrid(n), 256, 0, s, n, c, dinv, b, d, x);
}
#define INST(T)                                                               \
  template void vk_fill<T>(long, T, T*, hipStream_t);                         \
  template void vk_copy<T>(long, const T*, T*, hipStream_t);                  \
  template void vk_axpy<T>(long, T, const T*, T*, hipStream_t);               \
  template void vk_axpby<T>(long, T, const T*, T, T*, hipStream_t);           \
  template void vk_scale<T>(long, T, const T*, T*, hipStream_t);              \
  template void vk_xmy_scale<T>(long, const T*, const T*, T*, hipStream_t);   \
  template void vk_residual<T>(long, const T*, T*, hipStream_t);              \
  template void vk_cheb_update<T>(long, T, T, const T*, const T*, T*, T*, hipStream_t); \
  template void vk_cheb_first<T>(long, T, const T*, const T*, T*, T*, hipStream_t);
INST(double)
INST(float)
#undef INST

// y = x / sqrt(ss[0]) with the squared norm ss[0] read on the device (no host round trip)
__global__ void k_scale_inv_sqrt(long n, const double* __restrict__ ss, const double* __restrict__ x,
                                 double* __restrict__ y) {
  const double f = 1.0 / sqrt(ss[0]);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = f * x[i];
}
void vk_scale_inv_sqrt(long n, const double* ss, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_inv_sqrt, vgrid(n), 256, 0, s, n, ss, x, y);
}

void vk_d2f(long n, double scale, const double* x, float* y, hipStream_t s) {
  hipLaunchKernelGGL(k_d2f, vgrid(n), 256, 0, s, n, scale, x, y);
}
void vk_f2d_axpy(long n, double scale, const float* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_f2d_axpy, vgrid(n), 256, 0, s, n, scale, x, y);
}
void vk_f2d(long n, double scale, const float* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_f2d, vgrid(n), 256, 0, s, n, scale, x, y);
}

// y (+)= sum_j c[j] X[j]; c and X live in device memory; accumulation in double
template <typename T>
__global__ void k_multi_axpy(long n, int k, const T* const* X, const double* c, T* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double acc = (double)y[i];
    for (int j = 0; j < k; j++) acc += c[j] * (double)X[j][i];
    y[i] = (T)acc;
  }
}
template <typename T>
void vk_multi_axpy(long n, int k, const T* const* X, const double* c, T* y, hipStream_t s) {
  hipLaunchKernelGGL(k_multi_axpy<T>, vgrid(n), 256, 0, s, n, k, X, c, y);
}
template void vk_multi_axpy<double>(long, int, const double* const*, const double*, double*, hipStream_t);
template void vk_multi_axpy<float>(long, int, const float* const*, const double*, float*, hipStream_t);

// deterministic multi-dot: fixed grid of RB blocks; in each block every wave reduces its
// lanes by a fixed butterfly, the four wave sums are added in wave order, and one block then
// sums the RB partials in order.  Products and sums in double for any input precision (an
// FP32 Krylov basis keeps FP64 reductions).  Up to KM = 32 vectors share one pass over x.
#define RB 1024
#define KMAX 32
int vk_reduce_blocks() { return RB; }
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// the block's KM sums, in sm[KM][4] (per wave) then partial[j * RB + block] (thread j)
template <int KM>
__device__ inline void block_sums(double (&acc)[KM], int k, double (*sm)[4], double* partial) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < KM; j++) {
    const double t = wave_sum(acc[j]);
    if (lane == 0) sm[j][wv] = t;
  }
  __syncthreads();
  if (threadIdx.x < KM && threadIdx.x < k) {
    const int j = threadIdx.x;
    partial[j * RB + blockIdx.x] = ((sm[j][0] + sm[j][1]) + sm[j][2]) + sm[j][3];
  }
}
template <int KM, typename TX, typename TY>
__global__ __launch_bounds__(256) void k_mdot(long n, int k, const TX* __restrict__ x, const TY* const* Y,
                                              double* partial) {
  __shared__ double sm[KM][4];
  double acc[KM];
#pragma unroll
  for (int j = 0; j < KM; j++) acc[j] = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double xv = (double)x[i];
#pragma unroll
    for (int j = 0; j < KM; j++)
      if (j < k) acc[j] += xv * (double)Y[j][i];
  }
  block_sums<KM>(acc, k, sm, partial);
}
// k_mdot over element pairs (one 2-wide load per vector and thread: half the load
// instructions); the Gram-Schmidt step's vectors are whole allocations, so the pairs are aligned.
// An odd last element is added by the grid's first thread.
template <typename T> struct Pair;
template <> struct Pair<double> { using V = double2; };
template <> struct Pair<float> { using V = float2; };
template <int KM, typename TX, typename TY>
__global__ __launch_bounds__(256) void k_mdot_pairs(long n, int k, const TX* __restrict__ x, const TY* const* Y,
                                                    double* partial) {
  using XV = typename Pair<TX>::V;
  using YV = typename Pair<TY>::V;
  __shared__ double sm[KM][4];
  double acc[KM];
#pragma unroll
  for (int j = 0; j < KM; j++) acc[j] = 0.0;
  const long n2 = n >> 1;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x) {
    const XV xv = reinterpret_cast<const XV*>(x)[i];
    const double x0 = (double)xv.x, x1 = (double)xv.y;
#pragma unroll
    for (int j = 0; j < KM; j++)
      if (j < k) {
        const YV yv = reinterpret_cast<const YV*>(Y[j])[i];
        acc[j] += x0 * (double)yv.x;
        acc[j] += x1 * (double)yv.y;
      }
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const double xl = (double)x[n - 1];
#pragma unroll
    for (int j = 0; j < KM; j++)
      if (j < k) acc[j] += xl * (double)Y[j][n - 1];
  }
  block_sums<KM>(acc, k, sm, partial);
}
// x.x with the one-vector k_mdot's grid, order and partial layout (bit-identical to it)
template <typename T>
__global__ __launch_bounds__(256) void k_sqnorm(long n, const T* __restrict__ x, double* partial) {
  __shared__ double sm[1][4];
  double acc[1] = {0.0};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double xv = (double)x[i];
    acc[0] += xv * (double)x[i];
  }
  block_sums<1>(acc, 1, sm, partial);
}

// one block per dot product (block j sums the RB partials of dot j in a fixed order)
__global__ void k_mdot_final(const double* partial, double* out) {
  __shared__ double sm[256];
  const int j = blockIdx.x;
  double a = 0.0;
  for (int b = threadIdx.x; b < RB; b += 256) a += partial[j * RB + b];
  sm[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = sm[0];
}
template <typename TX, typename TY>
static void mdot_pass(long n, int k, const TX* x, const TY* const* Y, double* partial, hipStream_t s) {
  if (k <= 1)
    hipLaunchKernelGGL((k_mdot<1, TX, TY>), RB, 256, 0, s, n, k, x, Y, partial);
  else if (k <= 4)
    hipLaunchKernelGGL((k_mdot<4, TX, TY>), RB, 256, 0, s, n, k, x, Y, partial);
  else if (k <= 8)
    hipLaunchKernelGGL((k_mdot<8, TX, TY>), RB, 256, 0, s, n, k, x, Y, partial);
  else if (k <= 16)
    hipLaunchKernelGGL((k_mdot<16, TX, TY>), RB, 256, 0, s, n, k, x, Y, partial);
  else
    hipLaunchKernelGGL((k_mdot<32, TX, TY>), RB, 256, 0, s, n, k, x, Y, partial);
}
template <typename TX, typename TY>
static void mdot_pairs_pass(long n, int k, const TX* x, const TY* const* Y, double* partial, hipStream_t s) {
  if (k <= 8)
    hipLaunchKernelGGL((k_mdot_pairs<8, TX, TY>), RB, 256, 0, s, n, k, x, Y, partial);
  else if (k <= 16)
    hipLaunchKernelGGL((k_mdot_pairs<16, TX, TY>), RB, 256, 0, s, n, k, x, Y, partial);
  else
    hipLaunchKernelGGL((k_mdot_pairs<32, TX, TY>), RB, 256, 0, s, n, k, x, Y, partial);
}
template <typename T>
void vk_multi_dot(long n, int k, const T* x, const T* const* Y, double* partial, double* out, hipStream_t s) {
  if (k < 1) return;
  for (int j0 = 0; j0 < k; j0 += KMAX) {
    const int kk = std::min(KMAX, k - j0);
    mdot_pass<T, T>(n, kk, x, Y + j0, partial, s);
    hipLaunchKernelGGL(k_mdot_final, kk, 256, 0, s, partial, out + j0);
  }
  UWG_HIP(hipGetLastError());
}

// Classical Gram-Schmidt update on the device with the coefficients left in device memory:
// w -= sum_j h[j] V[j], accumulated in double (k_multi_axpy with c = -h), over element pairs
// (2-wide loads and stores; whole allocations, so aligned); an odd last element is done by the
// grid's first thread.
template <typename TW, typename TV>
__global__ void k_gs_update_pairs(long n, int k, const TV* const* V, const double* __restrict__ h,
                                  TW* __restrict__ w) {
  using WV = typename Pair<TW>::V;
  using VV = typename Pair<TV>::V;
  const long n2 = n >> 1;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x) {
    const WV wv = reinterpret_cast<const WV*>(w)[i];
    double a0 = (double)wv.x, a1 = (double)wv.y;
    for (int j = 0; j < k; j++) {
      const VV v = reinterpret_cast<const VV*>(V[j])[i];
      const double c = -h[j];
      a0 += c * (double)v.x;
      a1 += c * (double)v.y;
    }
    WV o;
    o.x = (TW)a0;
    o.y = (TW)a1;
    reinterpret_cast<WV*>(w)[i] = o;
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    double acc = (double)w[n - 1];
    for (int j = 0; j < k; j++) acc += (-h[j]) * (double)V[j][n - 1];
    w[n - 1] = (TW)acc;
  }
}

// v = w / sqrt(max(0, *dot)) (the scale rounded to TW as vk_scale takes it, the product rounded
// to the basis type TV); nothing is written when the norm is zero
template <typename TW, typename TV>
__global__ void k_scale_by_norm(long n, const double* __restrict__ dot, const TW* __restrict__ w,
                                TV* __restrict__ v) {
  const double hn = sqrt(fmax(0.0, *dot));
  if (!(hn > 0.0)) return;
  const TW a = (TW)(1.0 / hn);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[i] = (TV)(a * w[i]);
}

// The update runs on the streaming grid (a reduction-sized grid holds too few loads in flight
// for the k-vector gather); |w|^2 is then the one-vector dot of vk_multi_dot.  V[j] and w must
// be whole allocations (16-byte aligned) for the paired loads of the dot products.
template <typename TW, typename TV>
void vk_gs_step(long n, int k, const TV* const* V, TW* w, TV* v_next, double* h, double* partial, hipStream_t s) {
  for (int j0 = 0; j0 < k; j0 += KMAX) {
    const int kk = std::min(KMAX, k - j0);
    mdot_pairs_pass<TW, TV>(n, kk, w, V + j0, partial, s);
    hipLaunchKernelGGL(k_mdot_final, kk, 256, 0, s, partial, h + j0);
  }
  hipLaunchKernelGGL((k_gs_update_pairs<TW, TV>), vgrid(std::max(n >> 1, 1L)), 256, 0, s, n, k, V, h, w);
  hipLaunchKernelGGL(k_sqnorm<TW>, RB, 256, 0, s, n, w, partial);
  hipLaunchKernelGGL(k_mdot_final, 1, 256, 0, s, partial, h + k);
  hipLaunchKernelGGL((k_scale_by_norm<TW, TV>), vgrid(n), 256, 0, s, n, h + k, w, v_next);
  UWG_HIP(hipGetLastError());
}
template void vk_gs_step<double, double>(long, int, const double* const*, double*, double*, double*, double*,
                                         hipStream_t);
template void vk_gs_step<float, float>(long, int, const float* const*, float*, float*, double*, double*, hipStream_t);
template void vk_gs_step<double, float>(long, int, const float* const*, double*, float*, double*, double*,
                                        hipStream_t);
template void vk_multi_dot<double>(long, int, const double*, const double* const*, double*, double*, hipStream_t);
template void vk_multi_dot<float>(long, int, const float*, const float* const*, double*, double*, hipStream_t);

// dense y = A x (A symmetric, n x n): one block per row, coalesced along the row
__global__ __launch_bounds__(256) void k_dense_mv(int n, const double* A, const double* x, double* y) {
  __shared__ double sm[256];
  const int r = blockIdx.x;
  double acc = 0.0;
  for (int c = threadIdx.x; c < n; c += 256) acc += A[(long)r * n + c] * x[c];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[r] = sm[0];
}
// Dense operator of the coarsest multigrid level, one thread per entry (r, c), r = node_r d + i,
// c = node_c d + j: K[r][c] = sum over the elements holding both nodes (fixed order) and their
// 3^d Gauss points of W_q (delta_ij dN_a . dN_b + dN_a/dx_j dN_b/dx_i), physical derivatives;
// identity rows and columns on Dirichlet dofs.  A gather, so the sums are deterministic.
__global__ void k_coarse_dense(Geom g, const double* __restrict__ W, const uint8_t* __restrict__ bc,
                               double* __restrict__ K) {
  const int dim = g.dim, n = (int)(g.N * dim);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * n) return;
  const int r = (int)(t / n), c = (int)(t % n);
  const int nr = r / dim, i = r % dim, nc = c / dim, j = c % dim;
  if (((bc[nr] >> i) & 1) || ((bc[nc] >> j) & 1)) {
    K[t] = r == c ? 1.0 : 0.0;
    return;
  }
  const int nx = g.nn[0], ny = g.nn[1];
  const int rx = nr % nx, ry = (nr / nx) % ny, rz = dim == 3 ? nr / (nx * ny) : 0;
  const int cx = nc % nx, cy = (nc / nx) % ny, cz = dim == 3 ? nc / (nx * ny) : 0;
  if (abs(rx - cx) > 1 || abs(ry - cy) > 1 || abs(rz - cz) > 1) {
    K[t] = 0.0;
    return;
  }
  double h[3] = {1.0, 1.0, 1.0};
  for (int d = 0; d < dim; d++) h[d] = (g.hi[d] - g.lo[d]) / g.res[d];
  const double ga = sqrt(0.6);
  double acc = 0.0;
  // elements around node_r: lower corner (rx - 1 + ex, ...) for ex, ey, ez in {0, 1}
  for (int ez = 0; ez < (dim == 3 ? 2 : 1); ez++)
    for (int ey = 0; ey < 2; ey++)
      for (int ex = 0; ex < 2; ex++) {
        const int ei = rx - 1 + ex, ej = ry - 1 + ey, ek = dim == 3 ? rz - 1 + ez : 0;
        if (ei < 0 || ei >= g.res[0] || ej < 0 || ej >= g.res[1] || ek < 0 || ek >= (dim == 3 ? g.res[2] : 1)) continue;
        // local corners of node_r and node_c in this element
        const int ax = rx - ei, ay = ry - ej, az = dim == 3 ? rz - ek : 0;
        const int bx = cx - ei, by = cy - ej, bz = dim == 3 ? cz - ek : 0;
        if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
        const long e = ei + (long)g.res[0] * (ej + (long)g.res[1] * ek);
        const double sa[3] = {ax ? 1.0 : -1.0, ay ? 1.0 : -1.0, az ? 1.0 : -1.0};
        const double sb[3] = {bx ? 1.0 : -1.0, by ? 1.0 : -1.0, bz ? 1.0 : -1.0};
        const int Q = dim == 3 ? 27 : 9;
        for (int q = 0; q < Q; q++) {
          const double X[3] = {(q % 3 - 1) * ga, ((q / 3) % 3 - 1) * ga, (q / 9 - 1) * ga};
          double da[3], db[3];
          if (dim == 2) {
            da[0] = 0.25 * sa[0] * (1 + sa[1] * X[1]) * 2.0 / h[0];
            da[1] = 0.25 * sa[1] * (1 + sa[0] * X[0]) * 2.0 / h[1];
            db[0] = 0.25 * sb[0] * (1 + sb[1] * X[1]) * 2.0 / h[0];
            db[1] = 0.25 * sb[1] * (1 + sb[0] * X[0]) * 2.0 / h[1];
          } else {
            da[0] = 0.125 * sa[0] * (1 + sa[1] * X[1]) * (1 + sa[2] * X[2]) * 2.0 / h[0];
            da[1] = 0.125 * sa[1] * (1 + sa[0] * X[0]) * (1 + sa[2] * X[2]) * 2.0 / h[1];
            da[2] = 0.125 * sa[2] * (1 + sa[0] * X[0]) * (1 + sa[1] * X[1]) * 2.0 / h[2];
            db[0] = 0.125 * sb[0] * (1 + sb[1] * X[1]) * (1 + sb[2] * X[2]) * 2.0 / h[0];
            db[1] = 0.125 * sb[1] * (1 + sb[0] * X[0]) * (1 + sb[2] * X[2]) * 2.0 / h[1];
            db[2] = 0.125 * sb[2] * (1 + sb[0] * X[0]) * (1 + sb[1] * X[1]) * 2.0 / h[2];
          }
          double dot = 0.0;
          for (int kk = 0; kk < dim; kk++) dot += da[kk] * db[kk];
          acc += W[(long)q * g.Es + e] * ((i == j ? dot : 0.0) + da[j] * db[i]);
        }
      }
  K[t] = acc;
}

// One Gauss-Jordan step of the inverse of a symmetric positive definite n x n matrix
// (row-major), pivot k, from A into B (two buffers, so no entry is read after another
// workgroup overwrote it): with row = A[k][:] / p, row[k] = 1 / p, col = A[:][k],
//   B[k][j] = row[j],   B[i][j] = (j == k ? 0 : A[i][j]) - col[i] row[j]  (i != k).
// After steps k = 0 .. n-1 the last buffer holds the inverse (no pivoting, which SPD matrices
// do not need).  flag[0] = 1 on a non-positive pivot.
__global__ void k_gj_step(int n, int k, const double* __restrict__ A, double* __restrict__ B, int* __restrict__ flag) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * n) return;
  const double p = A[(long)k * n + k];
  if (t == 0 && !(p > 0.0)) flag[0] = 1;
  const int i = (int)(t / n), j = (int)(t % n);
  const double ip = 1.0 / p;
  const double rj = j == k ? ip : A[(long)k * n + j] * ip;
  B[t] = i == k ? rj : (j == k ? 0.0 : A[t]) - A[(long)i * n + k] * rj;
}

// Ainv (n x n) = inverse of the coarsest level's operator; work: a second n x n buffer
void launch_coarse_inverse(const Geom& g, const double* W, const uint8_t* bc, double* Ainv, double* work, int* flag,
                           hipStream_t s) {
  const long n = g.N * g.dim;
  // n steps alternate between the two buffers: start in the one that ends in Ainv
  double* bufs[2] = {(n % 2) ? work : Ainv, (n % 2) ? Ainv : work};
  hipLaunchKernelGGL(k_coarse_dense, nblocks(n * n, 256), 256, 0, s, g, W, bc, bufs[0]);
  for (int k = 0; k < (int)n; k++)
    hipLaunchKernelGGL(k_gj_step, nblocks(n * n, 256), 256, 0, s, (int)n, k, bufs[k & 1], bufs[(k + 1) & 1], flag);
  UWG_HIP(hipGetLastError());
}

void launch_dense_mv(int n, const double* A, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_dense_mv, n, 256, 0, s, n, A, x, y);
  UWG_HIP(hipGetLastError());
}

}  // namespace uwg
