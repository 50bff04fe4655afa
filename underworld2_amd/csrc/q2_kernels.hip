// q2_kernels.hip -- the Q2/dPc1 element pair (27-node velocity, d+1 discontinuous linear
// pressure values per element) on a regular box mesh.
//
// Layouts (device memory):
//   Q2 vertices        n = I + nx (J + ny K), nx = 2 rx + 1 (C2Generator.c:147-260)
//   element-local      a = ax + 3 ay + 9 az, vertex (2i+ax, 2j+ay, 2k+az)
//   velocity vectors   [N][dim] interleaved
//   pressure vectors   [E][dim+1] (Inner2DGenerator.c:258-283)
//   quadrature weights W [E][NQ], NQ = 5^dim: the element stiffness is
//                      K_e = sum_q W_q B(g_q)^T diag(2,2,2,1,1,1) B(g_q)
//                      at the tensor 5-point Gauss nodes g_q.
// Particle quadrature (ConstitutiveMatrixCartesian.cpp:238-429) integrates products of Q2
// derivatives, polynomials of degree <= 4 per direction, so any point set is reproduced
// exactly by 5^dim weights at the Gauss nodes: W_q = sum_p eta_p w_p detJ L_q(xi_p), with L_q
// the tensor Lagrange polynomials on the Gauss nodes.  The Gauss-swarm path (no voronoi
// swarm) gives W_q = eta(g_q) w_q detJ directly.
//
// The matrix-free apply runs per element with one workgroup of 128 lanes: lanes map to the
// 125 (3D) or 5 x 25 (2D) quadrature points for the gradient/stress pass and to the 81 (3D)
// or 5 x 18 (2D) element dofs for the transposed pass; elements of one parity class
// ("colour") share no vertex, so 2^dim launches accumulate into y without atomics and in a
// fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "uwg_internal.h"

namespace uwg {

namespace {
inline unsigned q2_blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
}  // namespace

// 5-point Gauss-Legendre nodes and weights: +-sqrt(5 -+ 2 sqrt(10/7)) / 3, 0 and
// (322 +- 13 sqrt(70)) / 900, 128 / 225
#define Q2_GA 0.53846931010568309104
#define Q2_GB 0.90617984593866399280
#define Q2_WA 0.47862867049936646804
#define Q2_WB 0.23692688505618908751
#define Q2_W0 0.56888888888888888889

__host__ __device__ __forceinline__ double g5(int m) {
  return m == 0 ? -Q2_GB : (m == 1 ? -Q2_GA : (m == 2 ? 0.0 : (m == 3 ? Q2_GA : Q2_GB)));
}
__host__ __device__ __forceinline__ double w5(int m) {
  return m == 0 || m == 4 ? Q2_WB : (m == 1 || m == 3 ? Q2_WA : Q2_W0);
}
// 1-D quadratic Lagrange basis on {-1, 0, 1} (Biquadratic.c:137-160) and its derivative
__host__ __device__ __forceinline__ double ph(int m, double x) {
  return m == 0 ? 0.5 * x * (x - 1.0) : (m == 1 ? 1.0 - x * x : 0.5 * x * (x + 1.0));
}
__host__ __device__ __forceinline__ double dph(int m, double x) {
  return m == 0 ? x - 0.5 : (m == 1 ? -2.0 * x : x + 0.5);
}
// degree-4 Lagrange polynomial of Gauss node m
__device__ __forceinline__ double lag5(int m, double x) {
  double num = 1.0, den = 1.0;
  const double gm = g5(m);
#pragma unroll
  for (int j = 0; j < 5; j++)
    if (j != m) {
      num *= x - g5(j);
      den *= gm - g5(j);
    }
  return num / den;
}
// quadratic Lagrange polynomial of 3-point Gauss node r (nodes -c, 0, c; c = sqrt(3/5))
__device__ __forceinline__ double lag3(int r, double x) {
  const double c2 = 0.6, c = 0.77459666924148337704;
  return r == 0 ? x * (x - c) / (2.0 * c2) : (r == 1 ? (c2 - x * x) / c2 : x * (x + c) / (2.0 * c2));
}

struct Q2G {
  int rx, ry, rz;  // elements per direction
  int nx, ny, nz;  // vertices per direction (2 r + 1)
  long E, N;
  double s[3];     // 2 / h
  double detJ;
};

static Q2G q2_pack(const Geom& g) {
  Q2G a{};
  a.rx = g.res[0];
  a.ry = g.res[1];
  a.rz = g.dim == 3 ? g.res[2] : 1;
  a.nx = 2 * a.rx + 1;
  a.ny = 2 * a.ry + 1;
  a.nz = g.dim == 3 ? 2 * a.rz + 1 : 1;
  a.E = g.E;
  a.N = (long)a.nx * a.ny * a.nz;
  a.detJ = 1.0;
  for (int d = 0; d < 3; d++) {
    const double h = d < g.dim ? (g.hi[d] - g.lo[d]) / g.res[d] : 2.0;
    a.s[d] = 2.0 / h;
    if (d < g.dim) a.detJ *= 0.5 * h;
  }
  return a;
}

template <int DIM>
__device__ __forceinline__ long q2_vertex(const Q2G& G, long e, int a) {
  const int i = (int)(e % G.rx), j = (int)((e / G.rx) % G.ry), k = DIM == 3 ? (int)(e / ((long)G.rx * G.ry)) : 0;
  const int I = 2 * i + a % 3, J = 2 * j + (a / 3) % 3, K = DIM == 3 ? 2 * k + a / 9 : 0;
  return I + (long)G.nx * (J + (long)G.ny * K);
}

// elements around vertex index I along one direction (r elements): element and local index,
// lower element first
__device__ __forceinline__ int q2_adj(int I, int r, int* el, int* loc) {
  if (I & 1) {
    el[0] = I >> 1;
    loc[0] = 1;
    return 1;
  }
  int c = 0;
  if ((I >> 1) - 1 >= 0) {
    el[c] = (I >> 1) - 1;
    loc[c] = 2;
    c++;
  }
  if ((I >> 1) < r) {
    el[c] = I >> 1;
    loc[c] = 0;
    c++;
  }
  return c;
}

// ---------------------------------------------------------------------------------------
// quadrature weights and element body force
// ---------------------------------------------------------------------------------------
template <int DIM, bool GAUSS>
__global__ void k_q2_weights(Q2G G, const int* __restrict__ cell_ptr, const int* __restrict__ cell_pid,
                             const double* __restrict__ xi_slot, const double* __restrict__ w_slot,
                             const int* __restrict__ mat, int nmat, const double* __restrict__ eta_tab,
                             const double* __restrict__ force_tab, const double* __restrict__ eta_p,
                             const double* __restrict__ force_p, const double* __restrict__ eta_g,
                             const double* __restrict__ force_g, double* __restrict__ W, double* __restrict__ fe) {
  constexpr int NQ = DIM == 3 ? 125 : 25, NEN = DIM == 3 ? 27 : 9;
  const long e = blockIdx.x;
  const int t = threadIdx.x;
  const int qx = t % 5, qy = (t / 5) % 5, qz = DIM == 3 ? t / 25 : 0;
  const int fa = t / DIM, fi = t % DIM;
  const int ax = fa % 3, ay = (fa / 3) % 3, az = DIM == 3 ? fa / 9 : 0;
  double wacc = 0.0, facc = 0.0;
  if (GAUSS) {
    if (t < NQ) {
      const double wq = w5(qx) * w5(qy) * (DIM == 3 ? w5(qz) : 1.0);
      wacc = eta_g[e * NQ + t] * (wq * G.detJ);
    }
    if (t < NEN * DIM)
      for (int q = 0; q < NQ; q++) {
        const int px = q % 5, py = (q / 5) % 5, pz = DIM == 3 ? q / 25 : 0;
        const double wq = w5(px) * w5(py) * (DIM == 3 ? w5(pz) : 1.0);
        const double Na = ph(ax, g5(px)) * ph(ay, g5(py)) * (DIM == 3 ? ph(az, g5(pz)) : 1.0);
        facc += (wq * G.detJ) * force_g[(e * NQ + q) * DIM + fi] * Na;
      }
  } else {
    const int s0 = cell_ptr[e], s1 = cell_ptr[e + 1];
    for (int s = s0; s < s1; s++) {
      const int p = cell_pid[s];
      double xi[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < DIM; d++) xi[d] = xi_slot[(long)s * DIM + d];
      const double w = w_slot[s] * G.detJ;
      double eta, f;
      if (mat) {
        int m = mat[p];
        if (m < 0 || m >= nmat) {  // index outside the table: poison the element
          m = 0;
          eta = __builtin_nan("");
        } else {
          eta = eta_tab[m];
        }
        f = force_tab[m * DIM + fi];
      } else {
        eta = eta_p[p];
        f = force_p[(long)p * DIM + fi];
      }
      if (t < NQ) wacc += eta * w * (lag5(qx, xi[0]) * lag5(qy, xi[1]) * (DIM == 3 ? lag5(qz, xi[2]) : 1.0));
      if (t < NEN * DIM) facc += w * f * (ph(ax, xi[0]) * ph(ay, xi[1]) * (DIM == 3 ? ph(az, xi[2]) : 1.0));
    }
  }
  if (t < NQ) W[e * NQ + t] = wacc;
  if (t < NEN * DIM) fe[e * NEN * DIM + t] = facc;
}

void launch_q2_weights(const Geom& g, const int* cell_ptr, const int* cell_pid, const double* xi_slot,
                       const double* w_slot, const int* mat, int nmat, const double* eta_tab, const double* force_tab,
                       const double* eta_p, const double* force_p, const double* eta_g, const double* force_g,
                       double* W, double* fe, hipStream_t s) {
  const Q2G G = q2_pack(g);
  const bool gauss = eta_g != nullptr;
  if (g.dim == 3) {
    if (gauss)
      hipLaunchKernelGGL((k_q2_weights<3, true>), dim3((unsigned)g.E), 128, 0, s, G, cell_ptr, cell_pid, xi_slot,
                         w_slot, mat, nmat, eta_tab, force_tab, eta_p, force_p, eta_g, force_g, W, fe);
    else
      hipLaunchKernelGGL((k_q2_weights<3, false>), dim3((unsigned)g.E), 128, 0, s, G, cell_ptr, cell_pid, xi_slot,
                         w_slot, mat, nmat, eta_tab, force_tab, eta_p, force_p, eta_g, force_g, W, fe);
  } else {
    if (gauss)
      hipLaunchKernelGGL((k_q2_weights<2, true>), dim3((unsigned)g.E), 64, 0, s, G, cell_ptr, cell_pid, xi_slot,
                         w_slot, mat, nmat, eta_tab, force_tab, eta_p, force_p, eta_g, force_g, W, fe);
    else
      hipLaunchKernelGGL((k_q2_weights<2, false>), dim3((unsigned)g.E), 64, 0, s, G, cell_ptr, cell_pid, xi_slot,
                         w_slot, mat, nmat, eta_tab, force_tab, eta_p, force_p, eta_g, force_g, W, fe);
  }
  UWG_HIP(hipGetLastError());
}

// f[n][i] = sum over the elements around vertex n of fe (fixed order); 0 on Dirichlet dofs
template <int DIM>
__global__ void k_q2_gather_f(Q2G G, const double* __restrict__ fe, const uint8_t* __restrict__ bc,
                              double* __restrict__ f) {
  constexpr int NEN = DIM == 3 ? 27 : 9;
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  const int I = (int)(n % G.nx), J = (int)((n / G.nx) % G.ny), K = DIM == 3 ? (int)(n / ((long)G.nx * G.ny)) : 0;
  int ex[2], lx[2], ey[2], ly[2], ez[2] = {0, 0}, lz[2] = {0, 0};
  const int cx = q2_adj(I, G.rx, ex, lx), cy = q2_adj(J, G.ry, ey, ly), cz = DIM == 3 ? q2_adj(K, G.rz, ez, lz) : 1;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < cz; c++)
    for (int b = 0; b < cy; b++)
      for (int a = 0; a < cx; a++) {
        const long e = ex[a] + (long)G.rx * (ey[b] + (long)G.ry * ez[c]);
        const int loc = lx[a] + 3 * ly[b] + 9 * lz[c];
        for (int d = 0; d < DIM; d++) acc[d] += fe[(e * NEN + loc) * DIM + d];
      }
  const uint8_t m = bc[n];
  for (int d = 0; d < DIM; d++) f[n * DIM + d] = ((m >> d) & 1) ? 0.0 : acc[d];
}

void launch_q2_gather_f(const Geom& g, const double* fe, const uint8_t* bc, double* f, hipStream_t s) {
  const Q2G G = q2_pack(g);
  if (g.dim == 3)
    hipLaunchKernelGGL(k_q2_gather_f<3>, q2_blocks(G.N, 256), 256, 0, s, G, fe, bc, f);
  else
    hipLaunchKernelGGL(k_q2_gather_f<2>, q2_blocks(G.N, 256), 256, 0, s, G, fe, bc, f);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// 3D K apply on the matrix cores.  Both contractions of the element operator are GEMMs with
// one constant matrix D[a][(q,j)] = dN_a/dx_j(g_q) (27 x 375, padded to 32 x 384):
//   G = X D      rows: (element, component), 5 elements per workgroup -> 15 (+1) rows
//   T = W (G + G^T) pointwise per quadrature point (in LDS)
//   Y = T D^T    scattered to the element's vertices (colour launches, no atomics)
// with v_mfma_{f32,f64}_16x16x4.  Operand layouts (checked on the device): A[i][k] in lane
// i + 16 k, B[k][j] in lane j + 16 k; the f32 result row of lane l, value v is 4 (l/16) + v,
// the f64 result row is l/16 + 4 v; the column is l % 16.
// ---------------------------------------------------------------------------------------
constexpr int Q2M_E = 5;      // elements per workgroup
constexpr int Q2M_K = 28;     // 27 vertices padded to a multiple of 4
constexpr int Q2M_C = 384;    // 125 points x 3 derivatives padded to 24 tiles of 16
constexpr int Q2M_CS = 388;   // LDS row stride (spreads the 16 rows of a column over the banks)

template <typename T> struct Q2Mma;
template <> struct Q2Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int l, int v) { return 4 * (l >> 4) + v; }
};
template <> struct Q2Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int l, int v) { return (l >> 4) + 4 * v; }
};

// REG: the wave's slice of D is loaded into registers once and the block loops over element
// batches (large grids, FP32); otherwise every MFMA reads its D operand through the cache
// (FP64, whose register budget the slice would exhaust, and small grids)
template <typename T, bool REG>
__global__ __launch_bounds__(256) void k_q2_apply_mfma(Q2G G, int cx, int cy, int cz, int ncx, int ncy, long nce,
                                                       const T* __restrict__ W, const T* __restrict__ Dm,
                                                       const T* __restrict__ x, T* __restrict__ y,
                                                       const uint8_t* __restrict__ bc) {
  using M = Q2Mma<T>;
  using v4 = typename M::v4;
  __shared__ T Xs[16][Q2M_K];
  __shared__ T Gs[16][Q2M_CS];
  T(*Red)[16][32] = reinterpret_cast<T(*)[16][32]>(&Gs[0][0]);  // reuses Gs once pass 4 has read it
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  auto elem_of = [&](long ce) -> long {
    const int i = cx + 2 * (int)(ce % ncx), j = cy + 2 * (int)((ce / ncx) % ncy);
    const int k = cz + 2 * (int)(ce / ((long)ncx * ncy));
    return i + (long)G.rx * (j + (long)G.ry * k);
  };
  // this wave's B operands of both passes, held in registers for every batch the block does
  auto b1 = [&](int c, int ks) -> T { return Dm[(4 * ks + (l >> 4)) * Q2M_C + (6 * wv + c) * 16 + (l & 15)]; };
  auto b2 = [&](int ks, int n) -> T { return Dm[(16 * n + (l & 15)) * Q2M_C + 4 * (24 * wv + ks) + (l >> 4)]; };
  T B1[REG ? 6 : 1][REG ? Q2M_K / 4 : 1], B2[REG ? 24 : 1][2];
  if constexpr (REG) {
#pragma unroll
    for (int c = 0; c < 6; c++)
#pragma unroll
      for (int ks = 0; ks < Q2M_K / 4; ks++) B1[c][ks] = b1(c, ks);
#pragma unroll
    for (int ks = 0; ks < 24; ks++) {
      B2[ks][0] = b2(ks, 0);
      B2[ks][1] = b2(ks, 1);
    }
  }
  const long nbatch = (nce + Q2M_E - 1) / Q2M_E;
  for (long bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const long ce0 = bt * Q2M_E;
    // 1. element dofs, rows r = 3 le + i (Dirichlet dofs read as zero; padding zero)
    for (int idx = t; idx < 16 * Q2M_K; idx += 256) {
      const int r = idx / Q2M_K, a = idx % Q2M_K, le = r / 3, i = r % 3;
      T v = (T)0;
      if (r < 3 * Q2M_E && a < 27 && ce0 + le < nce) {
        const long n = q2_vertex<3>(G, elem_of(ce0 + le), a);
        const uint8_t m = bc[n];
        const T xv = x[n * 3 + i];  // loaded unconditionally: the mask does not gate the load
        v = ((m >> i) & 1) ? (T)0 : xv;
      }
      Xs[r][a] = v;
    }
    __syncthreads();
    // 2. G = X D: wave wv computes column tiles 6 wv .. 6 wv + 5
    if constexpr (REG) {
#pragma unroll
      for (int c = 0; c < 6; c++) {
        v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < Q2M_K / 4; ks++) acc = M::mma(Xs[l & 15][4 * ks + (l >> 4)], B1[c][ks], acc);
#pragma unroll
        for (int v = 0; v < 4; v++) Gs[M::row(l, v)][(6 * wv + c) * 16 + (l & 15)] = acc[v];
      }
    } else {
#pragma unroll 1
      for (int c = 0; c < 6; c++) {
        v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < Q2M_K / 4; ks++) acc = M::mma(Xs[l & 15][4 * ks + (l >> 4)], b1(c, ks), acc);
#pragma unroll
        for (int v = 0; v < 4; v++) Gs[M::row(l, v)][(6 * wv + c) * 16 + (l & 15)] = acc[v];
      }
    }
    __syncthreads();
    // 3. tau_ij = W_q (g_ij + g_ji) in place
    for (int idx = t; idx < Q2M_E * 125; idx += 256) {
      const int le = idx / 125, q = idx % 125;
      if (ce0 + le >= nce) continue;
      const T wq = W[elem_of(ce0 + le) * 125 + q];
      T g[3][3];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) g[i][j] = Gs[3 * le + i][q * 3 + j];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Gs[3 * le + i][q * 3 + j] = wq * (g[i][j] + g[j][i]);
    }
    __syncthreads();
    // 4. Y = T D^T over the 384 columns, split into four quarters (one per wave)
    v4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
    if constexpr (REG) {
#pragma unroll
      for (int ks = 0; ks < 24; ks++) {
        const T av = Gs[l & 15][4 * (24 * wv + ks) + (l >> 4)];
        acc0 = M::mma(av, B2[ks][0], acc0);
        acc1 = M::mma(av, B2[ks][1], acc1);
      }
    } else {
#pragma unroll 2
      for (int ks = 0; ks < 24; ks++) {
        const T av = Gs[l & 15][4 * (24 * wv + ks) + (l >> 4)];
        acc0 = M::mma(av, b2(ks, 0), acc0);
        acc1 = M::mma(av, b2(ks, 1), acc1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 4; v++) {
      Red[wv][M::row(l, v)][l & 15] = acc0[v];
      Red[wv][M::row(l, v)][16 + (l & 15)] = acc1[v];
    }
    __syncthreads();
    // 5. sum the quarters, scatter to the free vertex dofs
    for (int idx = t; idx < 16 * 32; idx += 256) {
      const int r = idx / 32, a = idx % 32, le = r / 3, i = r % 3;
      if (r >= 3 * Q2M_E || a >= 27 || ce0 + le >= nce) continue;
      const T sacc = Red[0][r][a] + Red[1][r][a] + Red[2][r][a] + Red[3][r][a];
      const long n = q2_vertex<3>(G, elem_of(ce0 + le), a);
      if (!((bc[n] >> i) & 1)) y[n * 3 + i] += sacc;
    }
    __syncthreads();  // Xs / Gs are refilled by the next batch
  }
}

// D[a][(q,j)] for the matrix-core apply, [32][384] (zero padding), in T
void q2_mfma_matrix(const Geom& g, std::vector<double>& D) {
  const Q2G G = q2_pack(g);
  D.assign((size_t)32 * Q2M_C, 0.0);
  for (int a = 0; a < 27; a++) {
    const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
    for (int q = 0; q < 125; q++) {
      const double X = g5(q % 5), Y = g5((q / 5) % 5), Z = g5(q / 25);
      D[(size_t)a * Q2M_C + q * 3 + 0] = dph(ax, X) * G.s[0] * ph(ay, Y) * ph(az, Z);
      D[(size_t)a * Q2M_C + q * 3 + 1] = ph(ax, X) * dph(ay, Y) * G.s[1] * ph(az, Z);
      D[(size_t)a * Q2M_C + q * 3 + 2] = ph(ax, X) * ph(ay, Y) * dph(az, Z) * G.s[2];
    }
  }
}

// ---------------------------------------------------------------------------------------
// Sum-factorised 3D apply.  On the tensor 3 x 3 x 3 element and the 5 x 5 x 5 Gauss rule every
// shape-function derivative factorises into 1D tables, B[m][a] = phi_a(g_m) and
// D[m][a] = phi_a'(g_m) (times 2/h on its axis), so the gradient at the 125 points is three
// passes of 1D contractions (x, then y, then z; 2,070 multiply-adds per velocity component
// instead of 27 x 125 x 3) and the transposed contraction the same passes in reverse: 12,420
// multiply-adds, about 27 kflop per element with the pointwise stress, against 121.5 kflop for
// the dense contraction.  Five elements per 128-lane workgroup; the x and y passes and their
// transposes are spread over the lanes through LDS, each lane (element, qx, qy) runs the
// z pass, the stress at its five points and the transposed z pass in registers.  Elements of
// one parity class share no vertex, so the colour launches add into y without atomics.
// ---------------------------------------------------------------------------------------
constexpr int Q2S_E = 5;

template <typename T>
__global__ __launch_bounds__(128) void k_q2_apply_sf(Q2G G, int cx, int cy, int cz, int ncx, int ncy, long nce,
                                                     const T* __restrict__ W, const T* __restrict__ x,
                                                     T* __restrict__ y, const uint8_t* __restrict__ bc) {
  __shared__ T xs[Q2S_E][81];        // u[(ax + 3 ay + 9 az) * 3 + i]
  __shared__ T U[Q2S_E][2][5][27];   // x pass [B|D][qx][(ay + 3 az) * 3 + i]; reused for the y^T pass
  __shared__ T A[Q2S_E][25][27];     // after the z^T pass, per (qx, qy): [path * 9 + az * 3 + i]
  __shared__ T tb[3][5][3];          // B, D * (2/h_x), D * (2/h_y); D * (2/h_z) is formed in the z pass
  const int t = threadIdx.x;
  const long ce0 = (long)blockIdx.x * Q2S_E;
  auto elem_of = [&](long ce) -> long {
    const int i = cx + 2 * (int)(ce % ncx), j = cy + 2 * (int)((ce / ncx) % ncy);
    const int k = cz + 2 * (int)(ce / ((long)ncx * ncy));
    return i + (long)G.rx * (j + (long)G.ry * k);
  };
  if (t < 15) {
    const int m = t / 3, a = t % 3;
    tb[0][m][a] = (T)ph(a, g5(m));
    tb[1][m][a] = (T)dph(a, g5(m)) * (T)G.s[0];
    tb[2][m][a] = (T)dph(a, g5(m)) * (T)G.s[1];
  }
  // 1. element dofs (Dirichlet dofs read as zero)
  for (int idx = t; idx < Q2S_E * 81; idx += 128) {
    const int le = idx / 81, r = idx % 81;
    T v = (T)0;
    if (ce0 + le < nce) {
      const long n = q2_vertex<3>(G, elem_of(ce0 + le), r / 3);
      const int i = r % 3;
      const uint8_t m = bc[n];
      const T xv = x[n * 3 + i];  // loaded unconditionally: the mask does not gate the load
      v = ((m >> i) & 1) ? (T)0 : xv;
    }
    xs[le][r] = v;
  }
  __syncthreads();
  // 2. x pass: U[v][qx][L*3+i] = sum_ax (B|D)[qx][ax] u[(ax + 3 L) * 3 + i], L = ay + 3 az
  for (int idx = t; idx < Q2S_E * 270; idx += 128) {
    const int le = idx / 270, r = idx % 270, v = r / 135, qx = (r % 135) / 27, ln = r % 27;
    const int L = ln / 3, i = ln % 3;
    T acc = (T)0;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) acc += tb[v][qx][ax] * xs[le][(ax + 3 * L) * 3 + i];
    U[le][v][qx][ln] = acc;
  }
  __syncthreads();
  // 3. per (element, qx, qy): y pass, z pass, stress at the five points, z^T pass
  {
    const int le = t / 25, qq = t % 25, qx = qq % 5, qy = qq / 5;
    if (t < 125 && ce0 + le < nce) {
      const long e = elem_of(ce0 + le);
      T Vbb[3][3], Vdb[3][3], Vbd[3][3];  // [az][i]: z-derivative, x-derivative, y-derivative paths
#pragma unroll
      for (int az = 0; az < 3; az++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          T bb = (T)0, db = (T)0, bd = (T)0;
#pragma unroll
          for (int ay = 0; ay < 3; ay++) {
            const int ln = (ay + 3 * az) * 3 + i;
            bb += tb[0][qy][ay] * U[le][0][qx][ln];
            db += tb[0][qy][ay] * U[le][1][qx][ln];
            bd += tb[2][qy][ay] * U[le][0][qx][ln];
          }
          Vbb[az][i] = bb;
          Vdb[az][i] = db;
          Vbd[az][i] = bd;
        }
      T A1[3][3], A2[3][3], A3[3][3];
#pragma unroll
      for (int az = 0; az < 3; az++)
#pragma unroll
        for (int i = 0; i < 3; i++) A1[az][i] = A2[az][i] = A3[az][i] = (T)0;
#pragma unroll
      for (int qz = 0; qz < 5; qz++) {
        T bz[3], dz[3];
#pragma unroll
        for (int az = 0; az < 3; az++) {
          bz[az] = tb[0][qz][az];
          dz[az] = (T)dph(az, g5(qz)) * (T)G.s[2];
        }
        T g[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
          g[i][0] = bz[0] * Vdb[0][i] + bz[1] * Vdb[1][i] + bz[2] * Vdb[2][i];
          g[i][1] = bz[0] * Vbd[0][i] + bz[1] * Vbd[1][i] + bz[2] * Vbd[2][i];
          g[i][2] = dz[0] * Vbb[0][i] + dz[1] * Vbb[1][i] + dz[2] * Vbb[2][i];
        }
        const T w = W[e * 125 + qx + 5 * qy + 25 * qz];
        T tau[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++) tau[i][j] = w * (g[i][j] + g[j][i]);
#pragma unroll
        for (int az = 0; az < 3; az++)
#pragma unroll
          for (int i = 0; i < 3; i++) {
            A1[az][i] += bz[az] * tau[i][0];
            A2[az][i] += bz[az] * tau[i][1];
            A3[az][i] += dz[az] * tau[i][2];
          }
      }
#pragma unroll
      for (int az = 0; az < 3; az++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          A[le][qq][0 * 9 + az * 3 + i] = A1[az][i];
          A[le][qq][1 * 9 + az * 3 + i] = A2[az][i];
          A[le][qq][2 * 9 + az * 3 + i] = A3[az][i];
        }
    }
  }
  __syncthreads();
  // 4. y^T pass: C1[qx][L*3+i] = sum_qy B[qy][ay] A1, C2 = sum_qy (D[qy][ay] A2 + B[qy][ay] A3)
  for (int idx = t; idx < Q2S_E * 270; idx += 128) {
    const int le = idx / 270, r = idx % 270, v = r / 135, qx = (r % 135) / 27, ln = r % 27;
    const int L = ln / 3, i = ln % 3, ay = L % 3, az = L / 3;
    T acc = (T)0;
#pragma unroll
    for (int qy = 0; qy < 5; qy++) {
      const T* a = A[le][qx + 5 * qy];
      if (v == 0)
        acc += tb[0][qy][ay] * a[0 * 9 + az * 3 + i];
      else
        acc += tb[2][qy][ay] * a[1 * 9 + az * 3 + i] + tb[0][qy][ay] * a[2 * 9 + az * 3 + i];
    }
    U[le][v][qx][ln] = acc;
  }
  __syncthreads();
  // 5. x^T pass and scatter: y_ai += sum_qx D[qx][ax] C1 + B[qx][ax] C2
  for (int idx = t; idx < Q2S_E * 81; idx += 128) {
    const int le = idx / 81, r = idx % 81;
    if (ce0 + le >= nce) continue;
    const int a = r / 3, i = r % 3, ax = a % 3, L = a / 3;
    T acc = (T)0;
#pragma unroll
    for (int qx = 0; qx < 5; qx++)
      acc += tb[1][qx][ax] * U[le][0][qx][L * 3 + i] + tb[0][qx][ax] * U[le][1][qx][L * 3 + i];
    const long n = q2_vertex<3>(G, elem_of(ce0 + le), a);
    if (!((bc[n] >> i) & 1)) y[n * 3 + i] += acc;
  }
}

// 3D apply kernel: 2 sum-factorised (default), 1 matrix cores, 0 dense vector kernel
static int g_q2_kernel = 2;
void set_q2_mfma(int on) { g_q2_kernel = on ? 1 : 0; }
void set_q2_apply_kernel(int kind) { g_q2_kernel = kind; }

// ---------------------------------------------------------------------------------------
// matrix-free K apply
// ---------------------------------------------------------------------------------------
// y = (bc ? x : 0): identity rows on Dirichlet dofs, the element passes add the free rows
template <typename T, int DIM>
__global__ void k_q2_init(long N, const T* __restrict__ x, const uint8_t* __restrict__ bc, T* __restrict__ y) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const uint8_t m = bc[n];
  for (int d = 0; d < DIM; d++) y[n * DIM + d] = ((m >> d) & 1) ? x[n * DIM + d] : (T)0;
}

template <typename T, int DIM>
__global__ __launch_bounds__(128) void k_q2_apply(Q2G G, int cx, int cy, int cz, int ncx, int ncy, long nce,
                                                  const T* __restrict__ W, const T* __restrict__ x,
                                                  T* __restrict__ y, const uint8_t* __restrict__ bc) {
  constexpr int NQ = DIM == 3 ? 125 : 25, NEN = DIM == 3 ? 27 : 9, NV = NEN * DIM;
  constexpr int EPB = DIM == 3 ? 1 : 5, NS = DIM == 3 ? 6 : 3;
  __shared__ T xs[EPB][NV];
  __shared__ T tau[EPB][NQ][NS];
  const int t = threadIdx.x;
  const long ce0 = (long)blockIdx.x * EPB;
  auto elem_of = [&](long ce) -> long {
    const int i = cx + 2 * (int)(ce % ncx), j = cy + 2 * (int)((ce / ncx) % ncy);
    const int k = DIM == 3 ? cz + 2 * (int)(ce / ((long)ncx * ncy)) : 0;
    return i + (long)G.rx * (j + (long)G.ry * k);
  };
  // 1. gather the element dofs (Dirichlet dofs read as zero)
  for (int idx = t; idx < EPB * NV; idx += blockDim.x) {
    const int le = idx / NV, r = idx % NV;
    const long ce = ce0 + le;
    T v = (T)0;
    if (ce < nce) {
      const long n = q2_vertex<DIM>(G, elem_of(ce), r / DIM);
      const int i = r % DIM;
      const uint8_t m = bc[n];
      const T xv = x[n * DIM + i];  // loaded unconditionally: the mask does not gate the load
      v = ((m >> i) & 1) ? (T)0 : xv;
    }
    xs[le][r] = v;
  }
  __syncthreads();
  // 2. per quadrature point: gradient, tau = W (grad u + grad u^T)
  {
    const int le = t / NQ, q = t % NQ;
    const long ce = ce0 + le;
    if (le < EPB && ce < nce) {
      const int qx = q % 5, qy = (q / 5) % 5, qz = DIM == 3 ? q / 25 : 0;
      const T X = (T)g5(qx), Y = (T)g5(qy), Z = (T)g5(qz);
      T px[3], py[3], pz[3], dx[3], dy[3], dz[3];
#pragma unroll
      for (int m = 0; m < 3; m++) {
        px[m] = (T)ph(m, X);
        py[m] = (T)ph(m, Y);
        pz[m] = DIM == 3 ? (T)ph(m, Z) : (T)1;
        dx[m] = (T)dph(m, X) * (T)G.s[0];
        dy[m] = (T)dph(m, Y) * (T)G.s[1];
        dz[m] = (T)dph(m, Z) * (T)G.s[2];
      }
      T gr[3][3] = {};
#pragma unroll
      for (int a = 0; a < NEN; a++) {
        const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
        const T n0 = dx[ax] * py[ay] * pz[az];
        const T n1 = px[ax] * dy[ay] * pz[az];
        const T n2 = DIM == 3 ? px[ax] * py[ay] * dz[az] : (T)0;
#pragma unroll
        for (int i = 0; i < DIM; i++) {
          const T u = xs[le][a * DIM + i];
          gr[i][0] += u * n0;
          gr[i][1] += u * n1;
          if (DIM == 3) gr[i][2] += u * n2;
        }
      }
      const T w = W[elem_of(ce) * NQ + q];
      if (DIM == 3) {
        tau[le][q][0] = w * (gr[0][0] + gr[0][0]);
        tau[le][q][1] = w * (gr[1][1] + gr[1][1]);
        tau[le][q][2] = w * (gr[2][2] + gr[2][2]);
        tau[le][q][3] = w * (gr[0][1] + gr[1][0]);
        tau[le][q][4] = w * (gr[0][2] + gr[2][0]);
        tau[le][q][5] = w * (gr[1][2] + gr[2][1]);
      } else {
        tau[le][q][0] = w * (gr[0][0] + gr[0][0]);
        tau[le][q][1] = w * (gr[1][1] + gr[1][1]);
        tau[le][q][2] = w * (gr[0][1] + gr[1][0]);
      }
    }
  }
  __syncthreads();
  // 3. y_ai += sum_q sum_j dN_a/dx_j(g_q) tau_ij(q)
  for (int idx = t; idx < EPB * NV; idx += blockDim.x) {
    const int le = idx / NV, r = idx % NV;
    const long ce = ce0 + le;
    if (ce >= nce) continue;
    const int a = r / DIM, i = r % DIM;
    const int ax = a % 3, ay = (a / 3) % 3, az = DIM == 3 ? a / 9 : 0;
    T fx[5], fy[5], fz[5], gx[5], gy[5], gz[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const double gm = g5(m);
      fx[m] = (T)ph(ax, gm);
      fy[m] = (T)ph(ay, gm);
      fz[m] = DIM == 3 ? (T)ph(az, gm) : (T)1;
      gx[m] = (T)dph(ax, gm) * (T)G.s[0];
      gy[m] = (T)dph(ay, gm) * (T)G.s[1];
      gz[m] = (T)dph(az, gm) * (T)G.s[2];
    }
    // row i of the symmetric tau: (i,0), (i,1), (i,2)
    int c0, c1, c2;
    if (DIM == 3) {
      c0 = i == 0 ? 0 : (i == 1 ? 3 : 4);
      c1 = i == 0 ? 3 : (i == 1 ? 1 : 5);
      c2 = i == 0 ? 4 : (i == 1 ? 5 : 2);
    } else {
      c0 = i == 0 ? 0 : 2;
      c1 = i == 0 ? 2 : 1;
      c2 = 0;
    }
    T acc = (T)0;
#pragma unroll
    for (int qz = 0; qz < (DIM == 3 ? 5 : 1); qz++)
#pragma unroll
      for (int qy = 0; qy < 5; qy++)
#pragma unroll
        for (int qx = 0; qx < 5; qx++) {
          const int q = qx + 5 * qy + 25 * qz;
          acc += gx[qx] * fy[qy] * fz[qz] * tau[le][q][c0];
          acc += fx[qx] * gy[qy] * fz[qz] * tau[le][q][c1];
          if (DIM == 3) acc += fx[qx] * fy[qy] * gz[qz] * tau[le][q][c2];
        }
    const long n = q2_vertex<DIM>(G, elem_of(ce), a);
    if (!((bc[n] >> i) & 1)) y[n * DIM + i] += acc;
  }
}

template <typename T>
void launch_q2_apply(const Geom& g, const T* W, const T* x, T* y, const uint8_t* bc, hipStream_t s, const T* Dm) {
  const Q2G G = q2_pack(g);
  if (g.dim == 3)
    hipLaunchKernelGGL((k_q2_init<T, 3>), q2_blocks(G.N, 256), 256, 0, s, G.N, x, bc, y);
  else
    hipLaunchKernelGGL((k_q2_init<T, 2>), q2_blocks(G.N, 256), 256, 0, s, G.N, x, bc, y);
  UWG_HIP(hipGetLastError());
  const int ncol = g.dim == 3 ? 8 : 4, epb = g.dim == 3 ? 1 : 5;
  for (int c = 0; c < ncol; c++) {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    const int ncx = (G.rx - cx + 1) / 2, ncy = (G.ry - cy + 1) / 2, ncz = g.dim == 3 ? (G.rz - cz + 1) / 2 : 1;
    const long nce = (long)ncx * ncy * ncz;
    if (nce == 0) continue;
    if (g.dim == 3 && g_q2_kernel == 2) {
      hipLaunchKernelGGL((k_q2_apply_sf<T>), q2_blocks(nce, Q2S_E), 128, 0, s, G, cx, cy, cz, ncx, ncy, nce, W, x, y,
                         bc);
    } else if (g.dim == 3 && Dm && g_q2_kernel == 1) {
      const long nbatch = (nce + Q2M_E - 1) / Q2M_E;
      if (sizeof(T) == 4 && nbatch >= 2L * 2048) {
        // persistent blocks (2+ batches each) keep their slice of D in registers
        hipLaunchKernelGGL((k_q2_apply_mfma<T, true>), 2048, 256, 0, s, G, cx, cy, cz, ncx, ncy, nce, W, Dm, x, y, bc);
      } else if (sizeof(T) == 8) {
        hipLaunchKernelGGL((k_q2_apply_mfma<T, false>), (unsigned)nbatch, 256, 0, s, G, cx, cy, cz, ncx, ncy, nce, W,
                           Dm, x, y, bc);
      } else {  // small FP32 grids: the vector kernel is faster
        hipLaunchKernelGGL((k_q2_apply<T, 3>), q2_blocks(nce, epb), 128, 0, s, G, cx, cy, cz, ncx, ncy, nce, W, x, y,
                           bc);
      }
    }
    else if (g.dim == 3)
      hipLaunchKernelGGL((k_q2_apply<T, 3>), q2_blocks(nce, epb), 128, 0, s, G, cx, cy, cz, ncx, ncy, nce, W, x, y, bc);
    else
      hipLaunchKernelGGL((k_q2_apply<T, 2>), q2_blocks(nce, epb), 128, 0, s, G, cx, cy, cz, ncx, ncy, nce, W, x, y, bc);
    UWG_HIP(hipGetLastError());
  }
}
template void launch_q2_apply<double>(const Geom&, const double*, const double*, double*, const uint8_t*, hipStream_t,
                                      const double*);
template void launch_q2_apply<float>(const Geom&, const float*, const float*, float*, const uint8_t*, hipStream_t,
                                     const float*);


// ---------------------------------------------------------------------------------------
// Jacobi diagonal: K[ai][ai] = sum_e sum_q W_q (|grad N_a|^2 + (dN_a/dx_i)^2)
// ---------------------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ void q2_kdiag_elem(const Q2G& G, const double* __restrict__ W, long e, int a,
                                              double* acc) {
  constexpr int NQ = DIM == 3 ? 125 : 25;
  const int ax = a % 3, ay = (a / 3) % 3, az = DIM == 3 ? a / 9 : 0;
  for (int q = 0; q < NQ; q++) {
    const int qx = q % 5, qy = (q / 5) % 5, qz = DIM == 3 ? q / 25 : 0;
    const double X = g5(qx), Y = g5(qy), Z = g5(qz);
    const double fx = ph(ax, X), fy = ph(ay, Y), fz = DIM == 3 ? ph(az, Z) : 1.0;
    double dn[3];
    dn[0] = dph(ax, X) * G.s[0] * fy * fz;
    dn[1] = fx * dph(ay, Y) * G.s[1] * fz;
    dn[2] = DIM == 3 ? fx * fy * dph(az, Z) * G.s[2] : 0.0;
    const double s2 = dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2];
    const double w = W[e * NQ + q];
    for (int i = 0; i < DIM; i++) acc[i] += w * (s2 + dn[i] * dn[i]);
  }
}

template <int DIM>
__global__ void k_q2_diag(Q2G G, const double* __restrict__ W, const uint8_t* __restrict__ bc,
                          double* __restrict__ dinv) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  const int I = (int)(n % G.nx), J = (int)((n / G.nx) % G.ny), K = DIM == 3 ? (int)(n / ((long)G.nx * G.ny)) : 0;
  int ex[2], lx[2], ey[2], ly[2], ez[2] = {0, 0}, lz[2] = {0, 0};
  const int cx = q2_adj(I, G.rx, ex, lx), cy = q2_adj(J, G.ry, ey, ly), cz = DIM == 3 ? q2_adj(K, G.rz, ez, lz) : 1;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < cz; c++)
    for (int b = 0; b < cy; b++)
      for (int a = 0; a < cx; a++) {
        const long e = ex[a] + (long)G.rx * (ey[b] + (long)G.ry * ez[c]);
        q2_kdiag_elem<DIM>(G, W, e, lx[a] + 3 * ly[b] + 9 * lz[c], acc);
      }
  const uint8_t m = bc[n];
  for (int d = 0; d < DIM; d++) dinv[n * DIM + d] = ((m >> d) & 1) ? 1.0 : 1.0 / acc[d];
}

void launch_q2_diag(const Geom& g, const double* W, const uint8_t* bc, double* dinv, hipStream_t s) {
  const Q2G G = q2_pack(g);
  if (g.dim == 3)
    hipLaunchKernelGGL(k_q2_diag<3>, q2_blocks(G.N, 128), 128, 0, s, G, W, bc, dinv);
  else
    hipLaunchKernelGGL(k_q2_diag<2>, q2_blocks(G.N, 128), 128, 0, s, G, W, bc, dinv);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// G (dPc1 pressure -> velocity) and G^T; Gr = G_e [NEN*dim][dim+1], the same for every element
// ---------------------------------------------------------------------------------------
template <typename T, int DIM>
__global__ void k_q2_apply_G(Q2G G, const double* __restrict__ Gr, const T* __restrict__ p,
                             const uint8_t* __restrict__ bc, T* __restrict__ y) {
  constexpr int NP = DIM + 1;
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  const int I = (int)(n % G.nx), J = (int)((n / G.nx) % G.ny), K = DIM == 3 ? (int)(n / ((long)G.nx * G.ny)) : 0;
  int ex[2], lx[2], ey[2], ly[2], ez[2] = {0, 0}, lz[2] = {0, 0};
  const int cx = q2_adj(I, G.rx, ex, lx), cy = q2_adj(J, G.ry, ey, ly), cz = DIM == 3 ? q2_adj(K, G.rz, ez, lz) : 1;
  T acc[3] = {0, 0, 0};
  for (int c = 0; c < cz; c++)
    for (int b = 0; b < cy; b++)
      for (int a = 0; a < cx; a++) {
        const long e = ex[a] + (long)G.rx * (ey[b] + (long)G.ry * ez[c]);
        const int loc = lx[a] + 3 * ly[b] + 9 * lz[c];
        for (int r = 0; r < NP; r++) {
          const T pv = p[e * NP + r];
          for (int d = 0; d < DIM; d++) acc[d] += (T)Gr[(loc * DIM + d) * NP + r] * pv;
        }
      }
  const uint8_t m = bc[n];
  for (int d = 0; d < DIM; d++) y[n * DIM + d] = ((m >> d) & 1) ? (T)0 : acc[d];
}

template <typename T, int DIM>
__global__ void k_q2_apply_GT(Q2G G, const double* __restrict__ Gr, const T* __restrict__ u,
                              const uint8_t* __restrict__ bc, T* __restrict__ q) {
  constexpr int NP = DIM + 1, NEN = DIM == 3 ? 27 : 9;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G.E) return;
  T acc[4] = {0, 0, 0, 0};
  for (int a = 0; a < NEN; a++) {
    const long n = q2_vertex<DIM>(G, e, a);
    const uint8_t m = bc[n];
    for (int d = 0; d < DIM; d++) {
      if ((m >> d) & 1) continue;
      const T v = u[n * DIM + d];
      for (int r = 0; r < NP; r++) acc[r] += (T)Gr[(a * DIM + d) * NP + r] * v;
    }
  }
  for (int r = 0; r < NP; r++) q[e * NP + r] = acc[r];
}

template <typename T>
void launch_q2_apply_G(const Geom& g, const double* Gr, const T* p, const uint8_t* bc, T* y, hipStream_t s) {
  const Q2G G = q2_pack(g);
  if (g.dim == 3)
    hipLaunchKernelGGL((k_q2_apply_G<T, 3>), q2_blocks(G.N, 256), 256, 0, s, G, Gr, p, bc, y);
  else
    hipLaunchKernelGGL((k_q2_apply_G<T, 2>), q2_blocks(G.N, 256), 256, 0, s, G, Gr, p, bc, y);
  UWG_HIP(hipGetLastError());
}
template <typename T>
void launch_q2_apply_GT(const Geom& g, const double* Gr, const T* u, const uint8_t* bc, T* q, hipStream_t s) {
  const Q2G G = q2_pack(g);
  if (g.dim == 3)
    hipLaunchKernelGGL((k_q2_apply_GT<T, 3>), q2_blocks(G.E, 128), 128, 0, s, G, Gr, u, bc, q);
  else
    hipLaunchKernelGGL((k_q2_apply_GT<T, 2>), q2_blocks(G.E, 128), 128, 0, s, G, Gr, u, bc, q);
  UWG_HIP(hipGetLastError());
}
template void launch_q2_apply_G<double>(const Geom&, const double*, const double*, const uint8_t*, double*, hipStream_t);
template void launch_q2_apply_G<float>(const Geom&, const double*, const float*, const uint8_t*, float*, hipStream_t);
template void launch_q2_apply_GT<double>(const Geom&, const double*, const double*, const uint8_t*, double*, hipStream_t);
template void launch_q2_apply_GT<float>(const Geom&, const double*, const float*, const uint8_t*, float*, hipStream_t);

// ---------------------------------------------------------------------------------------
// approxS_e = G_e^T diag(K_e)^-1 G_e (UzawaPreconditionerTerm.c:124-209), stored inverted
// ---------------------------------------------------------------------------------------
template <int DIM>
__global__ void k_q2_approxS(Q2G G, const double* __restrict__ W, const double* __restrict__ Gr,
                             double* __restrict__ S, double* __restrict__ Sinv, int* __restrict__ bad) {
  constexpr int NP = DIM + 1, NEN = DIM == 3 ? 27 : 9;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G.E) return;
  double M[NP][NP] = {};
  for (int a = 0; a < NEN; a++) {
    double kd[3] = {0.0, 0.0, 0.0};
    q2_kdiag_elem<DIM>(G, W, e, a, kd);
    for (int i = 0; i < DIM; i++) {
      const double* gv = Gr + (a * DIM + i) * NP;
      const double inv = 1.0 / kd[i];
      for (int r = 0; r < NP; r++)
        for (int c = 0; c < NP; c++) M[r][c] += gv[r] * gv[c] * inv;
    }
  }
  for (int r = 0; r < NP; r++)
    for (int c = 0; c < NP; c++) S[e * NP * NP + r * NP + c] = M[r][c];
  // Gauss-Jordan inverse (symmetric positive definite block)
  double A[NP][2 * NP];
  for (int r = 0; r < NP; r++)
    for (int c = 0; c < NP; c++) {
      A[r][c] = M[r][c];
      A[r][NP + c] = r == c ? 1.0 : 0.0;
    }
  bool ok = true;
  for (int c = 0; c < NP; c++) {
    int piv = c;
    for (int r = c + 1; r < NP; r++)
      if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
    if (!(fabs(A[piv][c]) > 0.0)) {
      ok = false;
      break;
    }
    if (piv != c)
      for (int k = 0; k < 2 * NP; k++) {
        const double tmp = A[c][k];
        A[c][k] = A[piv][k];
        A[piv][k] = tmp;
      }
    const double inv = 1.0 / A[c][c];
    for (int k = 0; k < 2 * NP; k++) A[c][k] *= inv;
    for (int r = 0; r < NP; r++) {
      if (r == c) continue;
      const double f = A[r][c];
      for (int k = 0; k < 2 * NP; k++) A[r][k] -= f * A[c][k];
    }
  }
  if (!ok) bad[0] = 1;
  for (int r = 0; r < NP; r++)
    for (int c = 0; c < NP; c++) Sinv[e * NP * NP + r * NP + c] = ok ? A[r][NP + c] : 0.0;
}

void launch_q2_approxS(const Geom& g, const double* W, const double* Gr, double* S, double* Sinv, int* bad,
                       hipStream_t s) {
  const Q2G G = q2_pack(g);
  if (g.dim == 3)
    hipLaunchKernelGGL(k_q2_approxS<3>, q2_blocks(G.E, 64), 64, 0, s, G, W, Gr, S, Sinv, bad);
  else
    hipLaunchKernelGGL(k_q2_approxS<2>, q2_blocks(G.E, 64), 64, 0, s, G, W, Gr, S, Sinv, bad);
  UWG_HIP(hipGetLastError());
}

// z_e = -Sinv_e r_e  (block Schur preconditioner)
template <typename T, typename S>
__global__ void k_q2_schur_prec(long E, int np, const S* __restrict__ Sinv, const T* __restrict__ r,
                                T* __restrict__ z) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  T rv[4];
  for (int c = 0; c < np; c++) rv[c] = r[e * np + c];
  for (int a = 0; a < np; a++) {
    T acc = (T)0;
    for (int c = 0; c < np; c++) acc += (T)Sinv[e * np * np + a * np + c] * rv[c];
    z[e * np + a] = -acc;
  }
}

template <typename T>
void launch_q2_schur_prec(long E, int np, const T* Sinv, const T* r, T* z, hipStream_t s) {
  hipLaunchKernelGGL((k_q2_schur_prec<T, T>), q2_blocks(E, 256), 256, 0, s, E, np, Sinv, r, z);
  UWG_HIP(hipGetLastError());
}
template void launch_q2_schur_prec<double>(long, int, const double*, const double*, double*, hipStream_t);
template void launch_q2_schur_prec<float>(long, int, const float*, const float*, float*, hipStream_t);

// ---------------------------------------------------------------------------------------
// Q1 weights of the embedded trilinear space: the Galerkin operator P^T K_Q2 P (P the exact
// Q1 -> Q2 interpolation) has integrands of degree <= 2 per direction, reproduced by the
// 3-point Gauss weights W1_r = sum_q W_q L3_r(g_q).  Output in the Q1 layout W1[r][Es].
// ---------------------------------------------------------------------------------------
template <int DIM>
__global__ void k_q2_project_W(long E, long Es, const double* __restrict__ W, double* __restrict__ W1) {
  constexpr int NQ = DIM == 3 ? 125 : 25, NR = DIM == 3 ? 27 : 9;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= E * NR) return;
  const long e = idx / NR;
  const int r = (int)(idx % NR);
  const int rx = r % 3, ry = (r / 3) % 3, rz = DIM == 3 ? r / 9 : 0;
  double acc = 0.0;
  for (int q = 0; q < NQ; q++) {
    const int qx = q % 5, qy = (q / 5) % 5, qz = DIM == 3 ? q / 25 : 0;
    const double L = lag3(rx, g5(qx)) * lag3(ry, g5(qy)) * (DIM == 3 ? lag3(rz, g5(qz)) : 1.0);
    acc += W[e * NQ + q] * L;
  }
  W1[(long)r * Es + e] = acc;
}

void launch_q2_project_W(const Geom& g, const double* W, double* W1, hipStream_t s) {
  const int NR = g.dim == 3 ? 27 : 9;
  if (g.dim == 3)
    hipLaunchKernelGGL(k_q2_project_W<3>, q2_blocks(g.E * NR, 256), 256, 0, s, g.E, g.Es, W, W1);
  else
    hipLaunchKernelGGL(k_q2_project_W<2>, q2_blocks(g.E * NR, 256), 256, 0, s, g.E, g.Es, W, W1);
  UWG_HIP(hipGetLastError());
}

// G_e on the host: -sum_g w_g detJ dN_a/dx_i(g) N^p_r(g) at the 5^d Gauss points, with the
// reference's dPc1 basis (BilinearInnerElType.c / TrilinearInnerElType.c SF_allNodes)
void q2_element_G(const Geom& g, std::vector<double>& Gr) {
  const int dim = g.dim, NEN = dim == 3 ? 27 : 9, NP = dim + 1, NQ = dim == 3 ? 125 : 25;
  const Q2G G = q2_pack(g);
  Gr.assign((size_t)NEN * dim * NP, 0.0);
  for (int q = 0; q < NQ; q++) {
    const int qx = q % 5, qy = (q / 5) % 5, qz = dim == 3 ? q / 25 : 0;
    const double X = g5(qx), Y = g5(qy), Z = dim == 3 ? g5(qz) : 0.0;
    const double wq = w5(qx) * w5(qy) * (dim == 3 ? w5(qz) : 1.0) * G.detJ;
    double Np[4];
    if (dim == 2) {
      Np[0] = -X - 0.5 * Y + 0.25;
      Np[1] = X - 0.5 * Y + 0.25;
      Np[2] = Y + 0.5;
    } else {
      const double f = 1.0 / 9.0;
      Np[0] = f * (-4.0 * X - 20.0 * Y + 12.0 * Z + 3.0);
      Np[1] = f * (16.0 * X - 64.0 * Y + 60.0 * Z + 6.0);
      Np[2] = f * (36.0 * Y - 36.0 * Z);
      Np[3] = f * (-12.0 * X + 48.0 * Y - 36.0 * Z);
    }
    for (int a = 0; a < NEN; a++) {
      const int ax = a % 3, ay = (a / 3) % 3, az = dim == 3 ? a / 9 : 0;
      const double fx = ph(ax, X), fy = ph(ay, Y), fz = dim == 3 ? ph(az, Z) : 1.0;
      const double dn[3] = {dph(ax, X) * G.s[0] * fy * fz, fx * dph(ay, Y) * G.s[1] * fz,
                            dim == 3 ? fx * fy * dph(az, Z) * G.s[2] : 0.0};
      for (int i = 0; i < dim; i++)
        for (int r = 0; r < NP; r++) Gr[(size_t)(a * dim + i) * NP + r] -= wq * dn[i] * Np[r];
    }
  }
}

// nodal Q2 field interpolated at points (FeVariable_InterpolateValueAt with the quadratic
// shape functions): element search on the element grid, affine local coordinate
template <int DIM>
__global__ void k_q2_interp(Q2G G, double lox, double loy, double loz, double hix, double hiy, double hiz, long P,
                            const double* __restrict__ f, int nc, const double* __restrict__ x, double* __restrict__ v,
                            int* __restrict__ ok) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double lo[3] = {lox, loy, loz}, hi[3] = {hix, hiy, hiz};
  const int res[3] = {G.rx, G.ry, G.rz};
  int el[3] = {0, 0, 0};
  double xi[3] = {0.0, 0.0, 0.0};
  bool good = true;
  for (int d = 0; d < DIM; d++) {
    const double xd = x[p * DIM + d];
    if (!(xd >= lo[d] && xd <= hi[d])) {
      good = false;
      break;
    }
    const double h = (hi[d] - lo[d]) / res[d];
    int i = (int)floor((xd - lo[d]) / h);
    if (i >= res[d]) i = res[d] - 1;
    if (i < 0) i = 0;
    el[d] = i;
    xi[d] = 2.0 * (xd - (lo[d] + (i + 0.5) * h)) / h;
  }
  ok[p] = good ? 1 : 0;
  for (int c = 0; c < nc; c++) {
    double acc = 0.0;
    if (good)
      for (int a = 0; a < (DIM == 3 ? 27 : 9); a++) {
        const int ax = a % 3, ay = (a / 3) % 3, az = DIM == 3 ? a / 9 : 0;
        const long n = (2 * el[0] + ax) + (long)G.nx * ((2 * el[1] + ay) + (long)G.ny * (2 * el[2] + az));
        acc += f[n * nc + c] * (ph(ax, xi[0]) * ph(ay, xi[1]) * (DIM == 3 ? ph(az, xi[2]) : 1.0));
      }
    v[p * nc + c] = acc;
  }
}

void launch_q2_interpolate(const Geom& g, long P, const double* f, int nc, const double* x, double* v, int* ok,
                           hipStream_t s) {
  if (P == 0) return;
  const Q2G G = q2_pack(g);
  if (g.dim == 3)
    hipLaunchKernelGGL(k_q2_interp<3>, q2_blocks(P, 256), 256, 0, s, G, g.lo[0], g.lo[1], g.lo[2], g.hi[0], g.hi[1],
                       g.hi[2], P, f, nc, x, v, ok);
  else
    hipLaunchKernelGGL(k_q2_interp<2>, q2_blocks(P, 256), 256, 0, s, G, g.lo[0], g.lo[1], 0.0, g.hi[0], g.hi[1], 1.0,
                       P, f, nc, x, v, ok);
  UWG_HIP(hipGetLastError());
}

}  // namespace uwg
