// advdiff_kernels.hip -- SUPG advection-diffusion of a nodal scalar on a Q1 mesh
// (AdvectionDiffusion(method="SUPG"), PY/systems/_advectiondiffusion.py:838-1010).
//
// AD/ = LU/Underworld/Utils/AdvectionDiffusion/src/.
//  * k_supg_elem: the element residual of AD/Residual.cpp:223-343 with the per-element upwind
//    diffusivity of AD/Residual.cpp:430-513 -- one thread per element, the 3^dim Gauss points
//    of the Q1 mesh, written to [E][2^dim] (no atomics).
//  * k_supg_correct: vertex gather of the element residuals in a fixed element order, then the
//    explicit corrector of AD/Multicorrector.c:206-260: dphidot = R / M (lumped mass),
//    phi += dt gamma dphidot, phidot += dphidot on free dofs, nodal fields refreshed.
//  * k_supg_mass: the lumped mass of AD/LumpedMassMatrixForceTerm.c:88-129.
//  * k_supg_adv_dt: min over vertices and axes of |h_d / v_d| (AD/Timestep.c:77-110).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "uwg_geom.h"
#include "uwg_internal.h"

namespace uwg {

namespace {
inline unsigned ad_blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
}  // namespace

#define AD_ISQRT15 0.25819888974716112567
#define AD_MIN_DIFFUSIVITY 1.0e-20

struct ADGeom {
  int res[3];
  double lo[3], hi[3];
  long E, N;
};
static ADGeom ad_pack(const Geom& g) {
  ADGeom a{};
  for (int d = 0; d < 3; d++) {
    a.res[d] = g.res[d];
    a.lo[d] = g.lo[d];
    a.hi[d] = g.hi[d];
  }
  a.E = g.E;
  a.N = g.N;
  return a;
}

// upwind parameter xi(Pe): 0 exact coth(a) - 1/a, 1 doubly asymptotic, 2 critical
__device__ __forceinline__ double ad_upwind(double pe, int kind) {
  if (kind == 0) {
    if (fabs(pe) < 1.0e-8) return 0.33333333333333 * pe;
    if (pe < -20.0) return -1.0 - 1.0 / pe;
    if (pe > 20.0) return 1.0 - 1.0 / pe;
    return cosh(pe) / sinh(pe) - 1.0 / pe;
  }
  if (kind == 1) {
    if (pe <= -3.0) return -1.0;
    if (pe <= 3.0) return 0.33333333333333 * pe;
    return 1.0;
  }
  if (pe <= -1.0) return -1.0 - 1.0 / pe;
  if (pe <= 1.0) return 0.0;
  return 1.0 - 1.0 / pe;
}

template <int DIM>
__device__ __forceinline__ void ad_gauss(int q, double* X, double* w) {
  const double ga = 0.77459666924148337704;
  const double x1[3] = {-ga, 0.0, ga}, w1[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  X[0] = x1[q % 3];
  X[1] = x1[(q / 3) % 3];
  X[2] = DIM == 3 ? x1[q / 9] : 0.0;
  *w = w1[q % 3] * w1[(q / 3) % 3] * (DIM == 3 ? w1[q / 9] : 1.0);
}

template <int DIM>
__device__ __forceinline__ void ad_shape(const double* X, double* N, double (*dN)[8]) {
  constexpr int NN = DIM == 3 ? 8 : 4;
  for (int a = 0; a < NN; a++) {
    double s[3], f[3];
    for (int d = 0; d < DIM; d++) {
      s[d] = ((a >> d) & 1) ? 1.0 : -1.0;
      f[d] = 0.5 * (1.0 + s[d] * X[d]);
    }
    N[a] = f[0] * f[1] * (DIM == 3 ? f[2] : 1.0);
    for (int k = 0; k < DIM; k++) {
      double v = 0.5 * s[k];
      for (int d = 0; d < DIM; d++)
        if (d != k) v *= f[d];
      dN[k][a] = v;
    }
  }
}

template <int DIM>
__global__ void k_supg_elem(ADGeom G, const double* __restrict__ phi, const double* __restrict__ phidot,
                            const double* __restrict__ vel, const double* __restrict__ kap,
                            const double* __restrict__ src, int upwind, double* __restrict__ res_e) {
  constexpr int NN = DIM == 3 ? 8 : 4, Q = DIM == 3 ? 27 : 9;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G.E) return;
  double V[24];
  element_verts<DIM>(G.res, G.lo, G.hi, e, V);
  double h[3] = {1.0, 1.0, 1.0}, detJ = 1.0;
  for (int d = 0; d < DIM; d++) {
    h[d] = fabs(V[(NN - 1) * DIM + d] - V[d]);
    detJ *= 0.5 * h[d];
  }
  const long nx = G.res[0] + 1, ny = G.res[1] + 1;
  const long i = e % G.res[0], j = (e / G.res[0]) % G.res[1], k = DIM == 3 ? e / ((long)G.res[0] * G.res[1]) : 0;
  const long base = i + nx * (j + ny * k);
  long nodes[8];
  double ph[8], pd[8], vv[8][3];
  for (int a = 0; a < NN; a++) {
    nodes[a] = base + (a & 1) + ((a >> 1) & 1) * nx + (DIM == 3 ? ((a >> 2) & 1) * nx * ny : 0);
    ph[a] = phi[nodes[a]];
    pd[a] = phidot[nodes[a]];
    for (int d = 0; d < DIM; d++) vv[a][d] = vel[nodes[a] * DIM + d];
  }
  // upwind diffusivity (element mean diffusivity, velocity at the centre)
  double avg = 0.0;
  for (int q = 0; q < Q; q++) avg += kap[e * Q + q];
  avg /= (double)Q;
  if (avg < AD_MIN_DIFFUSIVITY) avg = AD_MIN_DIFFUSIVITY;
  double vc[3] = {0.0, 0.0, 0.0};
  {
    const double Xc[3] = {0.0, 0.0, 0.0};
    double Nc[8], dNc[3][8];
    ad_shape<DIM>(Xc, Nc, dNc);
    for (int a = 0; a < NN; a++)
      for (int d = 0; d < DIM; d++) vc[d] += vv[a][d] * Nc[a];
  }
  double up = 0.0;
  for (int d = 0; d < DIM; d++) {
    const double pe = vc[d] * h[d] / (2.0 * avg);
    up += ad_upwind(pe, upwind) * vc[d] * h[d];
  }
  up *= AD_ISQRT15;
  double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = 0; q < Q; q++) {
    double X[3], w, N[8], dN[3][8];
    ad_gauss<DIM>(q, X, &w);
    ad_shape<DIM>(X, N, dN);
    for (int d = 0; d < DIM; d++)
      for (int a = 0; a < NN; a++) dN[d][a] *= 2.0 / h[d];
    double v[3] = {0.0, 0.0, 0.0}, gphi[3] = {0.0, 0.0, 0.0}, pdot = 0.0;
    for (int a = 0; a < NN; a++) {
      for (int d = 0; d < DIM; d++) {
        v[d] += vv[a][d] * N[a];
        gphi[d] += ph[a] * dN[d][a];
      }
      pdot += pd[a] * N[a];
    }
    double udotu = 0.0, vg = 0.0;
    for (int d = 0; d < DIM; d++) {
      udotu += v[d] * v[d];
      vg += v[d] * gphi[d];
    }
    const double fac = w * detJ;
    const double kq = kap[e * Q + q];
    const double total = pdot + vg - (src ? src[e * Q + q] : 0.0);
    const bool plain = fabs(up) < AD_MIN_DIFFUSIVITY;
    const double supgf = plain ? 0.0 : up / udotu;
    for (int a = 0; a < NN; a++) {
      double pert = 0.0, dif = 0.0;
      for (int d = 0; d < DIM; d++) {
        pert += v[d] * dN[d][a];
        dif += dN[d][a] * gphi[d];
      }
      const double sn = plain ? N[a] : N[a] + supgf * pert;
      r[a] -= fac * (sn * total + kq * dif);
    }
  }
  for (int a = 0; a < NN; a++) res_e[e * NN + a] = r[a];
}

template <int DIM>
__global__ void k_supg_mass(ADGeom G, double* __restrict__ me) {
  constexpr int NN = DIM == 3 ? 8 : 4, Q = DIM == 3 ? 27 : 9;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G.E) return;
  double V[24];
  element_verts<DIM>(G.res, G.lo, G.hi, e, V);
  double detJ = 1.0;
  for (int d = 0; d < DIM; d++) detJ *= 0.5 * fabs(V[(NN - 1) * DIM + d] - V[d]);
  double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = 0; q < Q; q++) {
    double X[3], w, N[8], dN[3][8];
    ad_gauss<DIM>(q, X, &w);
    ad_shape<DIM>(X, N, dN);
    const double fac = detJ * w;
    for (int a = 0; a < NN; a++)
      for (int b = 0; b < NN; b++) m[a] += N[a] * N[b] * fac;
  }
  for (int a = 0; a < NN; a++) me[e * NN + a] = m[a];
}

// sum of the element entries around vertex n (fixed order: lower elements first)
template <int DIM>
__device__ __forceinline__ double ad_gather(const ADGeom& G, const double* __restrict__ ve, long n) {
  constexpr int NN = DIM == 3 ? 8 : 4;
  const long nx = G.res[0] + 1, ny = G.res[1] + 1;
  const int I = (int)(n % nx), J = (int)((n / nx) % ny), K = DIM == 3 ? (int)(n / (nx * ny)) : 0;
  double acc = 0.0;
  for (int oz = 0; oz < (DIM == 3 ? 2 : 1); oz++)
    for (int oy = 0; oy < 2; oy++)
      for (int ox = 0; ox < 2; ox++) {
        const int ei = I - 1 + ox, ej = J - 1 + oy, ek = DIM == 3 ? K - 1 + oz : 0;
        if (ei < 0 || ei >= G.res[0] || ej < 0 || ej >= G.res[1] || ek < 0 || ek >= G.res[2]) continue;
        const long e = ei + (long)G.res[0] * (ej + (long)G.res[1] * ek);
        const int a = (1 - ox) + 2 * (1 - oy) + (DIM == 3 ? 4 * (1 - oz) : 0);
        acc += ve[e * NN + a];
      }
  return acc;
}

template <int DIM>
__global__ void k_supg_gather_mass(ADGeom G, const double* __restrict__ me, double* __restrict__ M) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  M[n] = ad_gather<DIM>(G, me, n);
}

// predictor (AD/Multicorrector.c:171-200) on the solution vectors: phi_v = phi + dt (1-gamma)
// phidot, phidot_v = 0 on free dofs; the nodal fields are left as they are
__global__ void k_supg_predict(long N, const uint8_t* __restrict__ bc, const double* __restrict__ phi,
                               const double* __restrict__ phidot, double a, double* __restrict__ phi_v,
                               double* __restrict__ pdot_v) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  if (bc[n] & 1) return;
  phi_v[n] = phi[n] + a * phidot[n];
  pdot_v[n] = 0.0;
}

template <int DIM>
__global__ void k_supg_correct(ADGeom G, const uint8_t* __restrict__ bc, const double* __restrict__ res_e,
                               const double* __restrict__ M, double dtg, double* __restrict__ phi_v,
                               double* __restrict__ pdot_v, double* __restrict__ phi, double* __restrict__ phidot) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  if (bc[n] & 1) return;
  const double dp = ad_gather<DIM>(G, res_e, n) / M[n];
  const double pv = phi_v[n] + dtg * dp;
  const double dv = pdot_v[n] + dp;
  phi_v[n] = pv;
  pdot_v[n] = dv;
  phi[n] = pv;
  phidot[n] = dv;
}

// per-block minimum over vertices and axes of |sep_d / v_d| (zero components skipped)
template <int DIM>
__global__ void k_supg_adv_dt(long N, const double* __restrict__ vel, double sx, double sy, double sz,
                              double* __restrict__ partial) {
  __shared__ double sm[256];
  double m = HUGE_VAL;
  const double sep[3] = {sx, sy, sz};
  for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (long)gridDim.x * blockDim.x)
    for (int d = 0; d < DIM; d++) {
      const double v = vel[n * DIM + d];
      if (v == 0.0) continue;
      const double t = fabs(sep[d] / v);
      m = t < m ? t : m;
    }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

void launch_supg_elem(const Geom& g, const double* phi, const double* phidot, const double* vel, const double* kap,
                      const double* src, int upwind, double* res_e, hipStream_t s) {
  const ADGeom G = ad_pack(g);
  if (g.dim == 3)
    hipLaunchKernelGGL(k_supg_elem<3>, ad_blocks(g.E, 128), 128, 0, s, G, phi, phidot, vel, kap, src, upwind, res_e);
  else
    hipLaunchKernelGGL(k_supg_elem<2>, ad_blocks(g.E, 128), 128, 0, s, G, phi, phidot, vel, kap, src, upwind, res_e);
  UWG_HIP(hipGetLastError());
}

void launch_supg_mass(const Geom& g, double* me, double* M, hipStream_t s) {
  const ADGeom G = ad_pack(g);
  if (g.dim == 3) {
    hipLaunchKernelGGL(k_supg_mass<3>, ad_blocks(g.E, 128), 128, 0, s, G, me);
    hipLaunchKernelGGL(k_supg_gather_mass<3>, ad_blocks(g.N, 256), 256, 0, s, G, me, M);
  } else {
    hipLaunchKernelGGL(k_supg_mass<2>, ad_blocks(g.E, 128), 128, 0, s, G, me);
    hipLaunchKernelGGL(k_supg_gather_mass<2>, ad_blocks(g.N, 256), 256, 0, s, G, me, M);
  }
  UWG_HIP(hipGetLastError());
}

void launch_supg_predict(long N, const uint8_t* bc, const double* phi, const double* phidot, double a, double* phi_v,
                         double* pdot_v, hipStream_t s) {
  hipLaunchKernelGGL(k_supg_predict, ad_blocks(N, 256), 256, 0, s, N, bc, phi, phidot, a, phi_v, pdot_v);
  UWG_HIP(hipGetLastError());
}

void launch_supg_correct(const Geom& g, const uint8_t* bc, const double* res_e, const double* M, double dtg,
                         double* phi_v, double* pdot_v, double* phi, double* phidot, hipStream_t s) {
  const ADGeom G = ad_pack(g);
  if (g.dim == 3)
    hipLaunchKernelGGL(k_supg_correct<3>, ad_blocks(g.N, 256), 256, 0, s, G, bc, res_e, M, dtg, phi_v, pdot_v, phi,
                       phidot);
  else
    hipLaunchKernelGGL(k_supg_correct<2>, ad_blocks(g.N, 256), 256, 0, s, G, bc, res_e, M, dtg, phi_v, pdot_v, phi,
                       phidot);
  UWG_HIP(hipGetLastError());
}

void launch_supg_adv_dt(const Geom& g, const double* vel, const double* sep, double* partial, int nblk,
                        hipStream_t s) {
  if (g.dim == 3)
    hipLaunchKernelGGL(k_supg_adv_dt<3>, nblk, 256, 0, s, g.N, vel, sep[0], sep[1], sep[2], partial);
  else
    hipLaunchKernelGGL(k_supg_adv_dt<2>, nblk, 256, 0, s, g.N, vel, sep[0], sep[1], 1.0, partial);
  UWG_HIP(hipGetLastError());
}

}  // namespace uwg
