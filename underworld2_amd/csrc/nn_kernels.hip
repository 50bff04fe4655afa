// Learned coarse-grid correction: a 4-level convolutional U-net on the nodal grid of one
// multigrid level, evaluated and trained with the kernels in this file (FP32 storage and
// arithmetic, deterministic reductions).
//
// Input channels per node: the residual (dim components, scaled by 1/rms) and a log-viscosity
// feature.  Output: the correction (dim components, rescaled by rms), zero on Dirichlet dofs.
//
// Network (levels l = 0..3 on node grids n_l, n_{l+1} = (n_l + 1) / 2):
//   e_0 = relu(conv(relu(conv(x, Cin -> C))))            ; level 0
//   e_l = relu(conv(relu(conv(R e_{l-1} / 8))))           ; l = 1..3 (R: full weighting)
//   d_3 = e_3 ; d_l = relu(conv(P d_{l+1} + e_l))         ; l = 2..0 (P: trilinear)
//   y   = conv(d_0, C -> dim)
// R is the transpose of P, so the backward pass of a pooling is the other operator.
//
// Every convolution is 3x3x3 with zero padding; activations are [B][C][N] (N = nx*ny*nz,
// x fastest).  Grids here are at most 33^3 nodes, where direct convolution with weights
// broadcast from the scalar cache is latency-bound rather than FLOP-bound, so MFMA tiles
// would not pay.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "uwg_nn.h"

namespace uwg {

static inline unsigned nbk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct Grid3 {
  int nx, ny, nz;
  __host__ __device__ long N() const { return (long)nx * ny * nz; }
};
static Grid3 coarser(Grid3 g) { return Grid3{(g.nx + 1) / 2, (g.ny + 1) / 2, (g.nz + 1) / 2}; }

// ------------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------------
__global__ void k_conv_fwd(int B, int Ci, int Co, Grid3 g, const float* __restrict__ in, const float* __restrict__ w,
                           const float* __restrict__ bias, float* __restrict__ out, int relu) {
  const long N = g.N();
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)B * Co * N) return;
  const long p = t % N;
  const int co = (int)((t / N) % Co);
  const int b = (int)(t / (N * Co));
  const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((long)g.nx * g.ny));
  float acc = bias[co];
  const float* wb = w + (long)co * Ci * 27;
  for (int ci = 0; ci < Ci; ci++) {
    const float* src = in + ((long)b * Ci + ci) * N;
    const float* wc = wb + ci * 27;
#pragma unroll
    for (int dz = -1; dz <= 1; dz++) {
      int kk = k + dz;
      if (kk < 0 || kk >= g.nz) continue;
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) {
        int jj = j + dy;
        if (jj < 0 || jj >= g.ny) continue;
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          int ii = i + dx;
          if (ii < 0 || ii >= g.nx) continue;
          acc += wc[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)] * src[ii + (long)g.nx * (jj + (long)g.ny * kk)];
        }
      }
    }
  }
  out[t] = relu ? fmaxf(acc, 0.0f) : acc;
}

// Channel-blocked forward: each thread produces COB output channels of one node, so every
// input value fetched serves COB FMAs.  The block's weights for one input channel (27 taps x
// COB output channels) are staged in LDS and read as broadcasts (every lane the same address,
// 16 bytes per read).  Read straight from global memory they were wave-uniform scalar loads,
// one per weight (weights of consecutive output channels are Ci*27 apart), and the 432 loaded
// values per input channel overflowed the scalar registers: about 4,400 instructions per
// channel for 432 FMAs.  The FMA order (input channel, then taps) is unchanged, so the sums are
// the same bit for bit.
template <int COB>
__global__ __launch_bounds__(256) void k_conv_fwd_blk(int B, int Ci, int Co, Grid3 g, const float* __restrict__ in,
                                                      const float* __restrict__ w, const float* __restrict__ bias,
                                                      float* __restrict__ out, int relu) {
  __shared__ __align__(16) float ws[27 * COB];  // [tap][c]
  const long N = g.N();
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int co0 = blockIdx.y * COB;
  // threads past the last node still stage weights and reach every barrier
  const bool active = t < (long)B * N;
  const long p = active ? t % N : 0;
  const int b = active ? (int)(t / N) : 0;
  const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((long)g.nx * g.ny));
  float acc[COB];
#pragma unroll
  for (int c = 0; c < COB; c++) acc[c] = bias[co0 + c];
  const long cs = (long)Ci * 27;
  for (int ci = 0; ci < Ci; ci++) {
    __syncthreads();  // the previous channel's weights are no longer read
    for (int q = threadIdx.x; q < 27 * COB; q += blockDim.x) {
      const int c = q % COB, tap = q / COB;
      ws[q] = w[(long)(co0 + c) * cs + ci * 27 + tap];
    }
    __syncthreads();
    if (!active) continue;
    const float* src = in + ((long)b * Ci + ci) * N;
    // dz and dy stay loops: fully unrolled, the compiler hoists all 27 taps' weight reads and
    // the kernel needs 496 registers per lane (one wave per SIMD)
#pragma unroll 1
    for (int dz = -1; dz <= 1; dz++) {
      const int kk = k + dz;
      const bool okz = kk >= 0 && kk < g.nz;
#pragma unroll 1
      for (int dy = -1; dy <= 1; dy++) {
        const int jj = j + dy;
        const bool oky = okz && jj >= 0 && jj < g.ny;
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          const int ii = i + dx;
          const float v = (oky && ii >= 0 && ii < g.nx) ? src[ii + (long)g.nx * (jj + (long)g.ny * kk)] : 0.0f;
          const int tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
#pragma unroll
          for (int c = 0; c < COB; c++) acc[c] += ws[tap * COB + c] * v;
        }
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int c = 0; c < COB; c++) {
    const float a = acc[c];
    out[((long)b * Co + co0 + c) * N + p] = relu ? fmaxf(a, 0.0f) : a;
  }
}

// Split-reduction forward for inference-sized batches: a 256-thread block computes COB output
// channels of 16 nodes; the Ci*27-term reduction is split 16 ways across the block (input
// channel ci goes to split ci % 16) and the partial sums are added in LDS in split order
// (deterministic).  The per-thread serial chain is Ci/16*27 loads instead of Ci*27, which is
// what bounds the small grids of the coarse-correction network.
template <int COB>
__global__ __launch_bounds__(256) void k_conv_fwd_ks(int B, int Ci, int Co, Grid3 g, const float* __restrict__ in,
                                                     const float* __restrict__ w, const float* __restrict__ bias,
                                                     float* __restrict__ out, int relu) {
  __shared__ float part[COB][16][17];
  const long N = g.N();
  const int tid = threadIdx.x, nl = tid & 15, sp = tid >> 4;
  const long t = (long)blockIdx.x * 16 + nl;  // (batch, node) index
  const int co0 = blockIdx.y * COB;  // the launch guarantees co0 + COB <= Co
  float acc[COB];
#pragma unroll
  for (int c = 0; c < COB; c++) acc[c] = 0.0f;
  if (t < (long)B * N) {
    const long p = t % N;
    const int b = (int)(t / N);
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((long)g.nx * g.ny));
    const long cs = (long)Ci * 27;
    for (int ci = sp; ci < Ci; ci += 16) {
      const float* src = in + ((long)b * Ci + ci) * N;
      const float* wc = w + (long)co0 * cs + ci * 27;
#pragma unroll
      for (int dz = -1; dz <= 1; dz++) {
        const int kk = k + dz;
        const bool okz = kk >= 0 && kk < g.nz;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
          const int jj = j + dy;
          const bool oky = okz && jj >= 0 && jj < g.ny;
#pragma unroll
          for (int dx = -1; dx <= 1; dx++) {
            const int ii = i + dx;
            const float v = (oky && ii >= 0 && ii < g.nx) ? src[ii + (long)g.nx * (jj + (long)g.ny * kk)] : 0.0f;
            const int tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
#pragma unroll
            for (int c = 0; c < COB; c++) acc[c] += wc[c * cs + tap] * v;
          }
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < COB; c++) part[c][sp][nl] = acc[c];
  __syncthreads();
  if (tid < COB * 16) {
    const int c = tid >> 4, n2 = tid & 15;
    const long t2 = (long)blockIdx.x * 16 + n2;
    if (t2 < (long)B * N) {
      float s = bias[co0 + c];
#pragma unroll
      for (int q = 0; q < 16; q++) s += part[c][q][n2];
      const long p = t2 % N;
      const int b = (int)(t2 / N);
      out[((long)b * Co + co0 + c) * N + p] = relu ? fmaxf(s, 0.0f) : s;
    }
  }
}

// Channel-blocked input gradient: each thread produces CIB input channels of one node.
template <int CIB>
__global__ __launch_bounds__(256) void k_conv_bwd_data_blk(int B, int Ci, int Co, Grid3 g,
                                                           const float* __restrict__ dout,
                                                           const float* __restrict__ w, float* __restrict__ din) {
  const long N = g.N();
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ci0 = blockIdx.y * CIB;
  if (t >= (long)B * N) return;
  const long p = t % N;
  const int b = (int)(t / N);
  const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((long)g.nx * g.ny));
  float acc[CIB];
#pragma unroll
  for (int c = 0; c < CIB; c++) acc[c] = 0.0f;
  for (int co = 0; co < Co; co++) {
    const float* src = dout + ((long)b * Co + co) * N;
    const float* wc = w + ((long)co * Ci + ci0) * 27;
#pragma unroll
    for (int dz = -1; dz <= 1; dz++) {
      const int kk = k - dz;
      const bool okz = kk >= 0 && kk < g.nz;
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) {
        const int jj = j - dy;
        const bool oky = okz && jj >= 0 && jj < g.ny;
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          const int ii = i - dx;
          const float v = (oky && ii >= 0 && ii < g.nx) ? src[ii + (long)g.nx * (jj + (long)g.ny * kk)] : 0.0f;
          const int tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
#pragma unroll
          for (int c = 0; c < CIB; c++) acc[c] += wc[c * 27 + tap] * v;
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CIB; c++) din[((long)b * Ci + ci0 + c) * N + p] = acc[c];
}

// din[b][ci][p] = sum_co sum_tap w[co][ci][tap] dout[b][co][p - off(tap)]
__global__ void k_conv_bwd_data(int B, int Ci, int Co, Grid3 g, const float* __restrict__ dout,
                                const float* __restrict__ w, float* __restrict__ din) {
  const long N = g.N();
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)B * Ci * N) return;
  const long p = t % N;
  const int ci = (int)((t / N) % Ci);
  const int b = (int)(t / (N * Ci));
  const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((long)g.nx * g.ny));
  float acc = 0.0f;
  for (int co = 0; co < Co; co++) {
    const float* src = dout + ((long)b * Co + co) * N;
    const float* wc = w + ((long)co * Ci + ci) * 27;
#pragma unroll
    for (int dz = -1; dz <= 1; dz++) {
      int kk = k - dz;
      if (kk < 0 || kk >= g.nz) continue;
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) {
        int jj = j - dy;
        if (jj < 0 || jj >= g.ny) continue;
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          int ii = i - dx;
          if (ii < 0 || ii >= g.nx) continue;
          acc += wc[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)] * src[ii + (long)g.nx * (jj + (long)g.ny * kk)];
        }
      }
    }
  }
  din[t] = acc;
}

// dw[co][ci][tap] = sum_b sum_p dout[b][co][p] in[b][ci][p + off(tap)] ; one block per (co, ci)
__global__ __launch_bounds__(256) void k_conv_bwd_w(int B, int Ci, int Co, Grid3 g, const float* __restrict__ dout,
                                                    const float* __restrict__ in, float* __restrict__ dw) {
  __shared__ float sm[27][256];
  const int co = blockIdx.x / Ci, ci = blockIdx.x % Ci;
  const long N = g.N();
  float acc[27];
#pragma unroll
  for (int q = 0; q < 27; q++) acc[q] = 0.0f;
  for (long t = threadIdx.x; t < (long)B * N; t += 256) {
    const int b = (int)(t / N);
    const long p = t % N;
    const float go = dout[((long)b * Co + co) * N + p];
    if (go == 0.0f) continue;
    const float* src = in + ((long)b * Ci + ci) * N;
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((long)g.nx * g.ny));
#pragma unroll
    for (int dz = -1; dz <= 1; dz++) {
      int kk = k + dz;
      bool okz = kk >= 0 && kk < g.nz;
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) {
        int jj = j + dy;
        bool oky = okz && jj >= 0 && jj < g.ny;
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          int ii = i + dx;
          if (oky && ii >= 0 && ii < g.nx)
            acc[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)] += go * src[ii + (long)g.nx * (jj + (long)g.ny * kk)];
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 27; q++) sm[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int q = 0; q < 27; q++) sm[q][threadIdx.x] += sm[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 27) dw[((long)co * Ci + ci) * 27 + threadIdx.x] = sm[threadIdx.x][0];
}

// db[co] = sum_b sum_p dout[b][co][p]
__global__ __launch_bounds__(256) void k_bias_grad(int B, int Co, long N, const float* __restrict__ dout,
                                                   float* __restrict__ db) {
  __shared__ float sm[256];
  const int co = blockIdx.x;
  float a = 0.0f;
  for (long t = threadIdx.x; t < (long)B * N; t += 256) a += dout[((long)(t / N) * Co + co) * N + t % N];
  sm[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = sm[0];
}

// dy *= (y > 0)
__global__ void k_relu_bwd(long n, const float* __restrict__ y, float* __restrict__ dy) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && !(y[t] > 0.0f)) dy[t] = 0.0f;
}

// out[B*C][coarse] = s * R in[B*C][fine]   (full weighting, weights 1, 1/2 per offset)
__global__ void k_restrict_nodes(long BC, Grid3 f, Grid3 c, float s, const float* __restrict__ in,
                                 float* __restrict__ out) {
  const long Nc = c.N(), Nf = f.N();
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= BC * Nc) return;
  const long bc = t / Nc, p = t % Nc;
  const int i = (int)(p % c.nx), j = (int)((p / c.nx) % c.ny), k = (int)(p / ((long)c.nx * c.ny));
  const float* src = in + bc * Nf;
  float acc = 0.0f;
  for (int oz = -1; oz <= 1; oz++) {
    int K = 2 * k + oz;
    if (K < 0 || K >= f.nz) continue;
    for (int oy = -1; oy <= 1; oy++) {
      int J = 2 * j + oy;
      if (J < 0 || J >= f.ny) continue;
      for (int ox = -1; ox <= 1; ox++) {
        int I = 2 * i + ox;
        if (I < 0 || I >= f.nx) continue;
        float w = (ox ? 0.5f : 1.0f) * (oy ? 0.5f : 1.0f) * (oz ? 0.5f : 1.0f);
        acc += w * src[I + (long)f.nx * (J + (long)f.ny * K)];
      }
    }
  }
  out[t] = s * acc;
}

// out[B*C][fine] (+)= s * P in[B*C][coarse]
__global__ void k_prolong_nodes(long BC, Grid3 f, Grid3 c, float s, const float* __restrict__ in,
                                float* __restrict__ out, int accumulate) {
  const long Nc = c.N(), Nf = f.N();
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= BC * Nf) return;
  const long bc = t / Nf, q = t % Nf;
  const int I = (int)(q % f.nx), J = (int)((q / f.nx) % f.ny), K = (int)(q / ((long)f.nx * f.ny));
  const float* src = in + bc * Nc;
  float acc = 0.0f;
  const int ni = (I & 1) ? 2 : 1, nj = (J & 1) ? 2 : 1, nk = (K & 1) ? 2 : 1;
  const float w = ((I & 1) ? 0.5f : 1.0f) * ((J & 1) ? 0.5f : 1.0f) * ((K & 1) ? 0.5f : 1.0f);
  for (int c3 = 0; c3 < nk; c3++)
    for (int b2 = 0; b2 < nj; b2++)
      for (int a1 = 0; a1 < ni; a1++) {
        int ci = (I >> 1) + a1, cj = (J >> 1) + b2, ck = (K >> 1) + c3;
        if (ci < c.nx && cj < c.ny && ck < c.nz) acc += src[ci + (long)c.nx * (cj + (long)c.ny * ck)];
      }
  out[t] = (accumulate ? out[t] : 0.0f) + s * w * acc;
}

__global__ void k_add_inplace(long n, const float* __restrict__ a, float* __restrict__ y) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] += a[t];
}

// loss_b = sum m (o - e)^2 / (sum m e^2 + tiny); grad = 2 m (o - e) / den / B.  One block per sample.
__global__ __launch_bounds__(256) void k_loss(int B, int Cout, long N, const float* __restrict__ o,
                                              const float* __restrict__ e, const float* __restrict__ m,
                                              float* __restrict__ grad, float* __restrict__ loss_b) {
  __shared__ float s1[256], s2[256];
  const int b = blockIdx.x;
  const long n = (long)Cout * N;
  const float* ob = o + (long)b * n;
  const float* eb = e + (long)b * n;
  const float* mb = m + (long)b * n;
  float a1 = 0.0f, a2 = 0.0f;
  for (long t = threadIdx.x; t < n; t += 256) {
    float d = ob[t] - eb[t];
    a1 += mb[t] * d * d;
    a2 += mb[t] * eb[t] * eb[t];
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  const float den = s2[0] + 1e-30f;
  if (threadIdx.x == 0) loss_b[b] = s1[0] / den;
  if (grad)
    for (long t = threadIdx.x; t < n; t += 256) grad[(long)b * n + t] = 2.0f * mb[t] * (ob[t] - eb[t]) / den / B;
}

__global__ void k_adam(long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2, float wd) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float gg = g[t] + wd * p[t];
  float a = b1 * m1[t] + (1.0f - b1) * gg;
  float v = b2 * m2[t] + (1.0f - b2) * gg * gg;
  m1[t] = a;
  m2[t] = v;
  p[t] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// inference helpers --------------------------------------------------------------------
// rms of the masked residual (single block, deterministic) -> scale[0] = rms, scale[1] = 1/rms
// one 1024-thread block: the coarse-level vectors are ~1e4-1e5 values, a single block with a
// fixed reduction tree keeps the result deterministic
__global__ __launch_bounds__(1024) void k_rms(long n, const float* __restrict__ r, float* __restrict__ scale) {
  __shared__ float sm[1024];
  float a = 0.0f;
  for (long t = threadIdx.x; t < n; t += 1024) a += r[t] * r[t];
  sm[threadIdx.x] = a;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float rms = sqrtf(sm[0] / (float)n);
    scale[0] = rms;
    scale[1] = rms > 0.0f ? 1.0f / rms : 0.0f;
  }
}
// interleaved [N][dim] residual -> channels 0..dim-1 of the input (scaled); channel dim = feature
__global__ void k_pack_input(long N, int dim, const float* __restrict__ r, const float* __restrict__ feat,
                             const float* __restrict__ scale, float* __restrict__ x) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const float s = scale[1];
  for (int d = 0; d < dim; d++) x[(long)d * N + t] = s * r[t * dim + d];
  x[(long)dim * N + t] = feat[t];
}
__global__ void k_unpack_output(long N, int dim, const float* __restrict__ y, const float* __restrict__ scale,
                                float out_scale, const uint8_t* __restrict__ bc, float* __restrict__ out) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const float s = scale[0] * out_scale;
  const uint8_t m = bc[t];
  for (int d = 0; d < dim; d++) out[t * dim + d] = ((m >> d) & 1) ? 0.0f : s * y[(long)d * N + t];
}

// ------------------------------------------------------------------------------------------
// the network
// ------------------------------------------------------------------------------------------
struct ConvSpec {
  int ci, co;
  long woff, boff;  // offsets into the flat parameter vector
};

// conv index layout for depth D: encoder 2l, 2l+1 (l = 0..D-1); decoder cd(l) = 2D + (D-2-l)
// (l = D-2..0); head 3D-1.
struct NNModel {
  int C = 32, depth = 4, cin = 4, cout = 3;
  float out_scale = 1.0f;  // 1 / reference viscosity of the level (outputs are non-dimensional)
  std::vector<Grid3> grid;
  std::vector<ConvSpec> conv;
  long nparams = 0;
  float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
  long step = 0;
  int bmax = 0;
  std::vector<float*> act, gact;    // per conv output
  std::vector<float*> pool, gpool;  // per level (l >= 1): R(e_{l-1})/8
  std::vector<float*> sum, gsum;    // per level (l <= D-2): P(up) + e_l
  float *xin = nullptr, *y = nullptr, *gy = nullptr, *lossb = nullptr;
  float *scale = nullptr, *feat = nullptr;
  int cd(int l) const { return 2 * depth + (depth - 2 - l); }
  int head() const { return 3 * depth - 1; }
  int level_of(int c) const {
    if (c < 2 * depth) return c / 2;
    if (c < 3 * depth - 1) return depth - 2 - (c - 2 * depth);
    return 0;
  }
};

template <typename T>
static T* nalloc(long n) {
  T* p = nullptr;
  UWG_HIP(hipMalloc(&p, sizeof(T) * (n > 0 ? n : 1)));
  return p;
}

static void build_specs(NNModel* m) {
  m->conv.clear();
  long off = 0;
  auto add = [&](int ci, int co) {
    ConvSpec cs{ci, co, off, 0};
    off += (long)ci * co * 27;
    cs.boff = off;
    off += co;
    m->conv.push_back(cs);
  };
  for (int l = 0; l < m->depth; l++) {
    add(l == 0 ? m->cin : m->C, m->C);
    add(m->C, m->C);
  }
  for (int l = m->depth - 2; l >= 0; l--) add(m->C, m->C);
  add(m->C, m->cout);
  m->nparams = off;
}

static void free_batch(NNModel* m) {
  for (auto v : {&m->act, &m->gact, &m->pool, &m->gpool, &m->sum, &m->gsum})
    for (auto p : *v) (void)hipFree(p);
  for (auto v : {&m->act, &m->gact, &m->pool, &m->gpool, &m->sum, &m->gsum}) v->clear();
  for (void* p : {(void*)m->xin, (void*)m->y, (void*)m->gy, (void*)m->lossb}) (void)hipFree(p);
  m->xin = m->y = m->gy = m->lossb = nullptr;
  m->bmax = 0;
}

static void ensure_batch(NNModel* m, int B) {
  if (B <= m->bmax) return;
  free_batch(m);
  const int D = m->depth;
  for (size_t c = 0; c < m->conv.size(); c++) {
    long n = (long)B * m->conv[c].co * m->grid[m->level_of((int)c)].N();
    m->act.push_back(nalloc<float>(n));
    m->gact.push_back(nalloc<float>(n));
  }
  for (int l = 0; l < D; l++) {
    long n = (long)B * m->C * m->grid[l].N();
    m->pool.push_back(l >= 1 ? nalloc<float>(n) : nullptr);
    m->gpool.push_back(l >= 1 ? nalloc<float>(n) : nullptr);
    m->sum.push_back(l <= D - 2 ? nalloc<float>(n) : nullptr);
    m->gsum.push_back(l <= D - 2 ? nalloc<float>(n) : nullptr);
  }
  m->xin = nalloc<float>((long)B * m->cin * m->grid[0].N());
  m->y = nalloc<float>((long)B * m->cout * m->grid[0].N());
  m->gy = nalloc<float>((long)B * m->cout * m->grid[0].N());
  m->lossb = nalloc<float>(B);
  m->bmax = B;
}

static void conv_fwd(NNModel* m, int c, int B, const float* in, float* out, bool relu, hipStream_t s) {
  const ConvSpec& cs = m->conv[c];
  Grid3 g = m->grid[m->level_of(c)];
  const long nb = (long)B * g.N();
  const float* w = m->params + cs.woff;
  const float* bias = m->params + cs.boff;
  if (cs.co % 4 == 0 && nb <= 65536)
    hipLaunchKernelGGL(k_conv_fwd_ks<4>, dim3(nbk(nb, 16), cs.co / 4), 256, 0, s, B, cs.ci, cs.co, g, in, w, bias, out,
                       relu ? 1 : 0);
  else if (cs.co == 3 && nb <= 65536)
    hipLaunchKernelGGL(k_conv_fwd_ks<3>, dim3(nbk(nb, 16), 1), 256, 0, s, B, cs.ci, cs.co, g, in, w, bias, out,
                       relu ? 1 : 0);
  else if (cs.co % 4 == 0)
    hipLaunchKernelGGL(k_conv_fwd_blk<4>, dim3(nbk(nb, 256), cs.co / 4), 256, 0, s, B, cs.ci, cs.co, g, in, w, bias,
                       out, relu ? 1 : 0);
  else if (cs.co == 3)
    hipLaunchKernelGGL(k_conv_fwd_blk<3>, dim3(nbk(nb, 256), 1), 256, 0, s, B, cs.ci, cs.co, g, in, w, bias, out,
                       relu ? 1 : 0);
  else
    hipLaunchKernelGGL(k_conv_fwd, nbk(nb * cs.co, 256), 256, 0, s, B, cs.ci, cs.co, g, in, w, bias, out,
                       relu ? 1 : 0);
}

// forward for a batch packed in m->xin; output in m->y
static void nn_forward(NNModel* m, int B, hipStream_t s) {
  const int D = m->depth, C = m->C;
  for (int l = 0; l < D; l++) {
    const float* in = m->xin;
    if (l > 0) {
      Grid3 g = m->grid[l];
      hipLaunchKernelGGL(k_restrict_nodes, nbk((long)B * C * g.N(), 256), 256, 0, s, (long)B * C, m->grid[l - 1], g,
                         0.125f, m->act[2 * l - 1], m->pool[l]);
      in = m->pool[l];
    }
    conv_fwd(m, 2 * l, B, in, m->act[2 * l], true, s);
    conv_fwd(m, 2 * l + 1, B, m->act[2 * l], m->act[2 * l + 1], true, s);
  }
  const float* up = m->act[2 * D - 1];
  for (int l = D - 2; l >= 0; l--) {
    Grid3 g = m->grid[l];
    long n = (long)B * C * g.N();
    hipLaunchKernelGGL(k_prolong_nodes, nbk(n, 256), 256, 0, s, (long)B * C, g, m->grid[l + 1], 1.0f, up, m->sum[l], 0);
    hipLaunchKernelGGL(k_add_inplace, nbk(n, 256), 256, 0, s, n, m->act[2 * l + 1], m->sum[l]);
    conv_fwd(m, m->cd(l), B, m->sum[l], m->act[m->cd(l)], true, s);
    up = m->act[m->cd(l)];
  }
  conv_fwd(m, m->head(), B, up, m->y, false, s);
  UWG_HIP(hipGetLastError());
}

// conv backward: relu mask (if outact), weight and bias gradients, input gradient (if gin)
static void conv_bwd(NNModel* m, int c, int B, const float* in, float* gout, float* gin, const float* outact,
                     hipStream_t s) {
  const ConvSpec& cs = m->conv[c];
  Grid3 g = m->grid[m->level_of(c)];
  if (outact) {
    long n = (long)B * cs.co * g.N();
    hipLaunchKernelGGL(k_relu_bwd, nbk(n, 256), 256, 0, s, n, outact, gout);
  }
  hipLaunchKernelGGL(k_conv_bwd_w, cs.co * cs.ci, 256, 0, s, B, cs.ci, cs.co, g, gout, in, m->grads + cs.woff);
  hipLaunchKernelGGL(k_bias_grad, cs.co, 256, 0, s, B, cs.co, g.N(), gout, m->grads + cs.boff);
  if (gin) {
    const long nb = (long)B * g.N();
    if (cs.ci % 4 == 0)
      hipLaunchKernelGGL(k_conv_bwd_data_blk<4>, dim3(nbk(nb, 256), cs.ci / 4), 256, 0, s, B, cs.ci, cs.co, g, gout,
                         m->params + cs.woff, gin);
    else
      hipLaunchKernelGGL(k_conv_bwd_data, nbk(nb * cs.ci, 256), 256, 0, s, B, cs.ci, cs.co, g, gout,
                         m->params + cs.woff, gin);
  }
}

// backward from m->gy (dL/dy); parameter gradients written into m->grads
static void nn_backward(NNModel* m, int B, hipStream_t s) {
  const int D = m->depth, C = m->C;
  // head
  conv_bwd(m, m->head(), B, m->act[m->cd(0)], m->gy, m->gact[m->cd(0)], nullptr, s);
  // decoder, shallow to deep
  for (int l = 0; l <= D - 2; l++) {
    const int c = m->cd(l);
    conv_bwd(m, c, B, m->sum[l], m->gact[c], m->gsum[l], m->act[c], s);
    // sum_l = P(up) + e_l : g_up = R g_sum  (up = next decoder output, or the encoder top)
    const int cup = (l == D - 2) ? 2 * D - 1 : m->cd(l + 1);
    Grid3 gc = m->grid[l + 1];
    hipLaunchKernelGGL(k_restrict_nodes, nbk((long)B * C * gc.N(), 256), 256, 0, s, (long)B * C, m->grid[l], gc, 1.0f,
                       m->gsum[l], m->gact[cup]);
  }
  // encoder, deep to shallow
  for (int l = D - 1; l >= 0; l--) {
    Grid3 g = m->grid[l];
    const long n = (long)B * C * g.N();
    const int c1 = 2 * l + 1, c0 = 2 * l;
    if (l <= D - 2) {
      // g_e_l = g_sum_l (skip) + P(g_pool_{l+1}) / 8 (pooling into the next level)
      hipLaunchKernelGGL(k_prolong_nodes, nbk(n, 256), 256, 0, s, (long)B * C, g, m->grid[l + 1], 0.125f,
                         m->gpool[l + 1], m->gact[c1], 0);
      hipLaunchKernelGGL(k_add_inplace, nbk(n, 256), 256, 0, s, n, m->gsum[l], m->gact[c1]);
    }
    conv_bwd(m, c1, B, m->act[c0], m->gact[c1], m->gact[c0], m->act[c1], s);
    conv_bwd(m, c0, B, l == 0 ? m->xin : m->pool[l], m->gact[c0], l == 0 ? nullptr : m->gpool[l], m->act[c0], s);
  }
  UWG_HIP(hipGetLastError());
}

static NNModel* nn_alloc(Grid3 g0, int channels, int depth, int cin, int cout) {
  NNModel* m = new NNModel();
  m->C = channels;
  m->depth = depth;
  m->cin = cin;
  m->cout = cout;
  m->grid.push_back(g0);
  for (int l = 1; l < depth; l++) m->grid.push_back(coarser(m->grid.back()));
  build_specs(m);
  m->params = nalloc<float>(m->nparams);
  m->grads = nalloc<float>(m->nparams);
  m->m1 = nalloc<float>(m->nparams);
  m->m2 = nalloc<float>(m->nparams);
  UWG_HIP(hipMemset(m->m1, 0, sizeof(float) * m->nparams));
  UWG_HIP(hipMemset(m->m2, 0, sizeof(float) * m->nparams));
  m->scale = nalloc<float>(2);
  m->feat = nalloc<float>(g0.N());
  UWG_HIP(hipMemset(m->feat, 0, sizeof(float) * g0.N()));
  return m;
}

void nn_destroy(NNModel* m) {
  if (!m) return;
  free_batch(m);
  for (void* p : {(void*)m->params, (void*)m->grads, (void*)m->m1, (void*)m->m2, (void*)m->scale, (void*)m->feat})
    (void)hipFree(p);
  delete m;
}

NNModel* nn_create(const Geom& g, int channels, int depth, const float* params, int nparams) {
  if (g.dim != 3) throw std::runtime_error("the learned correction is provided for 3D meshes");
  if (depth < 2 || depth > 6 || channels < 1 || channels > 128) throw std::runtime_error("bad network shape");
  NNModel* m = nn_alloc(Grid3{g.nn[0], g.nn[1], g.nn[2]}, channels, depth, 4, 3);
  if (nparams != m->nparams) {
    long want = m->nparams;
    nn_destroy(m);
    throw std::runtime_error("learned model has " + std::to_string(nparams) + " parameters, network needs " +
                             std::to_string(want));
  }
  UWG_HIP(hipMemcpy(m->params, params, sizeof(float) * nparams, hipMemcpyHostToDevice));
  ensure_batch(m, 1);
  return m;
}

void nn_set_feature(NNModel* m, const float* feat_host, double eta_ref) {
  m->out_scale = (float)(1.0 / eta_ref);
  UWG_HIP(hipMemcpy(m->feat, feat_host, sizeof(float) * m->grid[0].N(), hipMemcpyHostToDevice));
}

// Residual-minimising step length for a correction d of K x = b with residual r:
// omega = (r . Kd) / (Kd . Kd), clamped to [0, 2] (a prediction pointing the wrong way, or a
// non-finite one, is dropped); then x = (accumulate ? x : 0) + omega d.  Single block, FP64 sums, deterministic.
__global__ __launch_bounds__(1024) void k_line_search(long n, const float* __restrict__ r, const float* __restrict__ kd,
                                                      const float* __restrict__ d, float* __restrict__ x,
                                                      int accumulate) {
  __shared__ double s1[1024], s2[1024];
  double a1 = 0.0, a2 = 0.0;
  for (long t = threadIdx.x; t < n; t += 1024) {
    a1 += (double)r[t] * kd[t];
    a2 += (double)kd[t] * kd[t];
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s1[threadIdx.x] += s1[threadIdx.x + st];
      s2[threadIdx.x] += s2[threadIdx.x + st];
    }
    __syncthreads();
  }
  double om = s2[0] > 0.0 ? s1[0] / s2[0] : 0.0;
  if (!(om > 0.0)) om = 0.0;  // also catches a non-finite prediction
  if (om > 2.0) om = 2.0;
  const float w = (float)om;
  for (long t = threadIdx.x; t < n; t += 1024) x[t] = (accumulate ? x[t] : 0.0f) + w * d[t];
}

void nn_line_search(long n, const float* r, const float* kd, const float* d, float* x, bool accumulate,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_line_search, 1, 1024, 0, s, n, r, kd, d, x, accumulate ? 1 : 0);
  UWG_HIP(hipGetLastError());
}

void nn_coarse_correction(NNModel* m, const Geom& g, const float* r, float* x, const uint8_t* bc, hipStream_t s) {
  const long N = m->grid[0].N();
  hipLaunchKernelGGL(k_rms, 1, 1024, 0, s, N * 3, r, m->scale);
  hipLaunchKernelGGL(k_pack_input, nbk(N, 256), 256, 0, s, N, 3, r, m->feat, m->scale, m->xin);
  nn_forward(m, 1, s);
  hipLaunchKernelGGL(k_unpack_output, nbk(N, 256), 256, 0, s, N, 3, m->y, m->scale, m->out_scale, bc, x);
  UWG_HIP(hipGetLastError());
  (void)g;
}

long nn_param_count(NNModel* m) { return m->nparams; }

// ==========================================================================================
// Learned Schur-complement stencil.
//
// The Schur complement S = G^T K^-1 G of Q1/dQ0 is preconditioned in the reference by the
// diagonal approxS.  The modes approxS leaves with small eigenvalues are oscillatory on the
// element scale, where a 27-point stencil around each element can act on them; so the network
// maps local features of the element grid to a per-element 27-point stencil C(e), and the
// preconditioner is M = D^-1/2 C D^-1/2 with D = approxS (C = identity reproduces approxS^-1).
//
// Features per element (9 channels) from t = -ln(approxS_e), which is ln(eta_e) up to the
// element geometry: channels 0..5 tanh((t_nb - t_e) / 2) for the face neighbours x-, x+, y-,
// y+, z-, z+ (0 where there is none), channels 6..8 -1 on the lower wall layer, +1 on the
// upper wall layer, 0 elsewhere (x, y, z).  Network: 3x3x3 convolutions with ReLU, C -> C
// (L layers, the first 9 -> C), then a 1x1x1 head C -> 27; the stencil is identity + head.
// Trained on (features, r, e) with r = S e (scaled: r^ = D^-1/2 r, e^ = D^1/2 e), loss
// |C r^ - e^|^2 / |e^|^2.  FP32 storage and arithmetic, deterministic reductions.
// ==========================================================================================
__device__ __forceinline__ void grid_ijk(const Grid3& g, long e, int& i, int& j, int& k) {
  i = (int)(e % g.nx);
  j = (int)((e / g.nx) % g.ny);
  k = (int)(e / ((long)g.nx * g.ny));
}

constexpr int kSnnFeat = 9;  // input channels of the Schur-stencil network

__global__ void k_snn_features(Grid3 g, const double* __restrict__ S, float* __restrict__ feat) {
  const long N = g.N();
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N) return;
  int i, j, k;
  grid_ijk(g, e, i, j, k);
  const double t = -log(S[e]);
  const int ii[6] = {i - 1, i + 1, i, i, i, i}, jj[6] = {j, j, j - 1, j + 1, j, j}, kk[6] = {k, k, k, k, k - 1, k + 1};
#pragma unroll
  for (int c = 0; c < 6; c++) {
    float v = 0.0f;
    if (ii[c] >= 0 && ii[c] < g.nx && jj[c] >= 0 && jj[c] < g.ny && kk[c] >= 0 && kk[c] < g.nz) {
      const double tn = -log(S[ii[c] + (long)g.nx * (jj[c] + (long)g.ny * kk[c])]);
      v = tanhf((float)(0.5 * (tn - t)));
    }
    feat[c * N + e] = v;
  }
  feat[6 * N + e] = i == 0 ? -1.0f : (i == g.nx - 1 ? 1.0f : 0.0f);
  feat[7 * N + e] = j == 0 ? -1.0f : (j == g.ny - 1 ? 1.0f : 0.0f);
  feat[8 * N + e] = k == 0 ? -1.0f : (k == g.nz - 1 ? 1.0f : 0.0f);
}

// 1x1x1 convolution: out[b][co][p] = bias[co] + sum_ci w[co][ci] in[b][ci][p]; COB output
// channels per thread (weights wave-uniform)
template <int COB>
__global__ __launch_bounds__(256) void k_conv1_fwd(int B, int Ci, int Co, long N, const float* __restrict__ in,
                                                   const float* __restrict__ w, const float* __restrict__ bias,
                                                   float* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int co0 = blockIdx.y * COB;
  if (t >= (long)B * N) return;
  const long p = t % N;
  const int b = (int)(t / N);
  float acc[COB];
#pragma unroll
  for (int c = 0; c < COB; c++) acc[c] = bias[co0 + c];
  for (int ci = 0; ci < Ci; ci++) {
    const float v = in[((long)b * Ci + ci) * N + p];
#pragma unroll
    for (int c = 0; c < COB; c++) acc[c] += w[(long)(co0 + c) * Ci + ci] * v;
  }
#pragma unroll
  for (int c = 0; c < COB; c++) out[((long)b * Co + co0 + c) * N + p] = acc[c];
}

// gin[b][ci][p] = sum_co w[co][ci] gout[b][co][p]
__global__ void k_conv1_bwd_data(int B, int Ci, int Co, long N, const float* __restrict__ gout,
                                 const float* __restrict__ w, float* __restrict__ gin) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)B * N) return;
  const long p = t % N;
  const int b = (int)(t / N);
  for (int ci = 0; ci < Ci; ci++) {
    float a = 0.0f;
    for (int co = 0; co < Co; co++) a += w[(long)co * Ci + ci] * gout[((long)b * Co + co) * N + p];
    gin[((long)b * Ci + ci) * N + p] = a;
  }
}

// dw[co][ci] = sum_b sum_p gout[b][co][p] in[b][ci][p]; one block per (co, ci)
__global__ __launch_bounds__(256) void k_conv1_bwd_w(int B, int Ci, int Co, long N, const float* __restrict__ gout,
                                                     const float* __restrict__ in, float* __restrict__ dw) {
  __shared__ float sm[256];
  const int co = blockIdx.x / Ci, ci = blockIdx.x % Ci;
  float a = 0.0f;
  for (long t = threadIdx.x; t < (long)B * N; t += 256) {
    const int b = (int)(t / N);
    const long p = t % N;
    a += gout[((long)b * Co + co) * N + p] * in[((long)b * Ci + ci) * N + p];
  }
  sm[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) dw[(long)co * Ci + ci] = sm[0];
}

// out[b][e] = sum_s (delta_s,13 + y[b][s][e]) r[b][e + off(s)]  (0 outside the grid);
// diff = out - t; per-block partial sums of diff^2 and t^2 in FP64 (blockIdx.y = sample)
__global__ __launch_bounds__(256) void k_snn_stencil_diff(Grid3 g, const float* __restrict__ y,
                                                          const float* __restrict__ r, const float* __restrict__ t,
                                                          float* __restrict__ diff, double* __restrict__ part) {
  __shared__ double s1[256], s2[256];
  const long N = g.N();
  const int b = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double a1 = 0.0, a2 = 0.0;
  if (e < N) {
    int i, j, k;
    grid_ijk(g, e, i, j, k);
    const float* yb = y + (long)b * 27 * N;
    const float* rb = r + (long)b * N;
    float o = 0.0f;
#pragma unroll
    for (int s = 0; s < 27; s++) {
      const int ii = i + s % 3 - 1, jj = j + (s / 3) % 3 - 1, kk = k + s / 9 - 1;
      if (ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz)
        o += ((s == 13 ? 1.0f : 0.0f) + yb[s * N + e]) * rb[ii + (long)g.nx * (jj + (long)g.ny * kk)];
    }
    const float tv = t[(long)b * N + e];
    const float d = o - tv;
    diff[(long)b * N + e] = d;
    a1 = (double)d * d;
    a2 = (double)tv * tv;
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s1[threadIdx.x] += s1[threadIdx.x + st];
      s2[threadIdx.x] += s2[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((long)b * gridDim.x + blockIdx.x) * 2] = s1[0];
    part[((long)b * gridDim.x + blockIdx.x) * 2 + 1] = s2[0];
  }
}

// per sample: loss_b = sum diff^2 / sum t^2; gscale_b = 2 / (sum t^2 B)  (one block per sample)
__global__ __launch_bounds__(256) void k_snn_loss_reduce(int B, int nblk, const double* __restrict__ part,
                                                         float* __restrict__ loss_b, float* __restrict__ gscale) {
  __shared__ double s1[256], s2[256];
  const int b = blockIdx.x;
  double a1 = 0.0, a2 = 0.0;
  for (int q = threadIdx.x; q < nblk; q += 256) {
    a1 += part[((long)b * nblk + q) * 2];
    a2 += part[((long)b * nblk + q) * 2 + 1];
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s1[threadIdx.x] += s1[threadIdx.x + st];
      s2[threadIdx.x] += s2[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double den = s2[0] + 1e-300;
    loss_b[b] = (float)(s1[0] / den);
    gscale[b] = (float)(2.0 / (den * B));
  }
}

// dL/dy[b][s][e] = gscale_b diff[b][e] r[b][e + off(s)]
__global__ void k_snn_stencil_grad(int B, Grid3 g, const float* __restrict__ diff, const float* __restrict__ r,
                                   const float* __restrict__ gscale, float* __restrict__ gy) {
  const long N = g.N();
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)B * N) return;
  const long e = t % N;
  const int b = (int)(t / N);
  int i, j, k;
  grid_ijk(g, e, i, j, k);
  const float d = gscale[b] * diff[t];
  const float* rb = r + (long)b * N;
#pragma unroll
  for (int s = 0; s < 27; s++) {
    const int ii = i + s % 3 - 1, jj = j + (s / 3) % 3 - 1, kk = k + s / 9 - 1;
    const bool in = ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz;
    gy[((long)b * 27 + s) * N + e] = in ? d * rb[ii + (long)g.nx * (jj + (long)g.ny * kk)] : 0.0f;
  }
}

// bfloat16 (round to nearest, ties to even) of a finite float, and back
__device__ __forceinline__ uint16_t snn_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float snn_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// preconditioner coefficients a[s][e] = -(delta_s,13 + y[s][e]) / sqrt(S_e S_e+off(s)) (0 outside),
// stored as bfloat16: z = sum_s a_s q_e+s is -M q (the Schur block's sign convention, S
// approximated by -approxS); a preconditioner's coefficients need no more than bf16's 8 bits
__global__ void k_snn_coef(Grid3 g, const float* __restrict__ y, const double* __restrict__ S,
                           uint16_t* __restrict__ a) {
  const long N = g.N();
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N) return;
  int i, j, k;
  grid_ijk(g, e, i, j, k);
  const double de = 1.0 / sqrt(S[e]);
#pragma unroll
  for (int s = 0; s < 27; s++) {
    const int ii = i + s % 3 - 1, jj = j + (s / 3) % 3 - 1, kk = k + s / 9 - 1;
    float v = 0.0f;
    if (ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz) {
      const double dn = 1.0 / sqrt(S[ii + (long)g.nx * (jj + (long)g.ny * kk)]);
      v = (float)(-(double)((s == 13 ? 1.0f : 0.0f) + y[s * N + e]) * de * dn);
    }
    a[s * N + e] = snn_bf16(v);
  }
}

// z_e = sum_s a[s][e] q_e+off(s)  (T = double for the FP64 Krylov iteration, float for FP32).
// One thread per element, x fastest: the 27 coefficient planes stream coalesced; the
// neighbour values come through the caches.  Out-of-grid taps have zero coefficients and read
// the element itself (a valid address), so no tap is branched on.  32-bit indices (the element
// count is below 2^31 on every grid this is used on).
template <typename T>
__global__ __launch_bounds__(256) void k_snn_apply(Grid3 g, const uint16_t* __restrict__ a, const T* __restrict__ q,
                                                   T* __restrict__ z) {
  const int N = (int)g.N();
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= N) return;
  const int nx = g.nx, nxy = g.nx * g.ny;
  const int i = e % nx, j = (e / nx) % g.ny, k = e / nxy;
  const int dxm = i > 0 ? -1 : 0, dxp = i < nx - 1 ? 1 : 0;
  const int dym = j > 0 ? -nx : 0, dyp = j < g.ny - 1 ? nx : 0;
  const int dzm = k > 0 ? -nxy : 0, dzp = k < g.nz - 1 ? nxy : 0;
  const int ox[3] = {dxm, 0, dxp}, oy[3] = {dym, 0, dyp}, oz[3] = {dzm, 0, dzp};
  uint16_t c[27];
#pragma unroll
  for (int s = 0; s < 27; s++) c[s] = a[s * N + e];
  T acc = (T)0;
#pragma unroll
  for (int s = 0; s < 27; s++) acc += (T)snn_f32(c[s]) * q[e + ox[s % 3] + oy[(s / 3) % 3] + oz[s / 9]];
  z[e] = acc;
}

struct SchurNet {
  int C = 16, L = 2;
  long nparams = 0;
  std::vector<long> woff, boff;  // per layer (L 3x3x3 convs, then the head)
  float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
  long step = 0;
  // activations for up to bmax samples of n elements
  long ncap = 0;
  int bmax = 0;
  std::vector<float*> act, gact;
  float *xin = nullptr, *y = nullptr, *gy = nullptr, *diff = nullptr, *lossb = nullptr, *gscale = nullptr;
  double* part = nullptr;
  // training set (all samples on one grid)
  Grid3 dg{0, 0, 0};
  long nsamples = 0, scap = 0;
  float *dx = nullptr, *dr = nullptr, *dt = nullptr;
};

static void snn_specs(SchurNet* m) {
  m->woff.clear();
  m->boff.clear();
  long off = 0;
  for (int l = 0; l < m->L; l++) {
    const int ci = l == 0 ? kSnnFeat : m->C;
    m->woff.push_back(off);
    off += (long)ci * m->C * 27;
    m->boff.push_back(off);
    off += m->C;
  }
  m->woff.push_back(off);
  off += (long)m->C * 27;
  m->boff.push_back(off);
  off += 27;
  m->nparams = off;
}

static void snn_free_batch(SchurNet* m) {
  for (auto v : {&m->act, &m->gact})
    for (auto p : *v) (void)hipFree(p);
  m->act.clear();
  m->gact.clear();
  for (void* p : {(void*)m->xin, (void*)m->y, (void*)m->gy, (void*)m->diff, (void*)m->lossb, (void*)m->gscale,
                  (void*)m->part})
    (void)hipFree(p);
  m->xin = m->y = m->gy = m->diff = m->lossb = m->gscale = nullptr;
  m->part = nullptr;
  m->bmax = 0;
  m->ncap = 0;
}

static void snn_ensure(SchurNet* m, int B, long N, bool train) {
  if (B <= m->bmax && N <= m->ncap && (!train || m->gy)) return;
  snn_free_batch(m);
  const long n = (long)B * N;
  for (int l = 0; l < m->L; l++) {
    m->act.push_back(nalloc<float>(n * m->C));
    m->gact.push_back(train ? nalloc<float>(n * m->C) : nullptr);
  }
  m->xin = nalloc<float>(n * kSnnFeat);
  m->y = nalloc<float>(n * 27);
  if (train) {
    m->gy = nalloc<float>(n * 27);
    m->diff = nalloc<float>(n);
    m->part = nalloc<double>((long)B * nbk(N, 256) * 2);
  }
  m->lossb = nalloc<float>(B);
  m->gscale = nalloc<float>(B);
  m->bmax = B;
  m->ncap = N;
}

static void conv3_fwd(int B, int ci, int co, Grid3 g, const float* in, const float* w, const float* bias, float* out,
                      bool relu, hipStream_t s) {
  const long nb = (long)B * g.N();
  if (co % 16 == 0)
    hipLaunchKernelGGL(k_conv_fwd_blk<16>, dim3(nbk(nb, 256), co / 16), 256, 0, s, B, ci, co, g, in, w, bias, out,
                       relu ? 1 : 0);
  else if (co % 8 == 0)
    hipLaunchKernelGGL(k_conv_fwd_blk<8>, dim3(nbk(nb, 256), co / 8), 256, 0, s, B, ci, co, g, in, w, bias, out,
                       relu ? 1 : 0);
  else if (co % 4 == 0)
    hipLaunchKernelGGL(k_conv_fwd_blk<4>, dim3(nbk(nb, 256), co / 4), 256, 0, s, B, ci, co, g, in, w, bias, out,
                       relu ? 1 : 0);
  else
    hipLaunchKernelGGL(k_conv_fwd, nbk(nb * co, 256), 256, 0, s, B, ci, co, g, in, w, bias, out, relu ? 1 : 0);
}

// y[B][27][N] from xin[B][9][N]
static void snn_forward(SchurNet* m, int B, Grid3 g, hipStream_t s) {
  const float* in = m->xin;
  for (int l = 0; l < m->L; l++) {
    conv3_fwd(B, l == 0 ? kSnnFeat : m->C, m->C, g, in, m->params + m->woff[l], m->params + m->boff[l], m->act[l], true,
              s);
    in = m->act[l];
  }
  const long nb = (long)B * g.N();
  hipLaunchKernelGGL(k_conv1_fwd<9>, dim3(nbk(nb, 256), 3), 256, 0, s, B, m->C, 27, g.N(), in,
                     m->params + m->woff[m->L], m->params + m->boff[m->L], m->y);
  UWG_HIP(hipGetLastError());
}

// loss of the batch (xin, r, t) and, if grads, parameter gradients into m->grads
static void snn_loss_backward(SchurNet* m, int B, Grid3 g, const float* r, const float* t, bool grads,
                              hipStream_t s) {
  const long N = g.N(), nb = (long)B * N;
  const int nblk = (int)nbk(N, 256);
  hipLaunchKernelGGL(k_snn_stencil_diff, dim3(nblk, B), 256, 0, s, g, m->y, r, t, m->diff, m->part);
  hipLaunchKernelGGL(k_snn_loss_reduce, B, 256, 0, s, B, nblk, m->part, m->lossb, m->gscale);
  if (!grads) {
    UWG_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_snn_stencil_grad, nbk(nb, 256), 256, 0, s, B, g, m->diff, r, m->gscale, m->gy);
  // head (1x1x1)
  const int L = m->L, C = m->C;
  hipLaunchKernelGGL(k_conv1_bwd_w, 27 * C, 256, 0, s, B, C, 27, N, m->gy, m->act[L - 1], m->grads + m->woff[L]);
  hipLaunchKernelGGL(k_bias_grad, 27, 256, 0, s, B, 27, N, m->gy, m->grads + m->boff[L]);
  hipLaunchKernelGGL(k_conv1_bwd_data, nbk(nb, 256), 256, 0, s, B, C, 27, N, m->gy, m->params + m->woff[L],
                     m->gact[L - 1]);
  for (int l = L - 1; l >= 0; l--) {
    const int ci = l == 0 ? kSnnFeat : C;
    hipLaunchKernelGGL(k_relu_bwd, nbk(nb * C, 256), 256, 0, s, nb * C, m->act[l], m->gact[l]);
    const float* in = l == 0 ? m->xin : m->act[l - 1];
    hipLaunchKernelGGL(k_conv_bwd_w, C * ci, 256, 0, s, B, ci, C, g, m->gact[l], in, m->grads + m->woff[l]);
    hipLaunchKernelGGL(k_bias_grad, C, 256, 0, s, B, C, N, m->gact[l], m->grads + m->boff[l]);
    if (l > 0) {
      if (ci % 4 == 0)
        hipLaunchKernelGGL(k_conv_bwd_data_blk<4>, dim3(nbk(nb, 256), ci / 4), 256, 0, s, B, ci, C, g, m->gact[l],
                           m->params + m->woff[l], m->gact[l - 1]);
      else
        hipLaunchKernelGGL(k_conv_bwd_data, nbk(nb * ci, 256), 256, 0, s, B, ci, C, g, m->gact[l],
                           m->params + m->woff[l], m->gact[l - 1]);
    }
  }
  UWG_HIP(hipGetLastError());
}

static SchurNet* snn_alloc(int C, int L) {
  if (C < 1 || C > 128 || C % 8 != 0 || L < 1 || L > 6) throw std::runtime_error("bad Schur-stencil network shape");
  SchurNet* m = new SchurNet();
  m->C = C;
  m->L = L;
  snn_specs(m);
  m->params = nalloc<float>(m->nparams);
  m->grads = nalloc<float>(m->nparams);
  m->m1 = nalloc<float>(m->nparams);
  m->m2 = nalloc<float>(m->nparams);
  UWG_HIP(hipMemset(m->grads, 0, sizeof(float) * m->nparams));
  UWG_HIP(hipMemset(m->m1, 0, sizeof(float) * m->nparams));
  UWG_HIP(hipMemset(m->m2, 0, sizeof(float) * m->nparams));
  return m;
}

void schur_net_destroy(SchurNet* m) {
  if (!m) return;
  snn_free_batch(m);
  for (void* p : {(void*)m->params, (void*)m->grads, (void*)m->m1, (void*)m->m2, (void*)m->dx, (void*)m->dr,
                  (void*)m->dt})
    (void)hipFree(p);
  delete m;
}

SchurNet* schur_net_create(int C, int L, const float* params, long nparams) {
  SchurNet* m = snn_alloc(C, L);
  if (nparams != m->nparams) {
    const long want = m->nparams;
    schur_net_destroy(m);
    throw std::runtime_error("Schur-stencil model has " + std::to_string(nparams) + " parameters, the network needs " +
                             std::to_string(want));
  }
  UWG_HIP(hipMemcpy(m->params, params, sizeof(float) * nparams, hipMemcpyHostToDevice));
  return m;
}

void schur_net_features(const Geom& g, const double* approxS, float* feat, hipStream_t s) {
  const Grid3 eg{g.res[0], g.res[1], g.res[2]};
  hipLaunchKernelGGL(k_snn_features, nbk(eg.N(), 256), 256, 0, s, eg, approxS, feat);
  UWG_HIP(hipGetLastError());
}

void schur_net_coefficients(SchurNet* m, const Geom& g, const double* approxS, uint16_t* a, hipStream_t s) {
  if (g.dim != 3) throw std::runtime_error("the learned Schur stencil is provided for 3D meshes");
  const Grid3 eg{g.res[0], g.res[1], g.res[2]};
  const long N = eg.N();
  snn_ensure(m, 1, N, false);
  schur_net_features(g, approxS, m->xin, s);
  snn_forward(m, 1, eg, s);
  hipLaunchKernelGGL(k_snn_coef, nbk(N, 256), 256, 0, s, eg, m->y, approxS, a);
  UWG_HIP(hipGetLastError());
}

template <typename T>
void schur_stencil_apply(const Geom& g, const uint16_t* a, const T* q, T* z, hipStream_t s) {
  const Grid3 eg{g.res[0], g.res[1], g.res[2]};
  if (eg.N() >= (1L << 31) / 27) throw std::runtime_error("element grid too large for the stencil apply");
  hipLaunchKernelGGL(k_snn_apply<T>, nbk(eg.N(), 256), 256, 0, s, eg, a, q, z);
  UWG_HIP(hipGetLastError());
}
template void schur_stencil_apply<double>(const Geom&, const uint16_t*, const double*, double*, hipStream_t);
template void schur_stencil_apply<float>(const Geom&, const uint16_t*, const float*, float*, hipStream_t);

// training set: append one sample (features from approxS, r^ and e^ given on the device, FP64)
__global__ void k_d2f_copy(long n, const double* __restrict__ a, float* __restrict__ b) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}

void schur_net_append(SchurNet* m, const Geom& g, const double* approxS, const double* rhat, const double* ehat) {
  const Grid3 eg{g.res[0], g.res[1], g.res[2]};
  const long N = eg.N();
  if (m->nsamples > 0 && (eg.nx != m->dg.nx || eg.ny != m->dg.ny || eg.nz != m->dg.nz))
    throw std::runtime_error("all training samples of a Schur-stencil set must share one element grid");
  if (m->nsamples == m->scap) {
    const long cap = std::max<long>(8, m->scap * 2);
    float* nx_ = nalloc<float>(cap * kSnnFeat * N);
    float* nr = nalloc<float>(cap * N);
    float* nt = nalloc<float>(cap * N);
    if (m->nsamples) {
      UWG_HIP(hipMemcpy(nx_, m->dx, sizeof(float) * m->nsamples * kSnnFeat * N, hipMemcpyDeviceToDevice));
      UWG_HIP(hipMemcpy(nr, m->dr, sizeof(float) * m->nsamples * N, hipMemcpyDeviceToDevice));
      UWG_HIP(hipMemcpy(nt, m->dt, sizeof(float) * m->nsamples * N, hipMemcpyDeviceToDevice));
    }
    for (void* p : {(void*)m->dx, (void*)m->dr, (void*)m->dt}) (void)hipFree(p);
    m->dx = nx_;
    m->dr = nr;
    m->dt = nt;
    m->scap = cap;
  }
  m->dg = eg;
  const long k = m->nsamples;
  schur_net_features(g, approxS, m->dx + k * kSnnFeat * N, 0);
  hipLaunchKernelGGL(k_d2f_copy, nbk(N, 256), 256, 0, 0, N, rhat, m->dr + k * N);
  hipLaunchKernelGGL(k_d2f_copy, nbk(N, 256), 256, 0, 0, N, ehat, m->dt + k * N);
  UWG_HIP(hipGetLastError());
  UWG_HIP(hipDeviceSynchronize());
  m->nsamples++;
}

}  // namespace uwg

// ------------------------------------------------------------------------------------------
// C ABI for training and testing the network
// ------------------------------------------------------------------------------------------
#include "../../include/uwg.h"

using namespace uwg;
struct uwg_nn {
  NNModel* m;
  float* data_x = nullptr;
  float* data_t = nullptr;
  float* data_m = nullptr;
  long nsamples = 0;
};

#define NN_API_BEGIN try {
#define NN_API_END                      \
  }                                     \
  catch (const std::exception& ex) {    \
    uwg::set_error(ex.what());          \
    return UWG_ERR_DEVICE;              \
  }

extern "C" int uwg_nn_create(int nx, int ny, int nz, int channels, int depth, long seed, uwg_nn** out) {
  if (nx < 3 || ny < 3 || nz < 3 || depth < 2 || depth > 6 || channels < 1 || channels > 128) {
    uwg::set_error("bad network shape");
    return UWG_ERR_ARG;
  }
  NN_API_BEGIN
  uwg_nn* h = new uwg_nn();
  h->m = nn_alloc(Grid3{nx, ny, nz}, channels, depth, 4, 3);
  // He-normal initialisation from a 64-bit LCG (deterministic for a seed); small head
  std::vector<float> p(h->m->nparams, 0.0f);
  unsigned long long st = (unsigned long long)seed * 6364136223846793005ULL + 1442695040888963407ULL;
  auto unif = [&]() {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    return ((st >> 11) + 0.5) / 9007199254740992.0;
  };
  for (size_t c = 0; c < h->m->conv.size(); c++) {
    const ConvSpec& cs = h->m->conv[c];
    double sd = sqrt(2.0 / (cs.ci * 27.0)) * (c + 1 == h->m->conv.size() ? 0.1 : 1.0);
    for (long k = 0; k < (long)cs.ci * cs.co * 27; k++) {
      double u1 = unif(), u2 = unif();
      p[cs.woff + k] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
  }
  UWG_HIP(hipMemcpy(h->m->params, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice));
  *out = h;
  return 0;
  NN_API_END
}

extern "C" int uwg_nn_destroy(uwg_nn* h) {
  if (!h) return 0;
  nn_destroy(h->m);
  for (void* p : {(void*)h->data_x, (void*)h->data_t, (void*)h->data_m}) (void)hipFree(p);
  delete h;
  return 0;
}

extern "C" int uwg_nn_param_count(uwg_nn* h, long* n) {
  *n = h->m->nparams;
  return 0;
}
extern "C" int uwg_nn_get_params(uwg_nn* h, float* out) {
  NN_API_BEGIN
  UWG_HIP(hipMemcpy(out, h->m->params, sizeof(float) * h->m->nparams, hipMemcpyDeviceToHost));
  return 0;
  NN_API_END
}
extern "C" int uwg_nn_set_params(uwg_nn* h, const float* in) {
  NN_API_BEGIN
  UWG_HIP(hipMemcpy(h->m->params, in, sizeof(float) * h->m->nparams, hipMemcpyHostToDevice));
  return 0;
  NN_API_END
}

// y[B][3][N] = NN(x[B][4][N])
extern "C" int uwg_nn_forward(uwg_nn* h, int B, const float* x, float* y) {
  NN_API_BEGIN
  NNModel* m = h->m;
  ensure_batch(m, B);
  const long N = m->grid[0].N();
  UWG_HIP(hipMemcpy(m->xin, x, sizeof(float) * B * m->cin * N, hipMemcpyHostToDevice));
  nn_forward(m, B, 0);
  UWG_HIP(hipMemcpy(y, m->y, sizeof(float) * B * m->cout * N, hipMemcpyDeviceToHost));
  return 0;
  NN_API_END
}

// mean over the batch of |m (y - t)|^2 / |m t|^2 and its gradient w.r.t. the parameters
extern "C" int uwg_nn_loss_grad(uwg_nn* h, int B, const float* x, const float* t, const float* mask, float* loss,
                                float* grads) {
  NN_API_BEGIN
  NNModel* m = h->m;
  ensure_batch(m, B);
  const long N = m->grid[0].N();
  float* dt = nalloc<float>((long)B * m->cout * N);
  float* dm = nalloc<float>((long)B * m->cout * N);
  UWG_HIP(hipMemcpy(m->xin, x, sizeof(float) * B * m->cin * N, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(dt, t, sizeof(float) * B * m->cout * N, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(dm, mask, sizeof(float) * B * m->cout * N, hipMemcpyHostToDevice));
  nn_forward(m, B, 0);
  hipLaunchKernelGGL(k_loss, B, 256, 0, 0, B, m->cout, N, m->y, dt, dm, m->gy, m->lossb);
  nn_backward(m, B, 0);
  std::vector<float> lb(B);
  UWG_HIP(hipMemcpy(lb.data(), m->lossb, sizeof(float) * B, hipMemcpyDeviceToHost));
  double acc = 0.0;
  for (float v : lb) acc += v;
  *loss = (float)(acc / B);
  if (grads) UWG_HIP(hipMemcpy(grads, m->grads, sizeof(float) * m->nparams, hipMemcpyDeviceToHost));
  (void)hipFree(dt);
  (void)hipFree(dm);
  return 0;
  NN_API_END
}

// upload a training set: x[S][4][N], t[S][3][N], mask[S][3][N]
extern "C" int uwg_nn_set_data(uwg_nn* h, long S, const float* x, const float* t, const float* mask) {
  NN_API_BEGIN
  NNModel* m = h->m;
  const long N = m->grid[0].N();
  for (void* p : {(void*)h->data_x, (void*)h->data_t, (void*)h->data_m}) (void)hipFree(p);
  h->data_x = nalloc<float>(S * m->cin * N);
  h->data_t = nalloc<float>(S * m->cout * N);
  h->data_m = nalloc<float>(S * m->cout * N);
  UWG_HIP(hipMemcpy(h->data_x, x, sizeof(float) * S * m->cin * N, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(h->data_t, t, sizeof(float) * S * m->cout * N, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(h->data_m, mask, sizeof(float) * S * m->cout * N, hipMemcpyHostToDevice));
  h->nsamples = S;
  return 0;
  NN_API_END
}

// Adam on minibatches drawn (deterministically from seed) from the uploaded set.
// losses[steps] receives each step's batch loss.
extern "C" int uwg_nn_train(uwg_nn* h, int steps, int B, float lr, float weight_decay, long seed, float* losses) {
  if (h->nsamples < 1) {
    uwg::set_error("no training data");
    return UWG_ERR_ARG;
  }
  NN_API_BEGIN
  NNModel* m = h->m;
  ensure_batch(m, B);
  const long N = m->grid[0].N();
  float* bt = nalloc<float>((long)B * m->cout * N);
  float* bm = nalloc<float>((long)B * m->cout * N);
  unsigned long long st = (unsigned long long)seed * 2862933555777941757ULL + 3037000493ULL;
  std::vector<float> lb(B);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < steps; it++) {
    for (int b = 0; b < B; b++) {
      st = st * 2862933555777941757ULL + 3037000493ULL;
      long k = (long)((st >> 17) % (unsigned long long)h->nsamples);
      UWG_HIP(hipMemcpyAsync(m->xin + (long)b * m->cin * N, h->data_x + k * m->cin * N, sizeof(float) * m->cin * N,
                             hipMemcpyDeviceToDevice, 0));
      UWG_HIP(hipMemcpyAsync(bt + (long)b * m->cout * N, h->data_t + k * m->cout * N, sizeof(float) * m->cout * N,
                             hipMemcpyDeviceToDevice, 0));
      UWG_HIP(hipMemcpyAsync(bm + (long)b * m->cout * N, h->data_m + k * m->cout * N, sizeof(float) * m->cout * N,
                             hipMemcpyDeviceToDevice, 0));
    }
    nn_forward(m, B, 0);
    hipLaunchKernelGGL(k_loss, B, 256, 0, 0, B, m->cout, N, m->y, bt, bm, m->gy, m->lossb);
    nn_backward(m, B, 0);
    m->step++;
    const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, nbk(m->nparams, 256), 256, 0, 0, m->nparams, m->params, m->grads, m->m1, m->m2, lr, b1,
                       b2, eps, c1, c2, weight_decay);
    if (losses) {
      UWG_HIP(hipMemcpy(lb.data(), m->lossb, sizeof(float) * B, hipMemcpyDeviceToHost));
      double acc = 0.0;
      for (float v : lb) acc += v;
      losses[it] = (float)(acc / B);
    }
  }
  UWG_HIP(hipDeviceSynchronize());
  (void)hipFree(bt);
  (void)hipFree(bm);
  return 0;
  NN_API_END
}

// mean loss over the uploaded set (evaluation, batches of B)
extern "C" int uwg_nn_eval(uwg_nn* h, int B, float* loss) {
  NN_API_BEGIN
  NNModel* m = h->m;
  ensure_batch(m, B);
  const long N = m->grid[0].N();
  double acc = 0.0;
  std::vector<float> lb(B);
  for (long s0 = 0; s0 < h->nsamples; s0 += B) {
    int nb = (int)std::min<long>(B, h->nsamples - s0);
    UWG_HIP(hipMemcpy(m->xin, h->data_x + s0 * m->cin * N, sizeof(float) * nb * m->cin * N, hipMemcpyDeviceToDevice));
    nn_forward(m, nb, 0);
    hipLaunchKernelGGL(k_loss, nb, 256, 0, 0, nb, m->cout, N, m->y, h->data_t + s0 * m->cout * N,
                       h->data_m + s0 * m->cout * N, nullptr, m->lossb);
    UWG_HIP(hipMemcpy(lb.data(), m->lossb, sizeof(float) * nb, hipMemcpyDeviceToHost));
    for (int b = 0; b < nb; b++) acc += lb[b];
  }
  *loss = (float)(acc / h->nsamples);
  return 0;
  NN_API_END
}

// ------------------------------------------------------------------------------------------
// C ABI for the learned Schur stencil: training, evaluation and test hooks
// ------------------------------------------------------------------------------------------
extern "C" int uwg_snn_create(int channels, int layers, long seed, uwg_snn** out) {
  if (channels < 8 || channels > 128 || channels % 8 != 0 || layers < 1 || layers > 6) {
    uwg::set_error("bad Schur-stencil network shape (channels a multiple of 8 in [8, 128], 1-6 layers)");
    return UWG_ERR_ARG;
  }
  NN_API_BEGIN
  uwg_snn* h = new uwg_snn();
  h->m = snn_alloc(channels, layers);
  SchurNet* m = h->m;
  // He-normal 3x3x3 layers from a 64-bit LCG (deterministic for a seed); zero head, so the
  // untrained network is the identity stencil (the plain approxS preconditioner)
  std::vector<float> p(m->nparams, 0.0f);
  unsigned long long st = (unsigned long long)seed * 6364136223846793005ULL + 1442695040888963407ULL;
  auto unif = [&]() {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    return ((st >> 11) + 0.5) / 9007199254740992.0;
  };
  for (int l = 0; l < m->L; l++) {
    const int ci = l == 0 ? kSnnFeat : m->C;
    const double sd = sqrt(2.0 / (ci * 27.0));
    for (long k = 0; k < (long)ci * m->C * 27; k++) {
      const double u1 = unif(), u2 = unif();
      p[m->woff[l] + k] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
  }
  UWG_HIP(hipMemcpy(m->params, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice));
  *out = h;
  return 0;
  NN_API_END
}

extern "C" int uwg_snn_destroy(uwg_snn* h) {
  if (!h) return 0;
  schur_net_destroy(h->m);
  delete h;
  return 0;
}

extern "C" int uwg_snn_param_count(uwg_snn* h, long* n) {
  *n = h->m->nparams;
  return 0;
}

extern "C" int uwg_snn_get_params(uwg_snn* h, float* out) {
  NN_API_BEGIN
  UWG_HIP(hipMemcpy(out, h->m->params, sizeof(float) * h->m->nparams, hipMemcpyDeviceToHost));
  return 0;
  NN_API_END
}

extern "C" int uwg_snn_set_params(uwg_snn* h, const float* in) {
  NN_API_BEGIN
  UWG_HIP(hipMemcpy(h->m->params, in, sizeof(float) * h->m->nparams, hipMemcpyHostToDevice));
  return 0;
  NN_API_END
}

// feat[9][N] of an element grid from approxS[N] (host arrays)
extern "C" int uwg_snn_features(int nx, int ny, int nz, const double* approxS, float* feat) {
  if (nx < 1 || ny < 1 || nz < 1) {
    uwg::set_error("bad element grid");
    return UWG_ERR_ARG;
  }
  NN_API_BEGIN
  const Grid3 eg{nx, ny, nz};
  const long N = eg.N();
  double* dS = nalloc<double>(N);
  float* df = nalloc<float>(kSnnFeat * N);
  UWG_HIP(hipMemcpy(dS, approxS, sizeof(double) * N, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_snn_features, nbk(N, 256), 256, 0, 0, eg, dS, df);
  UWG_HIP(hipGetLastError());
  UWG_HIP(hipMemcpy(feat, df, sizeof(float) * kSnnFeat * N, hipMemcpyDeviceToHost));
  (void)hipFree(dS);
  (void)hipFree(df);
  return 0;
  NN_API_END
}

// y[B][27][N] = network(x[B][9][N]) on an nx x ny x nz element grid (host arrays)
extern "C" int uwg_snn_forward(uwg_snn* h, int nx, int ny, int nz, int B, const float* x, float* y) {
  if (nx < 1 || ny < 1 || nz < 1 || B < 1) {
    uwg::set_error("bad grid or batch");
    return UWG_ERR_ARG;
  }
  NN_API_BEGIN
  SchurNet* m = h->m;
  const Grid3 g{nx, ny, nz};
  const long N = g.N();
  snn_ensure(m, B, N, false);
  UWG_HIP(hipMemcpy(m->xin, x, sizeof(float) * B * kSnnFeat * N, hipMemcpyHostToDevice));
  snn_forward(m, B, g, 0);
  UWG_HIP(hipMemcpy(y, m->y, sizeof(float) * B * 27 * N, hipMemcpyDeviceToHost));
  return 0;
  NN_API_END
}

// mean over the batch of |C(x) r - t|^2 / |t|^2 and its parameter gradient (host arrays)
extern "C" int uwg_snn_loss_grad(uwg_snn* h, int nx, int ny, int nz, int B, const float* x, const float* r,
                                 const float* t, float* loss, float* grads) {
  if (nx < 1 || ny < 1 || nz < 1 || B < 1) {
    uwg::set_error("bad grid or batch");
    return UWG_ERR_ARG;
  }
  NN_API_BEGIN
  SchurNet* m = h->m;
  const Grid3 g{nx, ny, nz};
  const long N = g.N();
  snn_ensure(m, B, N, true);
  float* dr = nalloc<float>((long)B * N);
  float* dt = nalloc<float>((long)B * N);
  UWG_HIP(hipMemcpy(m->xin, x, sizeof(float) * B * kSnnFeat * N, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(dr, r, sizeof(float) * B * N, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(dt, t, sizeof(float) * B * N, hipMemcpyHostToDevice));
  snn_forward(m, B, g, 0);
  snn_loss_backward(m, B, g, dr, dt, grads != nullptr, 0);
  std::vector<float> lb(B);
  UWG_HIP(hipMemcpy(lb.data(), m->lossb, sizeof(float) * B, hipMemcpyDeviceToHost));
  double acc = 0.0;
  for (float v : lb) acc += v;
  *loss = (float)(acc / B);
  if (grads) UWG_HIP(hipMemcpy(grads, m->grads, sizeof(float) * m->nparams, hipMemcpyDeviceToHost));
  (void)hipFree(dr);
  (void)hipFree(dt);
  return 0;
  NN_API_END
}

extern "C" int uwg_snn_sample_count(uwg_snn* h, long* n) {
  *n = h->m->nsamples;
  return 0;
}

extern "C" int uwg_snn_clear_data(uwg_snn* h) {
  SchurNet* m = h->m;
  for (void* p : {(void*)m->dx, (void*)m->dr, (void*)m->dt}) (void)hipFree(p);
  m->dx = m->dr = m->dt = nullptr;
  m->nsamples = m->scap = 0;
  return 0;
}

// Adam on minibatches of B samples drawn (deterministically from seed) from samples [0, ntrain)
// of the stored set; losses[steps] receives each step's batch loss (may be null)
extern "C" int uwg_snn_train(uwg_snn* h, int steps, int B, float lr, float weight_decay, long seed, long ntrain,
                             float* losses) {
  SchurNet* m = h->m;
  if (ntrain < 1 || ntrain > m->nsamples || B < 1 || steps < 0) {
    uwg::set_error("bad training arguments (ntrain must be in [1, stored samples])");
    return UWG_ERR_ARG;
  }
  NN_API_BEGIN
  const Grid3 g = m->dg;
  const long N = g.N();
  snn_ensure(m, B, N, true);
  float* br = nalloc<float>((long)B * N);
  float* bt = nalloc<float>((long)B * N);
  unsigned long long st = (unsigned long long)seed * 2862933555777941757ULL + 3037000493ULL;
  std::vector<float> lb(B);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < steps; it++) {
    for (int b = 0; b < B; b++) {
      st = st * 2862933555777941757ULL + 3037000493ULL;
      const long k = (long)((st >> 17) % (unsigned long long)ntrain);
      UWG_HIP(hipMemcpyAsync(m->xin + (long)b * kSnnFeat * N, m->dx + k * kSnnFeat * N, sizeof(float) * kSnnFeat * N,
                             hipMemcpyDeviceToDevice, 0));
      UWG_HIP(hipMemcpyAsync(br + (long)b * N, m->dr + k * N, sizeof(float) * N, hipMemcpyDeviceToDevice, 0));
      UWG_HIP(hipMemcpyAsync(bt + (long)b * N, m->dt + k * N, sizeof(float) * N, hipMemcpyDeviceToDevice, 0));
    }
    snn_forward(m, B, g, 0);
    snn_loss_backward(m, B, g, br, bt, true, 0);
    m->step++;
    const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, nbk(m->nparams, 256), 256, 0, 0, m->nparams, m->params, m->grads, m->m1, m->m2, lr, b1,
                       b2, eps, c1, c2, weight_decay);
    if (losses) {
      UWG_HIP(hipMemcpy(lb.data(), m->lossb, sizeof(float) * B, hipMemcpyDeviceToHost));
      double acc = 0.0;
      for (float v : lb) acc += v;
      losses[it] = (float)(acc / B);
    }
  }
  UWG_HIP(hipDeviceSynchronize());
  (void)hipFree(br);
  (void)hipFree(bt);
  return 0;
  NN_API_END
}

// mean loss over samples [first, first + count) of the stored set
extern "C" int uwg_snn_eval(uwg_snn* h, long first, long count, float* loss) {
  SchurNet* m = h->m;
  if (first < 0 || count < 1 || first + count > m->nsamples) {
    uwg::set_error("bad sample range");
    return UWG_ERR_ARG;
  }
  NN_API_BEGIN
  const Grid3 g = m->dg;
  const long N = g.N();
  snn_ensure(m, 1, N, true);
  double acc = 0.0;
  float lb = 0.0f;
  for (long k = first; k < first + count; k++) {
    UWG_HIP(hipMemcpy(m->xin, m->dx + k * kSnnFeat * N, sizeof(float) * kSnnFeat * N, hipMemcpyDeviceToDevice));
    snn_forward(m, 1, g, 0);
    snn_loss_backward(m, 1, g, m->dr + k * N, m->dt + k * N, false, 0);
    UWG_HIP(hipMemcpy(&lb, m->lossb, sizeof(float), hipMemcpyDeviceToHost));
    acc += lb;
  }
  *loss = (float)(acc / count);
  return 0;
  NN_API_END
}
