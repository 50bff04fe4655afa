// Bit-exact swarm kernels: random layout, cell search, coincident map (Newton),
// discrete-Voronoi weights, RK2 advection and interpolation.
// Compiled with -ffp-contract=off: every floating-point expression keeps the reference's
// operation order (citations in uwg_geom.h and below; LU/ = src/underworld/libUnderworld/).
#include "uwg_internal.h"
#include "uwg_geom.h"

namespace uwg {

struct GeomArgs {
  int res[3];
  double lo[3], hi[3];
  long E;
};
static GeomArgs pack(const Geom& g) {
  GeomArgs a;
  for (int d = 0; d < 3; d++) {
    a.res[d] = g.res[d];
    a.lo[d] = g.lo[d];
    a.hi[d] = g.hi[d];
  }
  a.E = g.E;
  return a;
}

static inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------------------
// A1 random layout.  The reference draws from one Park-Miller stream (Random.c:36-53),
// cell by cell, particle by particle, dimension by dimension
// (MeshParticleLayout.c:194-229).  The generator is multiplicative (x -> 16807 x mod
// 2^31-1, which Schrage's method evaluates exactly), so draw j equals
// 16807^(j+1) * seed0 mod m and every particle can jump straight to its own draws.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long mulmod(unsigned long long a, unsigned long long b) {
  return (a * b) % 2147483647ULL;
}
__device__ unsigned long long powmod(unsigned long long base, unsigned long long e) {
  unsigned long long r = 1;
  base %= 2147483647ULL;
  while (e) {
    if (e & 1ULL) r = mulmod(r, base);
    base = mulmod(base, base);
    e >>= 1;
  }
  return r;
}

template <int DIM>
__global__ void k_layout_random(GeomArgs g, int ppc, unsigned long long s0, double* coords, long P) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  long e = p / ppc;
  unsigned long long st = mulmod(powmod(16807ULL, (unsigned long long)p * DIM), s0);
  double xi[3], v[24], N[8];
  for (int d = 0; d < DIM; d++) {
    st = mulmod(st, 16807ULL);
    double r = (double)(long)st;
    xi[d] = (1.0 - -1.0) * r / (double)2147483647L + -1.0;
  }
  element_verts<DIM>(g.res, g.lo, g.hi, e, v);
  shape<DIM>(xi, N);
  double x[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < (DIM == 3 ? 8 : 4); a++)
    for (int d = 0; d < DIM; d++) x[d] += N[a] * v[a * DIM + d];
  for (int d = 0; d < DIM; d++) coords[p * DIM + d] = x[d];
}

void launch_layout_random(const Geom& g, int ppc, long seed, double* coords, hipStream_t s) {
  long P = g.E * ppc;
  unsigned long long s0 = (unsigned long long)(seed ^ 123459876L);
  if (g.dim == 3)
    hipLaunchKernelGGL(k_layout_random<3>, nblocks(P, 256), 256, 0, s, pack(g), ppc, s0, coords, P);
  else
    hipLaunchKernelGGL(k_layout_random<2>, nblocks(P, 256), 256, 0, s, pack(g), ppc, s0, coords, P);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// PerCellSpaceFillerLayout (layouts.py:146-181; MeshParticleLayout.c:194-229, filltype 0).
// Sequentially the reference draws, per particle, one Park-Miller number r (seeded as above,
// default seed 13) and then per axis: if r > 0.9 one Sobol number is skipped, and the next is
// the local coordinate.  Each axis' Sobol generator (SobolGenerator.c, WORDLENGTH 30) after n
// draws holds the XOR of its direction numbers v_b << (29 - b) over the set bits b of the
// Gray code n ^ (n >> 1), so particle p's value is a closed form of its draw count
// n_p = p + 1 + (skips of the particles before it) + (its own skip): one flag kernel, an
// exclusive scan of the flags, then every particle is independent.
// ---------------------------------------------------------------------------------------
__global__ void k_sobol_skip_flags(long P, unsigned long long s0, int* __restrict__ flags) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const unsigned long long st = mulmod(powmod(16807ULL, (unsigned long long)p + 1ULL), s0);
  const double randnum = (1.0 - 0.0) * (double)(long)st / (double)2147483647L + 0.0;
  flags[p] = randnum > 0.9 ? 1 : 0;
}

template <int DIM>
__global__ void k_layout_sobol(GeomArgs g, int ppc, SobolDirs dirs, const int* __restrict__ flags,
                               const long long* __restrict__ before, double* coords, long P) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  long e = p / ppc;
  const unsigned long long n = (unsigned long long)p + 1ULL + (unsigned long long)before[p] + (unsigned long long)flags[p];
  const unsigned long long gray = n ^ (n >> 1);
  double xi[3], v[24], N[8];
  for (int d = 0; d < DIM; d++) {
    unsigned int val = 0u;
    for (int b = 0; b < 30; b++)
      if ((gray >> b) & 1ULL) val ^= dirs.v[d][b] << (30 - b - 1);
    const double x = (double)val / (double)(1L << 30);
    xi[d] = (1.0 - -1.0) * x + -1.0;
  }
  element_verts<DIM>(g.res, g.lo, g.hi, e, v);
  shape<DIM>(xi, N);
  double x[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < (DIM == 3 ? 8 : 4); a++)
    for (int d = 0; d < DIM; d++) x[d] += N[a] * v[a * DIM + d];
  for (int d = 0; d < DIM; d++) coords[p * DIM + d] = x[d];
}

// Direction numbers of the generators of table entries 0, 1, 2 (polynomial degree 1/2/3,
// coefficient 0/1/1, initial numbers {1}, {1,1}, {1,3,7}; SobolGenerator.c table, built by
// MeshParticleLayout.c:58-60), by the generator's recursion (SobolGenerator_GetDirectionalNumber).
static SobolDirs sobol_dirs() {
  static const unsigned deg[3] = {1, 2, 3}, coef[3] = {0, 1, 1};
  static const unsigned init[3][3] = {{1, 0, 0}, {1, 1, 0}, {1, 3, 7}};
  SobolDirs D;
  for (int a = 0; a < 3; a++) {
    for (unsigned i = 0; i < deg[a]; i++) D.v[a][i] = init[a][i];
    for (unsigned i = deg[a]; i < 30; i++) {
      unsigned result = 0, dn = 0;
      for (unsigned c = 0; c < deg[a]; c++) {
        if (c == deg[a] - 1 || (coef[a] & (1u << (deg[a] - 2 - c)))) {
          dn = D.v[a][i - 1 - c];
          result ^= (1u << (c + 1)) * dn;
        }
      }
      D.v[a][i] = result ^ dn;
    }
  }
  return D;
}

void launch_layout_sobol(const Geom& g, int ppc, long seed, double* coords, hipStream_t s) {
  const long P = g.E * ppc;
  // the generator's word has 30 bits: a draw count must stay below 2^30 (at most 2 draws per particle)
  if (2 * P + 2 >= (1L << 30)) throw std::runtime_error("PerCellSpaceFillerLayout: too many particles for the 30-bit Sobol generator");
  const unsigned long long s0 = (unsigned long long)(seed ^ 123459876L);
  int* flags = nullptr;
  long long* before = nullptr;
  UWG_HIP(hipMalloc(&flags, sizeof(int) * (P + 1)));
  UWG_HIP(hipMalloc(&before, sizeof(long long) * (P + 1)));
  hipLaunchKernelGGL(k_sobol_skip_flags, nblocks(P, 256), 256, 0, s, P, s0, flags);
  UWG_HIP(hipGetLastError());
  scan_exclusive_flags(flags, before, P, s);
  const SobolDirs dirs = sobol_dirs();
  if (g.dim == 3)
    hipLaunchKernelGGL(k_layout_sobol<3>, nblocks(P, 256), 256, 0, s, pack(g), ppc, dirs, flags, before, coords, P);
  else
    hipLaunchKernelGGL(k_layout_sobol<2>, nblocks(P, 256), 256, 0, s, pack(g), ppc, dirs, flags, before, coords, P);
  UWG_HIP(hipGetLastError());
  UWG_HIP(hipStreamSynchronize(s));
  UWG_HIP(hipFree(flags));
  UWG_HIP(hipFree(before));
}

template <int DIM>
__global__ void k_mesh_coords(GeomArgs g, double* out, long N) {
  long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  long nx = g.res[0] + 1, ny = g.res[1] + 1;
  int idx[3] = {(int)(n % nx), (int)((n / nx) % ny), (int)(n / (nx * ny))};
  for (int d = 0; d < DIM; d++) out[n * DIM + d] = node_coord_1d(idx[d], g.res[d] + 1, g.lo[d], g.hi[d]);
}

void launch_mesh_coords(const Geom& g, double* out, hipStream_t s) {
  if (g.dim == 3)
    hipLaunchKernelGGL(k_mesh_coords<3>, nblocks(g.N, 256), 256, 0, s, pack(g), out, g.N);
  else
    hipLaunchKernelGGL(k_mesh_coords<2>, nblocks(g.N, 256), 256, 0, s, pack(g), out, g.N);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// A17 cell search for every particle (owner = E when outside the domain)
// ---------------------------------------------------------------------------------------
template <int DIM>
__global__ void k_cell_of(GeomArgs g, long P, const double* coords, int* owner) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double x[3];
  for (int d = 0; d < DIM; d++) x[d] = coords[p * DIM + d];
  owner[p] = (int)cell_of<DIM>(g.res, g.lo, g.hi, x, g.E);
}

void launch_cell_of(const Geom& g, long P, const double* coords, int* owner, hipStream_t s) {
  if (P == 0) return;
  if (g.dim == 3)
    hipLaunchKernelGGL(k_cell_of<3>, nblocks(P, 256), 256, 0, s, pack(g), P, coords, owner);
  else
    hipLaunchKernelGGL(k_cell_of<2>, nblocks(P, 256), 256, 0, s, pack(g), P, coords, owner);
  UWG_HIP(hipGetLastError());
}

// slot -> owning cell, from the CSR cell pointer
__global__ void k_slot_cell(long E, const int* cell_ptr, int* slot_cell) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  for (int s = cell_ptr[e]; s < cell_ptr[e + 1]; s++) slot_cell[s] = (int)e;
}
void launch_slot_cell(long E, const int* cell_ptr, int* slot_cell, hipStream_t s) {
  hipLaunchKernelGGL(k_slot_cell, nblocks(E, 256), 256, 0, s, E, cell_ptr, slot_cell);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// A3 coincident map: xi of every slot by Newton in its cell (CoincidentMapper.c:76-130)
// ---------------------------------------------------------------------------------------
template <int DIM>
__global__ void k_coincident(GeomArgs g, long P, const int* slot_cell, const int* cell_pid,
                             const double* coords, double* xi_slot) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P) return;
  double v[24], x[3], xi[3];
  element_verts<DIM>(g.res, g.lo, g.hi, slot_cell[s], v);
  long p = cell_pid[s];
  for (int d = 0; d < DIM; d++) x[d] = coords[p * DIM + d];
  newton_local<DIM>(v, x, xi);
  for (int d = 0; d < DIM; d++) xi_slot[s * DIM + d] = xi[d];
}

void launch_coincident_map(const Geom& g, long P, const int* slot_cell, const int* cell_pid,
                           const double* coords, double* xi_slot, hipStream_t s) {
  if (P == 0) return;
  if (g.dim == 3)
    hipLaunchKernelGGL(k_coincident<3>, nblocks(P, 256), 256, 0, s, pack(g), P, slot_cell, cell_pid, coords, xi_slot);
  else
    hipLaunchKernelGGL(k_coincident<2>, nblocks(P, 256), 256, 0, s, pack(g), P, slot_cell, cell_pid, coords, xi_slot);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// A4 discrete Voronoi cell weights, one thread per element, state in a per-thread
// scratch region.  Same growth as _DVCWeights_CreateVoronoi3D/2D (DVCWeights.c:676-765):
// seed, then rounds over particles 0..n-1 of ClaimCells (:256-311: claim unowned, steal
// only when DistanceTest > 0) followed by UpdateBchain (:388-438: neighbours S,N,E,W,U,D
// of newly claimed cells, first occurrence kept).  Centroids/weights as GetCentroids
// (:198-254).  A thread that would exceed its list capacity sets its flag and writes
// nothing; the host re-runs such elements with larger capacities.
// ---------------------------------------------------------------------------------------
struct DvcLayout {
  long owner_off, done_off, cnt_off, nb_off, claimed_off, bound_off, bytes;
};
__host__ __device__ inline DvcLayout dvc_layout(int ncell, int max_np, int cap) {
  DvcLayout L;
  long o = 0;
  auto al = [](long v) { return (v + 15) & ~15L; };
  L.owner_off = o; o = al(o + 2L * ncell);
  L.done_off = o; o = al(o + ncell);
  L.cnt_off = o; o = al(o + 4L * max_np);
  L.nb_off = o; o = al(o + 4L * max_np);
  L.claimed_off = o; o = al(o + 2L * cap);
  L.bound_off = o; o = al(o + 2L * cap * max_np);
  L.bytes = o;
  return L;
}

template <int DIM>
__global__ void k_dvc(int num, long E0, long nE, const int* elist, const int* cell_ptr, double* xi_slot,
                      double* w_slot, char* scratch, int max_np, int cap, int* flag) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nE) return;
  const long e = elist ? (long)elist[t] : E0 + t;
  const int nx = num, ny = num, nz = DIM == 3 ? num : 1;
  const int ncell = nx * ny * nz;
  const int nnb = DIM == 3 ? 6 : 4;
  const int s0 = cell_ptr[e];
  const int np = cell_ptr[e + 1] - s0;
  if (np <= 0) { flag[t] = 2; return; }     // under-resolved cell (reference firewall)
  if (np > max_np) { flag[t] = 1; return; }
  const DvcLayout L = dvc_layout(ncell, max_np, cap);
  char* base = scratch + (long)t * L.bytes;
  short* own = (short*)(base + L.owner_off);
  unsigned char* done = (unsigned char*)(base + L.done_off);
  int* cnt = (int*)(base + L.cnt_off);
  int* nbound = (int*)(base + L.nb_off);
  short* claimed = (short*)(base + L.claimed_off);
  short* bound = (short*)(base + L.bound_off);
  const double dx = (1.0 - -1.0) / nx, dy = (1.0 - -1.0) / ny, dz = (1.0 - -1.0) / nz;
  // cell centres: X = X + dx accumulation per axis (_DVCWeights_ConstructGrid :527-544)
  double cx[32], cy[32], cz[32];
  {
    double X = -1.0 - dx / 2.0;
    for (int i = 0; i < nx; i++) { X = X + dx; cx[i] = X; }
    double Y = -1.0 - dy / 2.0;
    for (int j = 0; j < ny; j++) { Y = Y + dy; cy[j] = Y; }
    double Z = -1.0 - dz / 2.0;
    for (int k = 0; k < nz; k++) { Z = Z + dz; cz[k] = Z; }
  }
  for (int c = 0; c < ncell; c++) { own[c] = -1; done[c] = 0; }
  const double* pos = xi_slot + (long)s0 * DIM;
  bool overflow = false;

  auto neighbour = [&](int c, int k) -> int {
    int i = c % nx, j = (c / nx) % ny, kk = c / (nx * ny);
    switch (k) {
      case 0: return j > 0 ? c - nx : -2;            // S
      case 1: return j < ny - 1 ? c + nx : -2;       // N
      case 2: return i < nx - 1 ? c + 1 : -2;        // E
      case 3: return i > 0 ? c - 1 : -2;             // W
      case 4: return kk < nz - 1 ? c + nx * ny : -2; // U
      default: return kk > 0 ? c - nx * ny : -2;     // D
    }
  };
  // UpdateBchain for particle pi from its claimed list (length ncl)
  auto update = [&](int pi, int ncl) {
    short* b = bound + (long)pi * cap;
    int nb = 0;
    for (int q = 0; q < ncl && !overflow; q++) {
      int c0 = claimed[q];
      for (int k = 0; k < nnb; k++) {
        int c1 = neighbour(c0, k);
        if (c1 != -2 && own[c1] != pi && done[c1] != 1) {
          if (nb >= cap) { overflow = true; break; }
          b[nb++] = (short)c1;
          done[c1] = 1;
        }
      }
    }
    for (int q = 0; q < nb; q++) done[b[q]] = 0;
    nbound[pi] = nb;
  };
  for (int i = 0; i < np && !overflow; i++) {
    int k = (int)((pos[i * DIM + 0] - -1.0) / dx);
    int j = (int)((pos[i * DIM + 1] - -1.0) / dy);
    int l = DIM == 3 ? (int)((pos[i * DIM + 2] - -1.0) / dz) : 0;
    if (k == nx) k--;
    if (j == ny) j--;
    if (DIM == 3 && l == nz) l--;
    // memory guard only: local coordinates lie in [-1, 1] up to roundoff
    k = k < 0 ? 0 : (k > nx - 1 ? nx - 1 : k);
    j = j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j);
    l = l < 0 ? 0 : (l > nz - 1 ? nz - 1 : l);
    int id = k + j * nx + l * nx * ny;
    own[id] = (short)i;
    claimed[0] = (short)id;
    update(i, 1);
  }
  int claimed_total = 1;
  while (claimed_total != 0 && !overflow) {
    claimed_total = 0;
    for (int pi = 0; pi < np && !overflow; pi++) {
      const short* b = bound + (long)pi * cap;
      int nb = nbound[pi], ncl = 0;
      for (int q = 0; q < nb; q++) {
        int c0 = b[q];
        int o = own[c0];
        if (o == -1) {
          claimed[ncl++] = (short)c0;
          own[c0] = (short)pi;
        } else if (o != pi) {
          int ci = c0 % nx, cj = (c0 / nx) % ny, ck = c0 / (nx * ny);
          double x0 = cx[ci], y0 = cy[cj], z0 = cz[ck];
          const double* p1 = pos + o * DIM;
          const double* p2 = pos + pi * DIM;
          double r = (p1[0] + p2[0] - x0 - x0) * (p1[0] - p2[0]) + (p1[1] + p2[1] - y0 - y0) * (p1[1] - p2[1]);
          if (DIM == 3) r = r + (p1[2] + p2[2] - z0 - z0) * (p1[2] - p2[2]);
          if (r > 0.0) {
            claimed[ncl++] = (short)c0;
            own[c0] = (short)pi;
          }
        }
      }
      claimed_total += ncl;
      update(pi, ncl);
    }
  }
  if (overflow) { flag[t] = 1; return; }
  // centroids and weights, grid-index order (_DVCWeights_GetCentroids3D :198-226)
  double da = DIM == 3 ? dx * dy * dz : dx * dy;
  double* out = xi_slot + (long)s0 * DIM;
  for (int i = 0; i < np; i++) {
    cnt[i] = 0;
    for (int d = 0; d < DIM; d++) out[i * DIM + d] = 0.0;
  }
  for (int c = 0; c < ncell; c++) {
    int p = own[c];
    if (p < 0) continue;
    int ci = c % nx, cj = (c / nx) % ny, ck = c / (nx * ny);
    out[p * DIM + 0] += cx[ci];
    out[p * DIM + 1] += cy[cj];
    if (DIM == 3) out[p * DIM + 2] += cz[ck];
    cnt[p]++;
  }
  for (int i = 0; i < np; i++) {
    w_slot[s0 + i] = cnt[i] * da;
    if (cnt[i] != 0)
      for (int d = 0; d < DIM; d++) out[i * DIM + d] /= cnt[i];
  }
  flag[t] = 0;
}

// One wavefront per element, grid-stride over elements.  The claim grid, a one-bit-per-cell
// mark and the particles' local coordinates live in LDS (about 8 KB, so many elements are in
// flight per CU); each particle's boundary chain lives in a per-workgroup global buffer
// (double-buffered) that the 64 lanes read and write contiguously.
//
// Why the lanes may work on a chain at once: within one particle's step every chain cell is
// examined once (the chain holds no duplicates) and its decision depends only on ownership
// before the step, so the set of cells the particle claims does not depend on the order of its
// chain; the rebuilt chain is the set of neighbours of those cells not owned by the particle,
// which does not depend on order either.  Ownership after every step - and therefore the
// weights and centroids, summed in grid order at the end - is exactly that of the sequential
// algorithm above, bit for bit; only the order of cells inside a chain may differ.  Duplicate
// neighbours are dropped by an atomic test-and-set on the mark bits.  Elements with more than
// DVCW_MAXNP particles or a chain longer than `cap` are flagged and left to k_dvc.
#define DVCW_MAXNC 4096
#define DVCW_MAXNP 128
template <int DIM, int NUM>
__global__ __launch_bounds__(64) void k_dvc_wave(int num_rt, long E, const int* __restrict__ cell_ptr,
                                                 double* __restrict__ xi_slot, double* __restrict__ w_slot,
                                                 short* __restrict__ gbound, int cap, int* __restrict__ flag) {
  const int num = NUM > 0 ? NUM : num_rt;  // compile-time grid size for the default resolution
  constexpr int NC = NUM > 0 ? (DIM == 3 ? NUM * NUM * NUM : NUM * NUM) : DVCW_MAXNC;  // claim grid cells
  constexpr int NCW = (NC + 31) / 32;                                                 // mark words
  __shared__ unsigned char own[NC];
  __shared__ unsigned int mark[NCW];
  __shared__ double ppos[DVCW_MAXNP * DIM];
  __shared__ int nbound[DVCW_MAXNP];
  __shared__ unsigned char which[DVCW_MAXNP];
  __shared__ double cen[3][32];
  const int lane = threadIdx.x;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int nx = num, ny = num, nz = DIM == 3 ? num : 1;
  const int nxy = nx * ny, ncell = nxy * nz;
  const int nnb = DIM == 3 ? 6 : 4;
  const double dx = (1.0 - -1.0) / nx, dy = (1.0 - -1.0) / ny, dz = (1.0 - -1.0) / nz;
  if (lane == 0) {
    // cell centres: X = X + dx accumulation per axis (_DVCWeights_ConstructGrid :527-544)
    double X = -1.0 - dx / 2.0;
    for (int i = 0; i < nx; i++) { X = X + dx; cen[0][i] = X; }
    double Y = -1.0 - dy / 2.0;
    for (int j = 0; j < ny; j++) { Y = Y + dy; cen[1][j] = Y; }
    double Z = -1.0 - dz / 2.0;
    for (int k = 0; k < nz; k++) { Z = Z + dz; cen[2][k] = Z; }
  }
  for (int q = lane; q < NCW; q += 64) mark[q] = 0u;
  // chains: buffer h (0/1) of particle p at gb + (2 p + h) * cap
  short* gb = gbound + (long)blockIdx.x * 2 * DVCW_MAXNP * cap;
  // neighbour k of cell c in the order S, N, E, W, U, D; -2 outside the grid.  Selects, not a
  // switch: k differs between lanes, and a branch per case costs scalar exec-mask work per lane
  // group
  auto neighbour = [&](int c, int k) -> int {
    const int i = c % nx, j = (c / nx) % ny, kk = c / nxy;
    const int off = k == 0 ? -nx : (k == 1 ? nx : (k == 2 ? 1 : (k == 3 ? -1 : (k == 4 ? nxy : -nxy))));
    const bool ok = k == 0 ? j > 0
                           : (k == 1 ? j < ny - 1
                                     : (k == 2 ? i < nx - 1 : (k == 3 ? i > 0 : (k == 4 ? kk < nz - 1 : kk > 0))));
    return ok ? c + off : -2;
  };
  for (long e = blockIdx.x; e < E; e += gridDim.x) {
    const int s0 = cell_ptr[e];
    const int np = cell_ptr[e + 1] - s0;
    if (np <= 0 || np > DVCW_MAXNP) {
      if (lane == 0) flag[e] = np <= 0 ? 2 : 1;
      continue;
    }
    for (int c = lane; c < ncell; c += 64) own[c] = 0xFF;
    for (int i = lane; i < np * DIM; i += 64) ppos[i] = xi_slot[(long)s0 * DIM + i];
    for (int i = lane; i < np; i += 64) which[i] = 0;
    __syncthreads();
    bool overflow = false;
    // rebuild the chain of particle pi from the ncl claimed cells in `cl`
    auto update = [&](int pi, const short* cl, int ncl) {
      short* nxt = gb + (long)(2 * pi + (which[pi] ^ 1)) * cap;
      const int ncand = ncl * nnb;
      int nb = 0;
      for (int base = 0; base < ncand; base += 64) {
        const int t = base + lane;
        bool keep = false;
        int c1 = -2;
        if (t < ncand) {
          const int q = t / nnb, k = t - q * nnb;
          c1 = neighbour(cl[q], k);
          if (c1 != -2 && own[c1] != pi) {
            const unsigned int bit = 1u << (c1 & 31);
            keep = (atomicOr(&mark[c1 >> 5], bit) & bit) == 0u;
          }
        }
        const unsigned long long m = __ballot(keep);
        const int rank = __popcll(m & lt);
        const int cntk = __popcll(m);
        if (nb + cntk > cap) {
          overflow = true;
          break;
        }
        if (keep) nxt[nb + rank] = (short)c1;
        nb += cntk;
      }
      __syncthreads();
      // every set mark bit belongs to a cell of the new chain: clear those words
      for (int q = lane; q < nb; q += 64) mark[nxt[q] >> 5] = 0u;
      if (lane == 0) {
        nbound[pi] = nb;
        which[pi] ^= 1;
      }
      __syncthreads();
    };
    // seeds, in particle order (_DVCWeights_CreateVoronoi3D :676-721)
    for (int i = 0; i < np && !overflow; i++) {
      short* seed = gb + (long)(2 * i + 0) * cap;  // the seed cell, read by update()
      if (lane == 0) {
        int k = (int)((ppos[i * DIM + 0] - -1.0) / dx);
        int j = (int)((ppos[i * DIM + 1] - -1.0) / dy);
        int l = DIM == 3 ? (int)((ppos[i * DIM + 2] - -1.0) / dz) : 0;
        if (k == nx) k--;
        if (j == ny) j--;
        if (DIM == 3 && l == nz) l--;
        k = k < 0 ? 0 : (k > nx - 1 ? nx - 1 : k);
        j = j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j);
        l = l < 0 ? 0 : (l > nz - 1 ? nz - 1 : l);
        const int id = k + j * nx + l * nxy;
        own[id] = (unsigned char)i;
        seed[0] = (short)id;
      }
      __syncthreads();
      update(i, seed, 1);
    }
    // rounds of claims over the particles until nothing changes (_DVCWeights_ClaimCells :256-311)
    int claimed_total = 1;
    while (claimed_total != 0 && !overflow) {
      claimed_total = 0;
      for (int pi = 0; pi < np && !overflow; pi++) {
        const int nb = nbound[pi];
        if (nb == 0) continue;  // an empty chain claims nothing and stays empty
        short* b = gb + (long)(2 * pi + which[pi]) * cap;
        int ncl = 0;
        for (int base = 0; base < nb; base += 64) {
          const int q = base + lane;
          // every lane evaluates the claim rule (lanes past the chain on its last entry, unowned
          // cells against the particle itself) and the predicate selects the result
          const int c0 = b[q < nb ? q : nb - 1];
          const int o = own[c0];
          const int oo = o == 0xFF ? pi : o;
          const int ci = c0 % nx, cj = (c0 / nx) % ny, ck = c0 / nxy;
          const double x0 = cen[0][ci], y0 = cen[1][cj], z0 = cen[2][ck];
          const double* p1 = ppos + oo * DIM;
          const double* p2 = ppos + pi * DIM;
          double r = (p1[0] + p2[0] - x0 - x0) * (p1[0] - p2[0]) + (p1[1] + p2[1] - y0 - y0) * (p1[1] - p2[1]);
          if (DIM == 3) r = r + (p1[2] + p2[2] - z0 - z0) * (p1[2] - p2[2]);
          const bool take = q < nb && (o == 0xFF || (o != pi && r > 0.0));
          const unsigned long long m = __ballot(take);
          const int rank = __popcll(m & lt);
          if (take) {
            own[c0] = (unsigned char)pi;
            b[ncl + rank] = (short)c0;  // claimed cells compacted in place (ncl + rank <= q)
          }
          ncl += __popcll(m);
        }
        claimed_total += ncl;
        if (ncl == 0) {
          // nothing claimed: the rebuilt chain is empty
          if (lane == 0) nbound[pi] = 0;
          __syncthreads();
          continue;
        }
        __syncthreads();
        update(pi, b, ncl);
      }
    }
    if (overflow) {
      // leave LDS marks clean for the next element
      for (int q = lane; q < NCW; q += 64) mark[q] = 0u;
      if (lane == 0) flag[e] = 1;
      __syncthreads();
      continue;
    }
    // centroids and weights, grid-index order (_DVCWeights_GetCentroids3D :198-226).  A
    // particle's cells lie inside their bounding box, so its lane scans only the box, in grid
    // order: the same cells in the same order as a scan of the whole grid, hence the same sums.
    // The boxes come from one pass over the grid rows (a lane per row, one atomic min/max per
    // run of equal owners); they live where the coordinates were (ppos is not read again for
    // this element).
    int* bb = reinterpret_cast<int*>(ppos);  // [2 DIM][DVCW_MAXNP]: lower i, j, k, then upper
    for (int p = lane; p < np; p += 64) {
      for (int d = 0; d < DIM; d++) {
        bb[d * DVCW_MAXNP + p] = 1 << 20;
        bb[(DIM + d) * DVCW_MAXNP + p] = -1;
      }
    }
    __syncthreads();
    for (int row = lane; row < ny * nz; row += 64) {
      const int j = row % ny, k = row / ny;
      int cur = 0xFF, i0 = 0;
      for (int i = 0; i <= nx; i++) {
        const int o = i < nx ? (int)own[i + nx * row] : -1;
        if (o != cur) {
          if (cur != 0xFF) {
            atomicMin(&bb[0 * DVCW_MAXNP + cur], i0);
            atomicMax(&bb[DIM * DVCW_MAXNP + cur], i - 1);
            atomicMin(&bb[1 * DVCW_MAXNP + cur], j);
            atomicMax(&bb[(DIM + 1) * DVCW_MAXNP + cur], j);
            if (DIM == 3) {
              atomicMin(&bb[2 * DVCW_MAXNP + cur], k);
              atomicMax(&bb[(DIM + 2) * DVCW_MAXNP + cur], k);
            }
          }
          cur = o;
          i0 = i;
        }
      }
    }
    __syncthreads();
    const double da = DIM == 3 ? dx * dy * dz : dx * dy;
    for (int p = lane; p < np; p += 64) {
      double sx = 0.0, sy = 0.0, sz = 0.0;
      int cnt = 0;
      const int i0 = bb[0 * DVCW_MAXNP + p], i1 = bb[DIM * DVCW_MAXNP + p];
      const int j0 = bb[1 * DVCW_MAXNP + p], j1 = bb[(DIM + 1) * DVCW_MAXNP + p];
      const int k0 = DIM == 3 ? bb[2 * DVCW_MAXNP + p] : 0, k1 = DIM == 3 ? bb[(DIM + 2) * DVCW_MAXNP + p] : 0;
      for (int ck = k0; ck <= k1; ck++)
        for (int cj = j0; cj <= j1; cj++)
          for (int ci = i0; ci <= i1; ci++) {
            if (own[ci + nx * cj + nxy * ck] != p) continue;
            sx += cen[0][ci];
            sy += cen[1][cj];
            if (DIM == 3) sz += cen[2][ck];
            cnt++;
          }
      double* out = xi_slot + (long)(s0 + p) * DIM;
      w_slot[s0 + p] = cnt * da;
      if (cnt != 0) {
        out[0] = sx / cnt;
        out[1] = sy / cnt;
        if (DIM == 3) out[2] = sz / cnt;
      } else {
        out[0] = 0.0;
        out[1] = 0.0;
        if (DIM == 3) out[2] = 0.0;
      }
    }
    if (lane == 0) flag[e] = 0;
    __syncthreads();
  }
}

bool dvc_wave_supported(int dim, int num) {
  const long nc = (long)num * num * (dim == 3 ? num : 1);
  return nc <= DVCW_MAXNC && num <= 32;
}

void launch_dvc_wave(int dim, int num, long E, const int* cell_ptr, double* xi_slot, double* w_slot, short* gbound,
                     int nblk, int cap, int* flag, hipStream_t s) {
  if (E == 0) return;
  if (dim == 3 && num == 15)
    hipLaunchKernelGGL((k_dvc_wave<3, 15>), nblk, 64, 0, s, num, E, cell_ptr, xi_slot, w_slot, gbound, cap, flag);
  else if (dim == 3)
    hipLaunchKernelGGL((k_dvc_wave<3, 0>), nblk, 64, 0, s, num, E, cell_ptr, xi_slot, w_slot, gbound, cap, flag);
  else if (num == 15)
    hipLaunchKernelGGL((k_dvc_wave<2, 15>), nblk, 64, 0, s, num, E, cell_ptr, xi_slot, w_slot, gbound, cap, flag);
  else
    hipLaunchKernelGGL((k_dvc_wave<2, 0>), nblk, 64, 0, s, num, E, cell_ptr, xi_slot, w_slot, gbound, cap, flag);
  UWG_HIP(hipGetLastError());
}

long dvc_wave_bound_shorts(int cap) { return 2L * DVCW_MAXNP * cap; }

long dvc_scratch_bytes(int dim, int num, int max_np, int cap) {
  int ncell = num * num * (dim == 3 ? num : 1);
  return dvc_layout(ncell, max_np, cap).bytes;
}

void launch_dvc_list(int dim, int num, long E0, long nE, const int* elist, const int* cell_ptr,
                     double* xi_slot, double* w_slot, void* scratch, int max_np, int cap, int* d_flag,
                     hipStream_t s) {
  if (nE == 0) return;
  if (dim == 3)
    hipLaunchKernelGGL(k_dvc<3>, nblocks(nE, 64), 64, 0, s, num, E0, nE, elist, cell_ptr, xi_slot, w_slot,
                       (char*)scratch, max_np, cap, d_flag);
  else
    hipLaunchKernelGGL(k_dvc<2>, nblocks(nE, 64), 64, 0, s, num, E0, nE, elist, cell_ptr, xi_slot, w_slot,
                       (char*)scratch, max_np, cap, d_flag);
  UWG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// A15/A16 RK2 advection (TimeIntegrand_SecondOrder, TimeIntegrand.c:249-319, fallback
// :441-476 since SwarmAdvector allows it).  Velocity lookup: regular search, Newton,
// sum_a v_a N_a in element-node order (FeVariable.c:389-403, 1007-1037).
// ---------------------------------------------------------------------------------------
template <int DIM>
__device__ bool velocity_at(const GeomArgs& g, const double* vel, const double* x, double* v) {
  int ijk[3];
  if (!search_element_ijk<DIM>(g.res, g.lo, g.hi, x, ijk)) return false;
  double verts[24], xi[3], N[8];
  element_verts_ijk<DIM>(g.res, g.lo, g.hi, ijk, verts);
  newton_local<DIM>(verts, x, xi);
  shape<DIM>(xi, N);
  long b = element_base_node_ijk<DIM>(g.res, ijk);
  long nx = g.res[0] + 1, nxy = nx * (g.res[1] + 1);
  for (int d = 0; d < DIM; d++) v[d] = 0.0;
  for (int a = 0; a < (DIM == 3 ? 8 : 4); a++) {
    long n = b + (a & 1) + ((a >> 1) & 1) * nx + (DIM == 3 ? ((a >> 2) & 1) * nxy : 0);
    for (int d = 0; d < DIM; d++) v[d] += vel[n * DIM + d] * N[a];
  }
  for (int d = 0; d < DIM; d++)
    if (isinf(v[d])) return false;
  return true;
}

// The same lookup in a Q2 velocity field (vertex grid of 2 res + 1 per direction): element search
// on the element grid, Newton on the element's 9 / 27 vertices with the quadratic shape
// functions (Biquadratic / Triquadratic element types), value summed over the vertices in
// incidence order.
template <int DIM>
__device__ bool velocity_at_q2(const GeomArgs& g, const double* vel, const double* x, double* v) {
  int ijk[3];
  if (!search_element_ijk<DIM>(g.res, g.lo, g.hi, x, ijk)) return false;
  constexpr int NN = DIM == 3 ? 27 : 9;
  double verts[NN * DIM], xi[3], N[NN];
  q2_element_verts_ijk<DIM>(g.res, g.lo, g.hi, ijk, verts);
  q2_newton_local<DIM>(verts, x, xi);
  q2_basis<DIM>(xi, N);
  const long i = ijk[0], j = ijk[1], k = ijk[2];
  const long nx = 2L * g.res[0] + 1, ny = 2L * g.res[1] + 1;
  for (int d = 0; d < DIM; d++) v[d] = 0.0;
  for (int a = 0; a < NN; a++) {
    const long n = (2 * i + a % 3) + nx * ((2 * j + (a / 3) % 3) + ny * (DIM == 3 ? 2 * k + a / 9 : 0));
    for (int d = 0; d < DIM; d++) v[d] += vel[n * DIM + d] * N[a];
  }
  for (int d = 0; d < DIM; d++)
    if (isinf(v[d])) return false;
  return true;
}

template <int DIM, int ORDER>
__device__ __forceinline__ bool velocity_lookup(const GeomArgs& g, const double* vel, const double* x, double* v) {
  if constexpr (ORDER == 2)
    return velocity_at_q2<DIM>(g, vel, x, v);
  else
    return velocity_at<DIM>(g, vel, x, v);
}

template <int DIM, int ORDER>
__global__ __launch_bounds__(256, 2) void k_advect_rk2(GeomArgs g, long P, const double* vel, double* coords, double dt,
                                                    int* counts) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double x[3], start[3], v0[3], v1[3];
  for (int d = 0; d < DIM; d++) start[d] = x[d] = coords[p * DIM + d];
  if (!velocity_lookup<DIM, ORDER>(g, vel, x, v0)) {
    atomicAdd(&counts[0], 1);
    return;
  }
  for (int d = 0; d < DIM; d++) x[d] = start[d] + 0.5 * dt * v0[d];
  if (velocity_lookup<DIM, ORDER>(g, vel, x, v1)) {
    for (int d = 0; d < DIM; d++) x[d] = start[d] + dt * v1[d];
  } else {
    for (int d = 0; d < DIM; d++) x[d] = start[d];
    velocity_lookup<DIM, ORDER>(g, vel, x, v0);
    for (int d = 0; d < DIM; d++) x[d] = start[d] + dt * v0[d];
    atomicAdd(&counts[1], 1);
  }
  for (int d = 0; d < DIM; d++) coords[p * DIM + d] = x[d];
}

// g: the element grid; order 2: vel lives on the Q2 vertex grid
void launch_advect_rk2(const Geom& g, long P, const double* vel, double* coords, double dt, int* d_counts,
                       hipStream_t s, int order) {
  if (P == 0) return;
  if (order == 2) {
    if (g.dim == 3)
      hipLaunchKernelGGL((k_advect_rk2<3, 2>), nblocks(P, 256), 256, 0, s, pack(g), P, vel, coords, dt, d_counts);
    else
      hipLaunchKernelGGL((k_advect_rk2<2, 2>), nblocks(P, 256), 256, 0, s, pack(g), P, vel, coords, dt, d_counts);
  } else {
    if (g.dim == 3)
      hipLaunchKernelGGL((k_advect_rk2<3, 1>), nblocks(P, 256), 256, 0, s, pack(g), P, vel, coords, dt, d_counts);
    else
      hipLaunchKernelGGL((k_advect_rk2<2, 1>), nblocks(P, 256), 256, 0, s, pack(g), P, vel, coords, dt, d_counts);
  }
  UWG_HIP(hipGetLastError());
}

template <int DIM>
__global__ void k_interp(GeomArgs g, long P, const double* vel, const double* x, double* v, int* ok) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double xx[3], vv[3];
  for (int d = 0; d < DIM; d++) xx[d] = x[p * DIM + d];
  bool f = velocity_at<DIM>(g, vel, xx, vv);
  ok[p] = f ? 1 : 0;
  for (int d = 0; d < DIM; d++) v[p * DIM + d] = f ? vv[d] : 0.0;
}

void launch_interpolate(const Geom& g, long P, const double* vel, const double* x, double* v, int* ok,
                        hipStream_t s) {
  if (P == 0) return;
  if (g.dim == 3)
    hipLaunchKernelGGL(k_interp<3>, nblocks(P, 256), 256, 0, s, pack(g), P, vel, x, v, ok);
  else
    hipLaunchKernelGGL(k_interp<2>, nblocks(P, 256), 256, 0, s, pack(g), P, vel, x, v, ok);
  UWG_HIP(hipGetLastError());
}

// A18 max over nodes of |v| (FeVariable_GetScalarAtNode, FeVariable.c:453-469): the max
// is order independent, so block partials give the exact reference value.
template <int DIM>
__global__ void k_speed_max(long N, const double* vel, double* partial) {
  __shared__ double sm[256];
  double m = 0.0;
  for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int d = 0; d < DIM; d++) s += vel[n * DIM + d] * vel[n * DIM + d];
    double c = sqrt(s);
    if (c > m) m = c;
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w && sm[threadIdx.x + w] > sm[threadIdx.x]) sm[threadIdx.x] = sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

void launch_node_speed_max(const Geom& g, const double* vel, double* partial, int nblk, hipStream_t s) {
  if (g.dim == 3)
    hipLaunchKernelGGL(k_speed_max<3>, nblk, 256, 0, s, g.N, vel, partial);
  else
    hipLaunchKernelGGL(k_speed_max<2>, nblk, 256, 0, s, g.N, vel, partial);
  UWG_HIP(hipGetLastError());
}

}  // namespace uwg

namespace uwg {
// particle arrays rebuilt after population control: coords/material of particle i from src[i]
__global__ void k_gather_particles(long P, int dim, const long* __restrict__ src, const double* __restrict__ coords,
                                   const int* __restrict__ mat, double* __restrict__ ncoords, int* __restrict__ nmat) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const long s = src[i];
  for (int d = 0; d < dim; d++) ncoords[i * dim + d] = coords[s * dim + d];
  nmat[i] = mat[s];
}
__global__ void k_scatter_coords(long n, int dim, const int* __restrict__ idx, const double* __restrict__ val,
                                 double* __restrict__ coords) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int d = 0; d < dim; d++) coords[(long)idx[k] * dim + d] = val[k * dim + d];
}
void launch_gather_particles(long P, int dim, const long* src, const double* coords, const int* mat, double* ncoords,
                             int* nmat, hipStream_t s) {
  if (P == 0) return;
  hipLaunchKernelGGL(k_gather_particles, (unsigned)((P + 255) / 256), 256, 0, s, P, dim, src, coords, mat, ncoords,
                     nmat);
  UWG_HIP(hipGetLastError());
}
void launch_scatter_coords(long n, int dim, const int* idx, const double* val, double* coords, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scatter_coords, (unsigned)((n + 255) / 256), 256, 0, s, n, dim, idx, val, coords);
  UWG_HIP(hipGetLastError());
}
}  // namespace uwg
