// Particle relocation after advection: new owners plus per-cell particle tables edited
// exactly as Swarm_UpdateAllParticleOwners does (LU/StgDomain/Swarm/src/SwarmClass.c:
// 947-1030): particles are visited in index order; a particle that changed cell is removed
// from its old table by swap-with-last (Swarm_RemoveParticleFromCell) and appended to its
// new one (Swarm_AddParticleToCell).  The edits that touch one cell are exactly that
// cell's removals and additions in particle-index order, so every cell can replay its own
// events independently: stable radix sorts group the moved particles by old and by new
// cell, and one thread per cell replays the merged sequence.
#include <hipcub/hipcub.hpp>

#include "uwg_internal.h"

namespace uwg {

struct RelocWork {
  long P, E;
  int* new_owner;
  uint8_t* moved;
  int* ids;        // 0..P-1
  int* mv;         // moved particle ids (ascending)
  int* nmv;        // [1] on device
  int* key_a;      // new owner of moved
  int* key_r;      // old owner of moved
  int* key_a_s;
  int* key_r_s;
  int* val_a_s;    // moved ids sorted by new owner (stable)
  int* val_r_s;    // moved ids sorted by old owner (stable)
  int* cnt_a;      // [E+1]
  int* cnt_r;      // [E+1]
  int* off_a;      // [E+1]
  int* off_r;      // [E+1]
  int* new_cnt;    // [E+1]
  int* new_ptr;    // [E+1]
  int* tmp_ptr;    // [E+1] scratch offsets (old + additions)
  int* tmp_cnt;
  int* tmp;        // scratch tables, size P + P
  int* new_pid;    // [P]
  int* outside;    // [1]
  void* cub_tmp;
  size_t cub_bytes;
};

static inline unsigned nb(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__global__ void k_iota(long n, int* x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (int)i;
}
__global__ void k_moved(long P, const int* oldo, const int* newo, uint8_t* moved) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P) moved[p] = oldo[p] != newo[p];
}
__global__ void k_keys(const int* nmv, const int* mv, const int* oldo, const int* newo, int* ka, int* kr) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < *nmv) {
    int p = mv[i];
    ka[i] = newo[p];
    kr[i] = oldo[p];
  }
}
__global__ void k_hist(const int* nmv, const int* keys, int* cnt) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < *nmv) atomicAdd(&cnt[keys[i]], 1);
}
__global__ void k_newcnt(long E, const int* ptr, const int* ca, const int* cr, int* nc, int* tc) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < E) {
    int old = ptr[e + 1] - ptr[e];
    nc[e] = old - cr[e] + ca[e];
    tc[e] = old + ca[e];
  }
}
// Per cell, replay removals/additions in particle-index order.  Eight lanes per cell copy the
// cell's list in and out (consecutive lanes, consecutive entries); the edits themselves are
// sequential (a removal moves the list's last entry into the hole) and run on the cell's first
// lane between two workgroup barriers, which also make the lanes' global writes of the list
// visible to each other.
constexpr int kReplayLanes = 8, kReplayBlock = 256;
__global__ __launch_bounds__(kReplayBlock) void k_replay(long E, const int* ptr, const int* pid, const int* off_a,
                                                         const int* cnt_a, const int* val_a, const int* off_r,
                                                         const int* cnt_r, const int* val_r, const int* tmp_ptr,
                                                         int* tmp, const int* new_ptr, int* new_pid) {
  __shared__ int nfin[kReplayBlock / kReplayLanes];
  const int lane = threadIdx.x & (kReplayLanes - 1), c = threadIdx.x / kReplayLanes;
  const long e = ((long)blockIdx.x * blockDim.x + threadIdx.x) / kReplayLanes;
  const bool live = e < E;
  int s0 = 0, n = 0, na = 0, nr = 0;
  int* out = nullptr;
  int* L = nullptr;
  if (live) {
    s0 = ptr[e];
    n = ptr[e + 1] - s0;
    na = cnt_a[e];
    nr = cnt_r[e];
    out = new_pid + new_ptr[e];
    L = tmp + tmp_ptr[e];
  }
  const bool edit = live && (na != 0 || nr != 0);
  if (live) {
    int* dst = edit ? L : out;
    for (int k = lane; k < n; k += kReplayLanes) dst[k] = pid[s0 + k];
  }
  __syncthreads();
  if (edit && lane == 0) {
    const int* A = val_a + off_a[e];
    const int* R = val_r + off_r[e];
    int ia = 0, ir = 0, m = n;
    while (ia < na || ir < nr) {
      bool take_r = ir < nr && (ia >= na || R[ir] < A[ia]);
      if (take_r) {
        int p = R[ir++];
        int k = 0;
        while (k < m && L[k] != p) k++;
        if (k < m) {
          L[k] = L[m - 1];
          m--;
        }
      } else {
        L[m++] = A[ia++];
      }
    }
    nfin[c] = m;
  }
  __syncthreads();
  if (edit) {
    const int m = nfin[c], nmax = new_ptr[e + 1] - new_ptr[e];
    for (int k = lane; k < m && k < nmax; k += kReplayLanes) out[k] = L[k];
  }
}
__global__ void k_count_outside(long P, const int* owner, long E, int* out) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P && owner[p] >= E) atomicAdd(out, 1);
}

RelocWork* reloc_create(long P, long E) {
  RelocWork* w = new RelocWork();
  w->P = P;
  w->E = E;
  long Pm = P > 0 ? P : 1;
  UWG_HIP(hipMalloc(&w->new_owner, sizeof(int) * Pm));
  UWG_HIP(hipMalloc(&w->moved, Pm));
  UWG_HIP(hipMalloc(&w->ids, sizeof(int) * Pm));
  UWG_HIP(hipMalloc(&w->mv, sizeof(int) * Pm));
  UWG_HIP(hipMalloc(&w->nmv, sizeof(int) * 2));
  UWG_HIP(hipMalloc(&w->key_a, sizeof(int) * Pm));
  UWG_HIP(hipMalloc(&w->key_r, sizeof(int) * Pm));
  UWG_HIP(hipMalloc(&w->key_a_s, sizeof(int) * Pm));
  UWG_HIP(hipMalloc(&w->key_r_s, sizeof(int) * Pm));
  UWG_HIP(hipMalloc(&w->val_a_s, sizeof(int) * Pm));
  UWG_HIP(hipMalloc(&w->val_r_s, sizeof(int) * Pm));
  for (int** a : {&w->cnt_a, &w->cnt_r, &w->off_a, &w->off_r, &w->new_cnt, &w->new_ptr, &w->tmp_ptr, &w->tmp_cnt})
    UWG_HIP(hipMalloc(a, sizeof(int) * (E + 2)));
  UWG_HIP(hipMalloc(&w->tmp, sizeof(int) * 2 * Pm));
  UWG_HIP(hipMalloc(&w->new_pid, sizeof(int) * Pm));
  UWG_HIP(hipMalloc(&w->outside, sizeof(int)));
  hipLaunchKernelGGL(k_iota, nb(Pm, 256), 256, 0, 0, Pm, w->ids);
  // temp storage: max over the primitives used
  size_t b1 = 0, b2 = 0, b3 = 0;
  hipcub::DeviceSelect::Flagged(nullptr, b1, w->ids, w->moved, w->mv, w->nmv, (int)Pm);
  hipcub::DeviceRadixSort::SortPairs(nullptr, b2, w->key_a, w->key_a_s, w->mv, w->val_a_s, (int)Pm);
  hipcub::DeviceScan::ExclusiveSum(nullptr, b3, w->cnt_a, w->off_a, (int)(E + 1));
  w->cub_bytes = b1;
  if (b2 > w->cub_bytes) w->cub_bytes = b2;
  if (b3 > w->cub_bytes) w->cub_bytes = b3;
  UWG_HIP(hipMalloc(&w->cub_tmp, w->cub_bytes));
  return w;
}

void reloc_destroy(RelocWork* w) {
  if (!w) return;
  for (void* p : {(void*)w->new_owner, (void*)w->moved, (void*)w->ids, (void*)w->mv, (void*)w->nmv, (void*)w->key_a,
                  (void*)w->key_r, (void*)w->key_a_s, (void*)w->key_r_s, (void*)w->val_a_s, (void*)w->val_r_s,
                  (void*)w->cnt_a, (void*)w->cnt_r, (void*)w->off_a, (void*)w->off_r, (void*)w->new_cnt,
                  (void*)w->new_ptr, (void*)w->tmp_ptr, (void*)w->tmp_cnt, (void*)w->tmp, (void*)w->new_pid,
                  (void*)w->outside, w->cub_tmp})
    (void)hipFree(p);
  delete w;
}

long relocate(const Geom& g, long P, const double* coords, int* owner, int* cell_ptr, int* cell_pid, RelocWork* w,
              hipStream_t s) {
  const long E = g.E;
  if (P == 0) return 0;
  launch_cell_of(g, P, coords, w->new_owner, s);
  hipLaunchKernelGGL(k_moved, nb(P, 256), 256, 0, s, P, owner, w->new_owner, w->moved);
  size_t bytes = w->cub_bytes;
  UWG_HIP(hipcub::DeviceSelect::Flagged(w->cub_tmp, bytes, w->ids, w->moved, w->mv, w->nmv, (int)P, s));
  int nmv = 0;
  UWG_HIP(hipMemcpyAsync(&nmv, w->nmv, sizeof(int), hipMemcpyDeviceToHost, s));
  UWG_HIP(hipStreamSynchronize(s));
  UWG_HIP(hipMemsetAsync(w->cnt_a, 0, sizeof(int) * (E + 2), s));
  UWG_HIP(hipMemsetAsync(w->cnt_r, 0, sizeof(int) * (E + 2), s));
  int bits = 1;
  while ((1L << bits) <= E) bits++;
  if (nmv > 0) {
    hipLaunchKernelGGL(k_keys, nb(nmv, 256), 256, 0, s, w->nmv, w->mv, owner, w->new_owner, w->key_a, w->key_r);
    bytes = w->cub_bytes;
    UWG_HIP(hipcub::DeviceRadixSort::SortPairs(w->cub_tmp, bytes, w->key_a, w->key_a_s, w->mv, w->val_a_s, nmv, 0,
                                               bits, s));
    bytes = w->cub_bytes;
    UWG_HIP(hipcub::DeviceRadixSort::SortPairs(w->cub_tmp, bytes, w->key_r, w->key_r_s, w->mv, w->val_r_s, nmv, 0,
                                               bits, s));
    hipLaunchKernelGGL(k_hist, nb(nmv, 256), 256, 0, s, w->nmv, w->key_a, w->cnt_a);
    hipLaunchKernelGGL(k_hist, nb(nmv, 256), 256, 0, s, w->nmv, w->key_r, w->cnt_r);
  }
  bytes = w->cub_bytes;
  UWG_HIP(hipcub::DeviceScan::ExclusiveSum(w->cub_tmp, bytes, w->cnt_a, w->off_a, (int)(E + 1), s));
  bytes = w->cub_bytes;
  UWG_HIP(hipcub::DeviceScan::ExclusiveSum(w->cub_tmp, bytes, w->cnt_r, w->off_r, (int)(E + 1), s));
  hipLaunchKernelGGL(k_newcnt, nb(E, 256), 256, 0, s, E, cell_ptr, w->cnt_a, w->cnt_r, w->new_cnt, w->tmp_cnt);
  UWG_HIP(hipMemsetAsync(w->new_cnt + E, 0, sizeof(int), s));
  UWG_HIP(hipMemsetAsync(w->tmp_cnt + E, 0, sizeof(int), s));
  bytes = w->cub_bytes;
  UWG_HIP(hipcub::DeviceScan::ExclusiveSum(w->cub_tmp, bytes, w->new_cnt, w->new_ptr, (int)(E + 1), s));
  bytes = w->cub_bytes;
  UWG_HIP(hipcub::DeviceScan::ExclusiveSum(w->cub_tmp, bytes, w->tmp_cnt, w->tmp_ptr, (int)(E + 1), s));
  hipLaunchKernelGGL(k_replay, nb(E * kReplayLanes, kReplayBlock), kReplayBlock, 0, s, E, cell_ptr, cell_pid, w->off_a,
                     w->cnt_a, w->val_a_s, w->off_r,
                     w->cnt_r, w->val_r_s, w->tmp_ptr, w->tmp, w->new_ptr, w->new_pid);
  UWG_HIP(hipGetLastError());
  UWG_HIP(hipMemcpyAsync(cell_ptr, w->new_ptr, sizeof(int) * (E + 1), hipMemcpyDeviceToDevice, s));
  UWG_HIP(hipMemcpyAsync(cell_pid, w->new_pid, sizeof(int) * P, hipMemcpyDeviceToDevice, s));
  UWG_HIP(hipMemcpyAsync(owner, w->new_owner, sizeof(int) * P, hipMemcpyDeviceToDevice, s));
  UWG_HIP(hipMemsetAsync(w->outside, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_count_outside, nb(P, 256), 256, 0, s, P, owner, E, w->outside);
  int outside = 0;
  UWG_HIP(hipMemcpyAsync(&outside, w->outside, sizeof(int), hipMemcpyDeviceToHost, s));
  UWG_HIP(hipStreamSynchronize(s));
  return outside;
}

// Exclusive prefix sum of 0/1 flags into 64-bit counts (the particle layout's skip counts).
void scan_exclusive_flags(const int* flags, long long* out, long n, hipStream_t s) {
  size_t bytes = 0;
  UWG_HIP(hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, flags, out, hipcub::Sum(), 0LL, (int)n, s));
  void* tmp = nullptr;
  UWG_HIP(hipMalloc(&tmp, bytes + 16));
  UWG_HIP(hipcub::DeviceScan::ExclusiveScan(tmp, bytes, flags, out, hipcub::Sum(), 0LL, (int)n, s));
  UWG_HIP(hipStreamSynchronize(s));
  UWG_HIP(hipFree(tmp));
}

}  // namespace uwg
