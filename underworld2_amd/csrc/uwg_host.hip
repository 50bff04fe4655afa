// Host side of libuwg: handles, the C ABI, assembly driver, the geometric multigrid
// hierarchy and the Krylov drivers.  All numerical work runs in the HIP kernels of
// fem_kernels.hip / swarm_kernels.hip / relocate.hip / nn_kernels.hip; this file only
// orchestrates launches and performs O(1)-sized host arithmetic (Hessenberg least
// squares, the dense inverse of the coarsest multigrid level).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>

#include "../../include/uwg.h"
#include "uwg_geom.h"
#include "uwg_internal.h"
#include "uwg_nn.h"

namespace uwg {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

template <typename T>
static T* dalloc(long n) {
  T* p = nullptr;
  if (n <= 0) n = 1;
  UWG_HIP(hipMalloc(&p, sizeof(T) * n));
  return p;
}
static void dfree(void* p) {
  if (p) (void)hipFree(p);
}

// Per-handle workspace: persistent device buffers reused across solves (grown on demand), so
// a solve performs no device allocation once the first solve has run.
enum WsSlot {
  WS_FU, WS_HP, WS_U, WS_P, WS_UBC_FLAG, WS_LIFT_T, WS_LIFT_Q, WS_FREEMASK,
  WS_RES_FU, WS_RES_HP, WS_RES_UF, WS_RES_T, WS_RES_GP, WS_RES_Q,
  WS_BLK_B, WS_BLK_X, WS_BLK_DS, WS_BLK_Q, WS_BLK_GP, WS_BLK_RP,
  WS_SCH_USTAR, WS_SCH_HHAT, WS_SCH_T, WS_SCH_Z, WS_SCH_DS,
  WS_MIX_B, WS_MIX_X, WS_MIX_R, WS_MIX_GP, WS_MIX_R32, WS_MIX_D32, WS_MIX_DS32, WS_MIX_Q32, WS_MIX_GP32,
  WS_Q2_ETAG, WS_Q2_FORCEG, WS_Q2_BAD, WS_COARSE_WORK,
  WS_COUNT
};
static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace uwg

using namespace uwg;

// ---------------------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------------------
struct uwg_mesh {
  Geom g;          // element grid (the Q1 vertex grid for order 1)
  Geom gn;         // vertex grid of the velocity discretisation: g for Q1, 2 res for Q2
  int order = 1;   // 1: Q1/dQ0, 2: Q2/dPc1
  double min_sep;
};

struct uwg_field {
  uwg_mesh* mesh;
  int ncomp;
  int on_elements;
  long n;  // entries
  double* d;
};

struct uwg_swarm {
  uwg_mesh* mesh;
  long P = 0, cap = 0;
  double* coords = nullptr;   // [cap][dim]
  int* owner = nullptr;       // [cap]
  int* cell_ptr = nullptr;    // [E+1]
  int* cell_pid = nullptr;    // [cap]
  int* mat = nullptr;         // [cap]
  double* xi_slot = nullptr;  // [cap][dim]
  double* w_slot = nullptr;   // [cap]
  int* slot_cell = nullptr;   // [cap]
  bool has_mat = false;
  bool has_integration = false;
  long state = 0;             // bumps whenever particles or their materials change (assembly cache)
  long geom_state = 0;        // bumps whenever particles move (integration-point cache)
  long integration_state = -1;
  RelocWork* reloc = nullptr;
  long reloc_P = -1;
  // swarm variables uploaded for device function evaluation: [P][var_nc] doubles per slot
  double* var[8] = {};
  int var_nc[8] = {};
  long var_rows[8] = {};
};

namespace uwg {

struct KrylovSpace;
template <typename T, typename TV = T> struct KSpaceT;

struct Level {
  Geom g;
  double* W = nullptr;
  float* W32 = nullptr;
  uint32_t* Wbf = nullptr;  // 3D FP32 cycle: the weights as bfloat16 pairs [14][Es] (when enabled)
  uint8_t* bc = nullptr;
  double* dinv = nullptr;
  float* dinv32 = nullptr;
  double lmax = 0.0;
  double *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr;
  float *x32 = nullptr, *b32 = nullptr, *r32 = nullptr, *d32 = nullptr;
  long n() const { return g.N * g.dim; }
};

struct Coarse {
  int n = 0;
  double* Ainv = nullptr;  // device, n x n
  long cap = 0;            // doubles allocated at Ainv
  void reserve(long nn) {
    if (nn <= cap && Ainv) return;
    dfree(Ainv);
    Ainv = dalloc<double>(nn);
    cap = nn;
  }
};

// Q2/dPc1 finest level: quadratic velocity on the vertex grid gn, d+1 pressure values per
// element.  Its multigrid correction is the Q1 hierarchy of the same element grid (L[0],
// Galerkin operator of the embedded trilinear space).
struct Q2Level {
  Geom gn;
  uint8_t* bc = nullptr;
  double* W = nullptr;      // [E][5^dim]
  float* W32 = nullptr;
  double* dinv = nullptr;
  float* dinv32 = nullptr;
  double lmax = 0.0;
  double *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr;
  float *x32 = nullptr, *b32 = nullptr, *r32 = nullptr, *d32 = nullptr;
  double* Gr = nullptr;     // G_e [27*dim or 9*dim][dim+1]
  double* S = nullptr;      // approxS blocks [E][dim+1][dim+1]
  double* Sinv = nullptr;   // their inverses
  float* Sinv32 = nullptr;
  double* fe = nullptr;     // [E][3^dim][dim]
  double* D = nullptr;      // 3D: derivative matrix of the matrix-core apply [32][384]
  float* D32 = nullptr;
  long n() const { return gn.N * gn.dim; }
};

}  // namespace uwg

struct uwg_stokes {
  uwg_field* vel;
  uwg_field* p;
  Geom g;
  std::vector<Level> L;
  Q2Level* q2 = nullptr;      // set for a Q2/dPc1 system
  std::vector<double> eta_g_prev;  // Gauss-point viscosities of the last Gauss-swarm assembly
  long nu = 0, np = 0;        // velocity dofs (all vertices x dim) and pressure dofs
  Coarse coarse;
  bool hierarchy_built = false;
  bool assembled = false;
  bool fp32_ready = false;
  bool fp32_bf16 = false;  // the FP32 data were built with bfloat16 cycle weights
  double Ge[24];
  // assembly products (level 0)
  double* fe = nullptr;       // [E][NN*dim]
  double* f = nullptr;        // [N][dim] rhs (BC dofs zero)
  double* approxS = nullptr;  // [E]
  double* ubc = nullptr;      // [N][dim] Dirichlet values (zero elsewhere)
  double* h = nullptr;        // [E] continuity rhs
  // materials
  int nmat = 0;
  double* eta_tab = nullptr;
  double* force_tab = nullptr;
  double* eta_p = nullptr;
  double* force_p = nullptr;
  long np_vals = 0;
  int mode = 0;               // 0 materials, 1 per-particle values
  uwg_swarm* swarm = nullptr;
  long assembled_state = -1;
  // Krylov workspace
  double* red_partial = nullptr;
  double* red_out = nullptr;
  double* dcoef = nullptr;
  double* dh = nullptr;  // Gram-Schmidt coefficients of the current Arnoldi step (device)
  const double** dptrs = nullptr;
  int levels_req = 0;
  // learned correction
  std::unique_ptr<NNModel, NNDeleter> nn;
  int nn_level = -1;
  // learned Schur stencil: network, its coefficients a[27][E] for the current assembly
  std::unique_ptr<SchurNet, SchurNetDeleter> snn;
  uint16_t* snn_a = nullptr;  // bfloat16
  long snn_stamp = -1;        // assembly_stamp the coefficients were computed for
  long assembly_stamp = 0;    // bumped by every assembly (approxS changes)
  int vcycles = 0;
  int nn_rejects = 0;
  uint8_t* bc0 = nullptr;     // level-0 Dirichlet mask (owned here, shared with L[0].bc)
  KrylovSpace* kin = nullptr; // Krylov workspaces, owned by the handle
  KrylovSpace* kout = nullptr;
  KrylovSpace* kb = nullptr;
  KSpaceT<float>* kf = nullptr;  // FP32 inner Krylov space of the mixed-precision solver
  KSpaceT<double, float>* kb32 = nullptr;  // FP64 Krylov space with an FP32 Arnoldi basis
  // learned-correction support
  bool nn_disabled = false;      // set when the FP64 residual shows the model is not helping
  int nn_window = 20;            // fallback monitor: over nn_window iterations the residual must
  double nn_min_drop = 0.01;     // fall by at least this fraction while the network is in use
  double nn_rise = 1.25;         // and, across a restart, may not rise above this factor of the
                                 // previous cycle's final estimate (0: no such test)
  double exact_vel_rtol = 0.0;   // > 0: the coupled FP64 preconditioner's velocity block is an
                                 // inner solve to this rtol instead of one V-cycle (diagnostic)
  int nn_mode = 0;               // 0: replace the coarse cycle, 1: add to it
  int capture_level = -1;        // record coarse vectors seen by the cycle
  int capture_kind = 0;          // 0: right-hand side reaching the level, 1: residual left by its cycle
  int capture_max = 0, captured = 0, capture_every = 1, capture_seen = 0;
  float* capture_buf = nullptr;
  // persistent per-solve buffers (see WsSlot)
  void* ws_ptr[WS_COUNT] = {};
  size_t ws_size[WS_COUNT] = {};
};

template <typename T>
static T* wsbuf(uwg_stokes* st, WsSlot slot, long n) {
  const size_t bytes = sizeof(T) * (size_t)(n > 0 ? n : 1);
  if (st->ws_size[slot] < bytes) {
    dfree(st->ws_ptr[slot]);
    st->ws_ptr[slot] = nullptr;
    st->ws_size[slot] = 0;
    UWG_HIP(hipMalloc(&st->ws_ptr[slot], bytes));
    st->ws_size[slot] = bytes;
  }
  return (T*)st->ws_ptr[slot];
}

// ---------------------------------------------------------------------------------------
// error wrapper
// ---------------------------------------------------------------------------------------
#define API_BEGIN try {
#define API_END                                 \
  }                                             \
  catch (const std::exception& ex) {            \
    set_error(ex.what());                       \
    return UWG_ERR_DEVICE;                      \
  }                                             \
  catch (...) {                                 \
    set_error("unknown error");                 \
    return UWG_ERR_DEVICE;                      \
  }

static int arg_error(const char* m) {
  set_error(m);
  return UWG_ERR_ARG;
}
static int arg_error(const std::string& m) { return arg_error(m.c_str()); }
static int domain_error(const char* m) {
  set_error(m);
  return UWG_ERR_DOMAIN;
}

extern "C" const char* uwg_last_error(void) { return g_err.c_str(); }

extern "C" int uwg_device_info(char* buf, int n) {
  API_BEGIN
  hipDeviceProp_t prop;
  int dev = 0;
  UWG_HIP(hipGetDevice(&dev));
  UWG_HIP(hipGetDeviceProperties(&prop, dev));
  snprintf(buf, n, "%s (%s), %d CUs, %.1f GB", prop.name, prop.gcnArchName, prop.multiProcessorCount,
           prop.totalGlobalMem / 1e9);
  return 0;
  API_END
}

// ---------------------------------------------------------------------------------------
// mesh and fields
// ---------------------------------------------------------------------------------------
extern "C" int uwg_mesh_create(int dim, const int* res, const double* minc, const double* maxc, uwg_mesh** out) {
  if (dim != 2 && dim != 3) return arg_error("dim must be 2 or 3");
  for (int d = 0; d < dim; d++) {
    if (res[d] < 1) return arg_error("element resolution must be >= 1");
    if (!(maxc[d] > minc[d])) return arg_error("maxCoord must exceed minCoord");
  }
  API_BEGIN
  uwg_mesh* m = new uwg_mesh();
  m->g = make_geom(dim, res, minc, maxc);
  // Mesh_HexType_GetMinimumSeparationGeneral: min over elements of pairwise node distances
  double ms = HUGE_VAL;
  const int nn = dim == 3 ? 8 : 4;
  for (long e = 0; e < m->g.E; e++) {
    // only distinct per-axis spacings matter on a tensor grid, but the reference visits
    // every element; we do the same on the host (setup only)
    double v[24];
    if (dim == 3)
      element_verts<3>(m->g.res, m->g.lo, m->g.hi, e, v);
    else
      element_verts<2>(m->g.res, m->g.lo, m->g.hi, e, v);
    for (int a = 0; a < nn - 1; a++)
      for (int b = a + 1; b < nn; b++) {
        double s = 0.0;
        for (int d = 0; d < dim; d++) {
          double t = v[a * dim + d] - v[b * dim + d];
          s += t * t;
        }
        s = sqrt(s);
        if (s < ms) ms = s;
      }
    if (e > 4096 && m->g.E > 100000) break;  // regular spacing: remaining elements repeat
  }
  m->min_sep = ms;
  m->gn = m->g;
  *out = m;
  return 0;
  API_END
}

// Q2 mesh (C2Generator): (2 res + 1)^dim vertices on the same element grid.  The minimum
// separation is again the minimum over elements of the pairwise distances of the element's
// vertices, now its 3^dim vertices (Mesh_HexType_GetMinimumSeparationGeneral).
extern "C" int uwg_mesh_create_q2(int dim, const int* res, const double* minc, const double* maxc, uwg_mesh** out) {
  int rc = uwg_mesh_create(dim, res, minc, maxc, out);
  if (rc) return rc;
  API_BEGIN
  uwg_mesh* m = *out;
  int r2[3];
  for (int d = 0; d < 3; d++) r2[d] = d < dim ? 2 * res[d] : 1;
  m->gn = make_geom(dim, r2, minc, maxc);
  m->order = 2;
  double ms = HUGE_VAL;
  const int nn = dim == 3 ? 27 : 9;
  for (long e = 0; e < m->g.E; e++) {
    double v[81];
    if (dim == 3)
      q2_element_verts<3>(m->g.res, m->g.lo, m->g.hi, e, v);
    else
      q2_element_verts<2>(m->g.res, m->g.lo, m->g.hi, e, v);
    for (int a = 0; a < nn - 1; a++)
      for (int b = a + 1; b < nn; b++) {
        double sq = 0.0;
        for (int d = 0; d < dim; d++) {
          double t = v[a * dim + d] - v[b * dim + d];
          sq += t * t;
        }
        sq = sqrt(sq);
        if (sq < ms) ms = sq;
      }
    if (e > 4096 && m->g.E > 100000) break;  // regular spacing: remaining elements repeat
  }
  m->min_sep = ms;
  return 0;
  API_END
}

extern "C" int uwg_mesh_destroy(uwg_mesh* m) {
  delete m;
  return 0;
}

extern "C" int uwg_mesh_coords(uwg_mesh* m, double* out) {
  API_BEGIN
  double* d = dalloc<double>(m->gn.N * m->gn.dim);
  launch_mesh_coords(m->gn, d, 0);
  UWG_HIP(hipMemcpy(out, d, sizeof(double) * m->gn.N * m->gn.dim, hipMemcpyDeviceToHost));
  dfree(d);
  return 0;
  API_END
}

extern "C" int uwg_mesh_min_separation(uwg_mesh* m, double* out) {
  *out = m->min_sep;
  return 0;
}

extern "C" int uwg_field_create(uwg_mesh* m, int ncomp, int on_elements, uwg_field** out) {
  if (ncomp < 1) return arg_error("ncomp must be >= 1");
  if (on_elements < 0 || on_elements > 2) return arg_error("on_elements must be 0 (vertices), 1 (dQ0) or 2 (dPc1)");
  API_BEGIN
  uwg_field* f = new uwg_field();
  f->mesh = m;
  f->ncomp = ncomp;
  f->on_elements = on_elements;
  const long nent = on_elements == 0 ? m->gn.N : (on_elements == 1 ? m->g.E : m->g.E * (m->g.dim + 1));
  f->n = nent * ncomp;
  f->d = dalloc<double>(f->n);
  UWG_HIP(hipMemset(f->d, 0, sizeof(double) * f->n));
  *out = f;
  return 0;
  API_END
}
extern "C" int uwg_field_destroy(uwg_field* f) {
  if (f) dfree(f->d);
  delete f;
  return 0;
}
extern "C" int uwg_field_set(uwg_field* f, const double* host) {
  API_BEGIN
  UWG_HIP(hipMemcpy(f->d, host, sizeof(double) * f->n, hipMemcpyHostToDevice));
  return 0;
  API_END
}
extern "C" int uwg_field_get(uwg_field* f, double* host) {
  API_BEGIN
  UWG_HIP(hipMemcpy(host, f->d, sizeof(double) * f->n, hipMemcpyDeviceToHost));
  return 0;
  API_END
}

// ---------------------------------------------------------------------------------------
// swarm
// ---------------------------------------------------------------------------------------
static void swarm_reserve(uwg_swarm* s, long P) {
  if (P <= s->cap) return;
  long cap = std::max(P, s->cap * 2);
  int dim = s->mesh->g.dim;
  auto grow = [&](auto*& ptr, long per) {
    using T = std::remove_reference_t<decltype(*ptr)>;
    T* n = dalloc<T>(cap * per);
    if (ptr && s->P > 0) UWG_HIP(hipMemcpy(n, ptr, sizeof(T) * s->P * per, hipMemcpyDeviceToDevice));
    dfree(ptr);
    ptr = n;
  };
  grow(s->coords, dim);
  grow(s->owner, 1);
  grow(s->cell_pid, 1);
  grow(s->mat, 1);
  grow(s->xi_slot, dim);
  grow(s->w_slot, 1);
  grow(s->slot_cell, 1);
  s->cap = cap;
  if (s->reloc) reloc_destroy(s->reloc);
  s->reloc = nullptr;
}

extern "C" int uwg_swarm_create(uwg_mesh* m, uwg_swarm** out) {
  API_BEGIN
  uwg_swarm* s = new uwg_swarm();
  s->mesh = m;
  s->cell_ptr = dalloc<int>(m->g.E + 1);
  UWG_HIP(hipMemset(s->cell_ptr, 0, sizeof(int) * (m->g.E + 1)));
  *out = s;
  return 0;
  API_END
}

extern "C" int uwg_swarm_destroy(uwg_swarm* s) {
  if (!s) return 0;
  for (void* p : {(void*)s->coords, (void*)s->owner, (void*)s->cell_ptr, (void*)s->cell_pid, (void*)s->mat,
                  (void*)s->xi_slot, (void*)s->w_slot, (void*)s->slot_cell})
    dfree(p);
  reloc_destroy(s->reloc);
  for (int v = 0; v < 8; v++) dfree(s->var[v]);
  delete s;
  return 0;
}

static void ensure_reloc(uwg_swarm* s) {
  if (!s->reloc || s->reloc_P != s->P) {
    reloc_destroy(s->reloc);
    s->reloc = reloc_create(s->P, s->mesh->g.E);
    s->reloc_P = s->P;
  }
}

static void populate_tables(uwg_swarm* s, int ppc);

extern "C" int uwg_swarm_populate_random(uwg_swarm* s, int ppc, long seed) {
  if (ppc < 1) return arg_error("particlesPerCell must be >= 1");
  if (s->P != 0) return arg_error("swarm already populated");
  API_BEGIN
  const Geom& g = s->mesh->g;
  long P = g.E * ppc;
  swarm_reserve(s, P);
  s->P = P;
  launch_layout_random(g, ppc, seed, s->coords, 0);
  populate_tables(s, ppc);
  return 0;
  API_END
}

extern "C" int uwg_swarm_populate_sobol(uwg_swarm* s, int ppc, long seed) {
  if (ppc < 1) return arg_error("particlesPerCell must be >= 1");
  if (s->P != 0) return arg_error("swarm already populated");
  const long P = s->mesh->g.E * ppc;
  if (2 * P + 2 >= (1L << 30)) return arg_error("PerCellSpaceFillerLayout: more particles than the 30-bit Sobol generator serves");
  API_BEGIN
  swarm_reserve(s, P);
  s->P = P;
  launch_layout_sobol(s->mesh->g, ppc, seed, s->coords, 0);
  populate_tables(s, ppc);
  return 0;
  API_END
}

// cell tables, owners and materials of a freshly laid-out swarm: particle index = cell * ppc + k
static void populate_tables(uwg_swarm* s, int ppc) {
  const Geom& g = s->mesh->g;
  const long P = s->P;
  std::vector<int> ptr(g.E + 1);
  for (long e = 0; e <= g.E; e++) ptr[e] = (int)(e * ppc);
  UWG_HIP(hipMemcpy(s->cell_ptr, ptr.data(), sizeof(int) * (g.E + 1), hipMemcpyHostToDevice));
  std::vector<int> ids(P);
  for (long p = 0; p < P; p++) ids[p] = (int)p;
  UWG_HIP(hipMemcpy(s->cell_pid, ids.data(), sizeof(int) * P, hipMemcpyHostToDevice));
  for (long p = 0; p < P; p++) ids[p] = (int)(p / ppc);
  UWG_HIP(hipMemcpy(s->owner, ids.data(), sizeof(int) * P, hipMemcpyHostToDevice));
  UWG_HIP(hipMemset(s->mat, 0, sizeof(int) * P));
  s->state++;
  s->geom_state++;
}

extern "C" int uwg_swarm_add_particles(uwg_swarm* s, long n, const double* coords) {
  if (n < 0) return arg_error("negative count");
  API_BEGIN
  const Geom& g = s->mesh->g;
  long P0 = s->P;
  swarm_reserve(s, P0 + n);
  UWG_HIP(hipMemcpy(s->coords + P0 * g.dim, coords, sizeof(double) * n * g.dim, hipMemcpyHostToDevice));
  UWG_HIP(hipMemset(s->mat + P0, 0, sizeof(int) * n));
  // new particles start unowned (owner = E) and are appended to their cells in index order
  std::vector<int> unowned(n, (int)g.E);
  UWG_HIP(hipMemcpy(s->owner + P0, unowned.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  s->P = P0 + n;
  ensure_reloc(s);
  long outside = relocate(g, s->P, s->coords, s->owner, s->cell_ptr, s->cell_pid, s->reloc, 0);
  s->state++;
  s->geom_state++;
  if (outside > 0) {
    set_error("particles outside the mesh were added");
    return UWG_ERR_DOMAIN;
  }
  return 0;
  API_END
}

extern "C" int uwg_swarm_count(uwg_swarm* s, long* P) {
  *P = s->P;
  return 0;
}
extern "C" int uwg_swarm_get_coords(uwg_swarm* s, double* out) {
  API_BEGIN
  UWG_HIP(hipMemcpy(out, s->coords, sizeof(double) * s->P * s->mesh->g.dim, hipMemcpyDeviceToHost));
  return 0;
  API_END
}
extern "C" int uwg_swarm_set_coords(uwg_swarm* s, const double* in, long* outside) {
  API_BEGIN
  UWG_HIP(hipMemcpy(s->coords, in, sizeof(double) * s->P * s->mesh->g.dim, hipMemcpyHostToDevice));
  ensure_reloc(s);
  long o = relocate(s->mesh->g, s->P, s->coords, s->owner, s->cell_ptr, s->cell_pid, s->reloc, 0);
  if (outside) *outside = o;
  s->state++;
  s->geom_state++;
  return o > 0 ? UWG_ERR_DOMAIN : 0;
  API_END
}
extern "C" int uwg_swarm_get_owners(uwg_swarm* s, int* out) {
  API_BEGIN
  UWG_HIP(hipMemcpy(out, s->owner, sizeof(int) * s->P, hipMemcpyDeviceToHost));
  return 0;
  API_END
}
extern "C" int uwg_swarm_get_cell_tables(uwg_swarm* s, int* cell_ptr, int* cell_pid) {
  API_BEGIN
  UWG_HIP(hipMemcpy(cell_ptr, s->cell_ptr, sizeof(int) * (s->mesh->g.E + 1), hipMemcpyDeviceToHost));
  UWG_HIP(hipMemcpy(cell_pid, s->cell_pid, sizeof(int) * s->P, hipMemcpyDeviceToHost));
  return 0;
  API_END
}
extern "C" int uwg_swarm_set_material(uwg_swarm* s, const int* mat) {
  API_BEGIN
  UWG_HIP(hipMemcpy(s->mat, mat, sizeof(int) * s->P, hipMemcpyHostToDevice));
  s->has_mat = true;
  s->state++;
  return 0;
  API_END
}
extern "C" int uwg_swarm_update_owners(uwg_swarm* s, long* outside) {
  API_BEGIN
  ensure_reloc(s);
  long o = relocate(s->mesh->g, s->P, s->coords, s->owner, s->cell_ptr, s->cell_pid, s->reloc, 0);
  if (outside) *outside = o;
  s->state++;
  s->geom_state++;
  if (o > 0) {
    set_error("Particles appear to have left the domain, but swarm flag `particleEscape` is False.");
    return UWG_ERR_DOMAIN;
  }
  return 0;
  API_END
}

static int run_dvc(uwg_swarm* s, int dvc_res) {
  const Geom& g = s->mesh->g;
  const long E = g.E;
  // per-cell counts for capacity planning
  std::vector<int> ptr(E + 1);
  UWG_HIP(hipMemcpy(ptr.data(), s->cell_ptr, sizeof(int) * (E + 1), hipMemcpyDeviceToHost));
  int maxc = 0;
  for (long e = 0; e < E; e++) {
    int c = ptr[e + 1] - ptr[e];
    if (c == 0) {
      set_error("Problem has an under resolved cell (Cell Id = " + std::to_string(e) +
                "), add more particles to your model");
      return UWG_ERR_UNDERRESOLVED;
    }
    maxc = std::max(maxc, c);
  }
  const int ncell = dvc_res * dvc_res * (g.dim == 3 ? dvc_res : 1);
  std::vector<int> retry;
  if (dvc_wave_supported(g.dim, dvc_res)) {
    // one wavefront per element; elements it cannot hold are flagged for the per-thread kernel
    hipDeviceProp_t prop;
    int dev = 0;
    UWG_HIP(hipGetDevice(&dev));
    UWG_HIP(hipGetDeviceProperties(&prop, dev));
    const int cap = std::min(512, ncell);
    // 16 workgroups per CU (19 fit in LDS; measured per pass at 128^3: 12 -> 659 ms, 16 -> 559 ms,
    // 19 -> 684 ms)
    const int nblk = (int)std::min<long>(E, (long)prop.multiProcessorCount * 16);
    short* gb = dalloc<short>(dvc_wave_bound_shorts(cap) * nblk);
    int* flag = dalloc<int>(E);
    launch_dvc_wave(g.dim, dvc_res, E, s->cell_ptr, s->xi_slot, s->w_slot, gb, nblk, cap, flag, 0);
    std::vector<int> hflag(E);
    UWG_HIP(hipMemcpy(hflag.data(), flag, sizeof(int) * E, hipMemcpyDeviceToHost));
    for (long e = 0; e < E; e++)
      if (hflag[e]) retry.push_back((int)e);
    dfree(gb);
    dfree(flag);
  } else {
    int max_np = std::min(std::max(64, ((maxc + 15) / 16) * 16), 32000);
    int cap = std::min(512, ncell);
    long per = dvc_scratch_bytes(g.dim, dvc_res, max_np, cap);
    long batch = std::max(1024L, std::min(E, (long)(3.0e9 / per)));
    char* scratch = dalloc<char>(per * batch);
    int* flag = dalloc<int>(batch);
    std::vector<int> hflag(batch);
    for (long e0 = 0; e0 < E; e0 += batch) {
      long nE = std::min(batch, E - e0);
      launch_dvc_list(g.dim, dvc_res, e0, nE, nullptr, s->cell_ptr, s->xi_slot, s->w_slot, scratch, max_np, cap,
                      flag, 0);
      UWG_HIP(hipMemcpy(hflag.data(), flag, sizeof(int) * nE, hipMemcpyDeviceToHost));
      for (long t = 0; t < nE; t++)
        if (hflag[t]) retry.push_back((int)(e0 + t));
    }
    dfree(scratch);
    dfree(flag);
  }
  if (!retry.empty()) {
    // larger lists for the few elements that overflowed
    int cap2 = ncell, np2 = std::max(maxc, 1);
    long per2 = dvc_scratch_bytes(g.dim, dvc_res, np2, cap2);
    long b2 = std::max(1L, std::min((long)retry.size(), (long)(2.0e9 / per2)));
    char* sc2 = dalloc<char>(per2 * b2);
    int* fl2 = dalloc<int>(b2);
    int* el = dalloc<int>(retry.size());
    UWG_HIP(hipMemcpy(el, retry.data(), sizeof(int) * retry.size(), hipMemcpyHostToDevice));
    std::vector<int> hf(b2);
    for (long i0 = 0; i0 < (long)retry.size(); i0 += b2) {
      long n = std::min(b2, (long)retry.size() - i0);
      launch_dvc_list(g.dim, dvc_res, 0, n, el + i0, s->cell_ptr, s->xi_slot, s->w_slot, sc2, np2, cap2, fl2, 0);
      UWG_HIP(hipMemcpy(hf.data(), fl2, sizeof(int) * n, hipMemcpyDeviceToHost));
      for (long t = 0; t < n; t++)
        if (hf[t]) {
          dfree(sc2); dfree(fl2); dfree(el);
          set_error("DVC list capacity exceeded");
          return UWG_ERR_DEVICE;
        }
    }
    dfree(sc2);
    dfree(fl2);
    dfree(el);
  }
  return 0;
}

extern "C" int uwg_swarm_integration_points(uwg_swarm* s, int use_dvc, int dvc_res) {
  API_BEGIN
  const Geom& g = s->mesh->g;
  if (dvc_res <= 0) dvc_res = 15;
  if (dvc_res > 32) return arg_error("DVC resolution must be <= 32");
  launch_slot_cell(g.E, s->cell_ptr, s->slot_cell, 0);
  launch_coincident_map(g, s->P, s->slot_cell, s->cell_pid, s->coords, s->xi_slot, 0);
  if (use_dvc) {
    int rc = run_dvc(s, dvc_res);
    if (rc) return rc;
  } else {
    // equal weights 2^dim / count per cell
    std::vector<int> ptr(g.E + 1);
    UWG_HIP(hipMemcpy(ptr.data(), s->cell_ptr, sizeof(int) * (g.E + 1), hipMemcpyDeviceToHost));
    std::vector<double> w(s->P);
    for (long e = 0; e < g.E; e++) {
      int c = ptr[e + 1] - ptr[e];
      for (int k = ptr[e]; k < ptr[e + 1]; k++) w[k] = (double)(1 << g.dim) / c;
    }
    UWG_HIP(hipMemcpy(s->w_slot, w.data(), sizeof(double) * s->P, hipMemcpyHostToDevice));
  }
  UWG_HIP(hipDeviceSynchronize());
  s->has_integration = true;
  s->integration_state = s->geom_state;
  return 0;
  API_END
}

extern "C" int uwg_swarm_get_integration(uwg_swarm* s, double* xi_slot, double* w_slot) {
  API_BEGIN
  int dim = s->mesh->g.dim;
  UWG_HIP(hipMemcpy(xi_slot, s->xi_slot, sizeof(double) * s->P * dim, hipMemcpyDeviceToHost));
  UWG_HIP(hipMemcpy(w_slot, s->w_slot, sizeof(double) * s->P, hipMemcpyDeviceToHost));
  return 0;
  API_END
}
extern "C" int uwg_swarm_set_integration(uwg_swarm* s, const double* xi_slot, const double* w_slot) {
  API_BEGIN
  int dim = s->mesh->g.dim;
  UWG_HIP(hipMemcpy(s->xi_slot, xi_slot, sizeof(double) * s->P * dim, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(s->w_slot, w_slot, sizeof(double) * s->P, hipMemcpyHostToDevice));
  s->has_integration = true;
  s->integration_state = s->geom_state;
  return 0;
  API_END
}

// ---------------------------------------------------------------------------------------
// advection
// ---------------------------------------------------------------------------------------
extern "C" int uwg_advect_rk2(uwg_swarm* s, uwg_field* vel, double dt, int update_owners, long* outside) {
  if (vel->mesh != s->mesh || vel->on_elements || vel->ncomp != s->mesh->g.dim)
    return arg_error("velocity field must be a nodal vector field on the swarm's mesh");
  API_BEGIN
  int* cnt = dalloc<int>(2);
  UWG_HIP(hipMemset(cnt, 0, sizeof(int) * 2));
  launch_advect_rk2(s->mesh->g, s->P, vel->d, s->coords, dt, cnt, 0, vel->mesh->order);
  int h[2];
  UWG_HIP(hipMemcpy(h, cnt, sizeof(int) * 2, hipMemcpyDeviceToHost));
  dfree(cnt);
  s->state++;
  s->geom_state++;
  if (h[0] > 0) {
    set_error("Unable to determine valid velocity for " + std::to_string(h[0]) +
              " particle(s): coordinate outside the global domain");
    return UWG_ERR_DOMAIN;
  }
  if (update_owners) return uwg_swarm_update_owners(s, outside);
  return 0;
  API_END
}

extern "C" int uwg_advect_max_dt(uwg_field* vel, double* dt) {
  API_BEGIN
  if (vel->on_elements || vel->ncomp != vel->mesh->g.dim) return arg_error("nodal velocity field required");
  const Geom& g = vel->mesh->gn;
  const int nblk = 512;
  double* part = dalloc<double>(nblk);
  launch_node_speed_max(g, vel->d, part, nblk, 0);
  std::vector<double> h(nblk);
  UWG_HIP(hipMemcpy(h.data(), part, sizeof(double) * nblk, hipMemcpyDeviceToHost));
  dfree(part);
  double vmax = 0.0;
  for (double v : h) vmax = std::max(vmax, v);
  *dt = 0.5 * vel->mesh->min_sep / vmax;
  return 0;
  API_END
}

extern "C" int uwg_field_interpolate(uwg_field* f, long n, const double* x, double* v, int* ok) {
  if (f->on_elements) return arg_error("nodal field required");
  if (f->ncomp < 1 || f->ncomp > 9) return arg_error("fields of 1 to 9 components can be interpolated");
  API_BEGIN
  const int dim = f->mesh->g.dim, nc = f->ncomp;
  double* dx = dalloc<double>(n * dim);
  double* dv = dalloc<double>(n * nc);
  int* dok = dalloc<int>(n);
  UWG_HIP(hipMemcpy(dx, x, sizeof(double) * n * dim, hipMemcpyHostToDevice));
  if (f->mesh->order == 2)
    launch_q2_interpolate(f->mesh->g, n, f->d, nc, dx, dv, dok, 0);
  else
    launch_field_interpolate(f->mesh->g, n, f->d, nc, dx, dv, dok, 0);
  UWG_HIP(hipMemcpy(v, dv, sizeof(double) * n * nc, hipMemcpyDeviceToHost));
  UWG_HIP(hipMemcpy(ok, dok, sizeof(int) * n, hipMemcpyDeviceToHost));
  dfree(dx);
  dfree(dv);
  dfree(dok);
  return 0;
  API_END
}

// ---------------------------------------------------------------------------------------
// Allocation-free interpolation.  uwg_field_interpolate stages the host points, values and
// flags in device memory it allocates per call; this variant stages them in a fixed
// file-scope device array instead, for callers that cannot allocate (code under graph
// capture, tight serving loops).  Each GPU holds its own copy of the array until the process
// ends; its device address is looked up once per device (the array's host-side name is not a
// device address).  Calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------------------
constexpr long kNoAllocMaxPoints = UWG_NOALLOC_MAX_POINTS;
constexpr size_t kScrX = 0;                                                   // double[max][3]
constexpr size_t kScrV = kScrX + (size_t)kNoAllocMaxPoints * 3 * sizeof(double);  // double[max][9]
constexpr size_t kScrOk = kScrV + (size_t)kNoAllocMaxPoints * 9 * sizeof(double);  // int[max]
constexpr size_t kScrBytes = kScrOk + (size_t)kNoAllocMaxPoints * sizeof(int);
static __device__ __align__(256) unsigned char scratch[kScrBytes];

static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[64] = {};

// device address of this device's scratch array (looked up on the device's first call)
static unsigned char* scratch_on_current_device() {
  int dev = 0;
  UWG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw std::runtime_error("device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !ptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
    g_scratch_ptr[dev] = (unsigned char*)ptr;
  }
  return g_scratch_ptr[dev];
}

extern "C" int uwg_field_interpolate_noalloc(uwg_field* f, long n, const double* x, double* v, int* ok) {
  if (f->on_elements) return arg_error("nodal field required");
  if (f->ncomp < 1 || f->ncomp > 9) return arg_error("fields of 1 to 9 components can be interpolated");
  if (n < 0 || n > kNoAllocMaxPoints)
    return arg_error("uwg_field_interpolate_noalloc takes at most " + std::to_string(kNoAllocMaxPoints) +
                     " points per call (got " + std::to_string(n) + "); split the points into batches");
  if (n == 0) return 0;
  API_BEGIN
  const int dim = f->mesh->g.dim, nc = f->ncomp;
  unsigned char* base = scratch_on_current_device();
  double* dx = (double*)(base + kScrX);
  double* dv = (double*)(base + kScrV);
  int* dok = (int*)(base + kScrOk);
  UWG_HIP(hipMemcpy(dx, x, sizeof(double) * n * dim, hipMemcpyHostToDevice));
  if (f->mesh->order == 2)
    launch_q2_interpolate(f->mesh->g, n, f->d, nc, dx, dv, dok, 0);
  else
    launch_field_interpolate(f->mesh->g, n, f->d, nc, dx, dv, dok, 0);
  UWG_HIP(hipMemcpy(v, dv, sizeof(double) * n * nc, hipMemcpyDeviceToHost));
  UWG_HIP(hipMemcpy(ok, dok, sizeof(int) * n, hipMemcpyDeviceToHost));
  return 0;
  API_END
}

// ---------------------------------------------------------------------------------------
// Stokes: assembly
// ---------------------------------------------------------------------------------------
// G_e at 3^d Gauss points (GradientStiffnessMatrixTerm.c:129-221); identical for every
// element of the regular mesh.
static void compute_Ge(const Geom& g, double* Ge) {
  const double a = sqrt(3.0 / 5.0);
  const double x1[3] = {-a, 0.0, a}, w1[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  const int dim = g.dim, NN = 1 << dim;
  double h[3];
  for (int d = 0; d < dim; d++) h[d] = (g.hi[d] - g.lo[d]) / g.res[d];
  double detJ = 1.0;
  for (int d = 0; d < dim; d++) detJ *= 0.5 * h[d];
  for (int q = 0; q < NN * dim; q++) Ge[q] = 0.0;
  for (int k = 0; k < (dim == 3 ? 3 : 1); k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        double xi[3] = {x1[i], x1[j], x1[k]};
        double w = w1[i] * w1[j] * (dim == 3 ? w1[k] : 1.0);
        double G[3][8];
        if (dim == 3)
          shape_local_derivs<3>(xi, G);
        else
          shape_local_derivs<2>(xi, G);
        for (int n = 0; n < NN; n++)
          for (int d = 0; d < dim; d++) Ge[n * dim + d] += w * (-detJ) * (G[d][n] * (2.0 / h[d]));
      }
}

static void build_levels(uwg_stokes* st, int levels_req) {
  // free previous levels, keeping the level-0 mask and quadrature weights
  double* W0 = st->L.empty() ? nullptr : st->L[0].W;
  for (size_t l = 0; l < st->L.size(); l++) {
    Level& A = st->L[l];
    for (void* p : {(void*)A.W, (void*)A.W32, (void*)A.Wbf, (void*)A.bc, (void*)A.dinv, (void*)A.dinv32, (void*)A.x,
                    (void*)A.b, (void*)A.r, (void*)A.d, (void*)A.x32, (void*)A.b32, (void*)A.r32, (void*)A.d32})
      if (p != (void*)st->bc0 && p != (void*)W0) dfree(p);
  }
  st->L.clear();
  Geom g = st->g;
  int nl = 1;
  std::vector<Geom> gs{g};
  while (true) {
    if (levels_req > 0 && nl >= levels_req) break;
    bool even = true;
    for (int d = 0; d < g.dim; d++) even = even && (g.res[d] % 2 == 0) && g.res[d] >= 4;
    if (!even) break;
    if (levels_req <= 0 && g.N * g.dim <= 800) break;
    int r2[3];
    for (int d = 0; d < 3; d++) r2[d] = d < g.dim ? g.res[d] / 2 : 1;
    g = make_geom(g.dim, r2, g.lo, g.hi);
    gs.push_back(g);
    nl++;
  }
  st->L.resize(nl);
  for (int l = 0; l < nl; l++) {
    Level& A = st->L[l];
    A.g = gs[l];
    const long Q = A.g.dim == 3 ? 27 : 9;
    if (l == 0 && W0) {
      A.W = W0;
    } else {
      A.W = dalloc<double>(Q * A.g.Es);
      UWG_HIP(hipMemset(A.W, 0, sizeof(double) * Q * A.g.Es));  // padding stays zero
    }
    A.dinv = dalloc<double>(A.n());
    A.x = dalloc<double>(A.n());
    A.b = dalloc<double>(A.n());
    A.r = dalloc<double>(A.n());
    A.d = dalloc<double>(A.n());
    A.bc = l > 0 ? dalloc<uint8_t>(A.g.N) : st->bc0;
  }
  st->hierarchy_built = true;
  st->fp32_ready = false;
}

// quadrature weights of a level, packed W[q][E] on the host (the device copy is padded to Es)
static void copy_W_to_host(const Level& A, double* out) {
  const long Q = A.g.dim == 3 ? 27 : 9;
  UWG_HIP(hipMemcpy2D(out, sizeof(double) * A.g.E, A.W, sizeof(double) * A.g.Es, sizeof(double) * A.g.E, Q,
                      hipMemcpyDeviceToHost));
}

// Whether the 3D FP32 multigrid cycle reads its element weights as bfloat16 (the element
// contractions then take bf16 inputs with FP32 arithmetic and accumulation).
static bool g_cycle_bf16 = true;
static bool g_basis_fp32 = true;  // FP32 Arnoldi basis in the coupled FP64 iteration at precision 1

static void ensure_fp32(uwg_stokes* st) {
  if (st->fp32_ready && st->fp32_bf16 == g_cycle_bf16) return;
  for (auto& A : st->L) {
    const long Q = A.g.dim == 3 ? 27 : 9;
    if (!A.W32) {
      A.W32 = dalloc<float>(Q * A.g.Es);
      A.dinv32 = dalloc<float>(A.n());
      A.x32 = dalloc<float>(A.n());
      A.b32 = dalloc<float>(A.n());
      A.r32 = dalloc<float>(A.n());
      A.d32 = dalloc<float>(A.n());
    }
    vk_d2f(Q * A.g.Es, 1.0, A.W, A.W32, 0);
    vk_d2f(A.n(), 1.0, A.dinv, A.dinv32, 0);
    if (A.g.dim == 3 && g_cycle_bf16) {
      if (!A.Wbf) A.Wbf = dalloc<uint32_t>(14 * A.g.Es);
      launch_pack_w3_bf16(A.g.Es, A.W, A.Wbf, 0);
    } else if (A.Wbf) {
      dfree(A.Wbf);
      A.Wbf = nullptr;
    }
  }
  st->fp32_bf16 = g_cycle_bf16;
  if (st->q2) {
    Q2Level& A = *st->q2;
    const long nq = st->g.E * (st->g.dim == 3 ? 125 : 25), n = A.n(), ns = st->np * (st->g.dim + 1);
    if (!A.W32) {
      A.W32 = dalloc<float>(nq);
      A.dinv32 = dalloc<float>(n);
      A.Sinv32 = dalloc<float>(ns);
      for (float** v : {&A.x32, &A.b32, &A.r32, &A.d32}) {
        *v = dalloc<float>(n);
        vk_fill<float>(n, 0.0f, *v, 0);
      }
    }
    vk_d2f(nq, 1.0, A.W, A.W32, 0);
    vk_d2f(n, 1.0, A.dinv, A.dinv32, 0);
    vk_d2f(ns, 1.0, A.Sinv, A.Sinv32, 0);
  }
  st->fp32_ready = true;
}

// Dense inverse of the coarsest operator (identity rows on Dirichlet dofs).
// K_e[ai][bj] = sum_g W_g (delta_ij sum_k dNa_k dNb_k + dNa_j dNb_i), physical derivatives.
static void build_coarse(uwg_stokes* st) {
  Level& A = st->L.back();
  const Geom& g = A.g;
  const int dim = g.dim, NN = 1 << dim, Q = dim == 3 ? 27 : 9;
  const int n = (int)(g.N * dim);
  st->coarse.n = n;
  if (n > 6000) {
    dfree(st->coarse.Ainv);
    st->coarse.Ainv = nullptr;
    st->coarse.cap = 0;
    return;
  }
  if (n <= 2048) {
    // on the device: dense operator by a deterministic gather, Gauss-Jordan (one launch per pivot)
    st->coarse.reserve((long)n * n);
    int* flag = wsbuf<int>(st, WS_UBC_FLAG, 1);
    double* work = wsbuf<double>(st, WS_COARSE_WORK, (long)n * n);
    UWG_HIP(hipMemsetAsync(flag, 0, sizeof(int), 0));
    launch_coarse_inverse(g, A.W, A.bc, st->coarse.Ainv, work, flag, 0);
    int bad = 0;
    UWG_HIP(hipMemcpy(&bad, flag, sizeof(int), hipMemcpyDeviceToHost));
    if (bad) throw std::runtime_error("coarse operator is not positive definite (check viscosity/BCs)");
    return;
  }
  std::vector<double> W((size_t)Q * g.E);
  copy_W_to_host(A, W.data());
  std::vector<uint8_t> bc(g.N);
  UWG_HIP(hipMemcpy(bc.data(), A.bc, g.N, hipMemcpyDeviceToHost));
  std::vector<double> K((size_t)n * n, 0.0);
  double h[3] = {1, 1, 1};
  for (int d = 0; d < dim; d++) h[d] = (g.hi[d] - g.lo[d]) / g.res[d];
  const double ga = sqrt(0.6);
  auto gp = [&](int i) { return i == 0 ? -ga : (i == 1 ? 0.0 : ga); };
  long nx = g.nn[0], ny = g.nn[1];
  for (long e = 0; e < g.E; e++) {
    long i = e % g.res[0], j = (e / g.res[0]) % g.res[1], k = dim == 3 ? e / ((long)g.res[0] * g.res[1]) : 0;
    long base = i + nx * (j + ny * k);
    long nodes[8];
    for (int a = 0; a < NN; a++) nodes[a] = base + (a & 1) + ((a >> 1) & 1) * nx + (dim == 3 ? ((a >> 2) & 1) * nx * ny : 0);
    for (int q = 0; q < Q; q++) {
      double X[3] = {gp(q % 3), gp((q / 3) % 3), gp(q / 9)};
      double dN[8][3];
      for (int a = 0; a < NN; a++) {
        double s[3] = {(a & 1) ? 1.0 : -1.0, (a & 2) ? 1.0 : -1.0, (a & 4) ? 1.0 : -1.0};
        if (dim == 2) {
          dN[a][0] = 0.25 * s[0] * (1 + s[1] * X[1]) * 2.0 / h[0];
          dN[a][1] = 0.25 * s[1] * (1 + s[0] * X[0]) * 2.0 / h[1];
        } else {
          dN[a][0] = 0.125 * s[0] * (1 + s[1] * X[1]) * (1 + s[2] * X[2]) * 2.0 / h[0];
          dN[a][1] = 0.125 * s[1] * (1 + s[0] * X[0]) * (1 + s[2] * X[2]) * 2.0 / h[1];
          dN[a][2] = 0.125 * s[2] * (1 + s[0] * X[0]) * (1 + s[1] * X[1]) * 2.0 / h[2];
        }
      }
      double w = W[(size_t)q * g.E + e];
      for (int a = 0; a < NN; a++)
        for (int b = 0; b < NN; b++) {
          double dot = 0.0;
          for (int kk = 0; kk < dim; kk++) dot += dN[a][kk] * dN[b][kk];
          for (int ii = 0; ii < dim; ii++)
            for (int jj = 0; jj < dim; jj++) {
              double v = w * ((ii == jj ? dot : 0.0) + dN[a][jj] * dN[b][ii]);
              K[(size_t)(nodes[a] * dim + ii) * n + nodes[b] * dim + jj] += v;
            }
        }
    }
  }
  for (long nd = 0; nd < g.N; nd++)
    for (int d = 0; d < dim; d++)
      if ((bc[nd] >> d) & 1) {
        long r = nd * dim + d;
        for (int c = 0; c < n; c++) {
          K[(size_t)r * n + c] = 0.0;
          K[(size_t)c * n + r] = 0.0;
        }
        K[(size_t)r * n + r] = 1.0;
      }
  // Cholesky K = L L^T, then inverse
  std::vector<double> Lm(K);
  for (int c = 0; c < n; c++) {
    double s = Lm[(size_t)c * n + c];
    for (int k2 = 0; k2 < c; k2++) s -= Lm[(size_t)c * n + k2] * Lm[(size_t)c * n + k2];
    if (!(s > 0.0)) throw std::runtime_error("coarse operator is not positive definite (check viscosity/BCs)");
    double dgl = sqrt(s);
    Lm[(size_t)c * n + c] = dgl;
    for (int r = c + 1; r < n; r++) {
      double t = Lm[(size_t)r * n + c];
      for (int k2 = 0; k2 < c; k2++) t -= Lm[(size_t)r * n + k2] * Lm[(size_t)c * n + k2];
      Lm[(size_t)r * n + c] = t / dgl;
    }
  }
  std::vector<double> Inv((size_t)n * n, 0.0), col(n);
  for (int c = 0; c < n; c++) {
    // solve L y = e_c ; L^T x = y
    for (int r = 0; r < n; r++) {
      double t = (r == c) ? 1.0 : 0.0;
      for (int k2 = 0; k2 < r; k2++) t -= Lm[(size_t)r * n + k2] * col[k2];
      col[r] = t / Lm[(size_t)r * n + r];
    }
    for (int r = n - 1; r >= 0; r--) {
      double t = col[r];
      for (int k2 = r + 1; k2 < n; k2++) t -= Lm[(size_t)k2 * n + r] * col[k2];
      col[r] = t / Lm[(size_t)r * n + r];
    }
    for (int r = 0; r < n; r++) Inv[(size_t)c * n + r] = col[r];
  }
  st->coarse.reserve((long)n * n);
  UWG_HIP(hipMemcpy(st->coarse.Ainv, Inv.data(), sizeof(double) * n * n, hipMemcpyHostToDevice));
}

// power iteration for lambda_max(D^-1 K) on each level
static void estimate_lmax(uwg_stokes* st, Level& A) {
  // 12 power iterations x <- D^-1 K x / |D^-1 K x| from a fixed pseudo-random start; the norm
  // stays on the device (one host read per level): after the first normalisation |x| = 1, so
  // lambda = |D^-1 K x| of the last iteration
  long n = A.n();
  launch_fill_hash(n, A.x, 0);
  const double* ptr[1] = {A.d};
  UWG_HIP(hipMemcpy(st->dptrs, ptr, sizeof(double*), hipMemcpyHostToDevice));
  for (int it = 0; it < 12; it++) {
    launch_apply_K<double>(A.g, A.W, A.x, A.r, A.bc, 0);
    vk_xmy_scale<double>(n, A.dinv, A.r, A.d, 0);
    vk_multi_dot(n, 1, A.d, st->dptrs, st->red_partial, st->red_out, 0);
    if (it < 11) vk_scale_inv_sqrt(n, st->red_out, A.d, A.x, 0);
  }
  double nd;
  UWG_HIP(hipMemcpy(&nd, st->red_out, sizeof(double), hipMemcpyDeviceToHost));
  A.lmax = sqrt(nd);
}

// Log-viscosity per node of a level: element mean viscosity sum_g W_g / volume (W reproduces
// the particle moments, so sum_g W_g = sum_p eta_p w_p detJ), log-averaged over the elements
// around each node; returned as log10(eta / eta_ref) with eta_ref the level's geometric mean.
static void level_feature(uwg_stokes* st, int level, std::vector<float>& out, double* eta_ref) {
  Level& A = st->L[level];
  const Geom& g = A.g;
  const int Q = g.dim == 3 ? 27 : 9;
  std::vector<double> W((size_t)Q * g.E);
  copy_W_to_host(A, W.data());
  double vol = 1.0;
  for (int d = 0; d < g.dim; d++) vol *= (g.hi[d] - g.lo[d]) / g.res[d];
  std::vector<double> le(g.E);
  for (long e = 0; e < g.E; e++) {
    double s = 0.0;
    for (int q = 0; q < Q; q++) s += W[(size_t)q * g.E + e];
    le[e] = log(std::max(s / vol, 1e-300));
  }
  std::vector<double> f(g.N, 0.0);
  for (long n = 0; n < g.N; n++) {
    long i = n % g.nn[0], j = (n / g.nn[0]) % g.nn[1], k = n / ((long)g.nn[0] * g.nn[1]);
    double acc = 0.0;
    int cnt = 0;
    for (int oz = (g.dim == 3 ? -1 : 0); oz <= 0; oz++)
      for (int oy = -1; oy <= 0; oy++)
        for (int ox = -1; ox <= 0; ox++) {
          long ei = i + ox, ej = j + oy, ek = k + oz;
          if (ei < 0 || ei >= g.res[0] || ej < 0 || ej >= g.res[1] || ek < 0 || ek >= g.res[2]) continue;
          acc += le[ei + g.res[0] * (ej + (long)g.res[1] * ek)];
          cnt++;
        }
    f[n] = acc / cnt;
  }
  double mean = 0.0;
  for (double v : f) mean += v;
  mean /= g.N;
  out.resize(g.N);
  for (long n = 0; n < g.N; n++) out[n] = (float)((f[n] - mean) / log(10.0));
  *eta_ref = exp(mean);
}

static void setup_hierarchy(uwg_stokes* st) {
  for (size_t l = 1; l < st->L.size(); l++) {
    Level& F = st->L[l - 1];
    Level& C = st->L[l];
    launch_restrict_W(F.g, C.g, F.W, C.W, 0);
    launch_bc_mask_from_walls(C.g, C.bc, 0, 0);
  }
  // coarse Dirichlet masks: a coarse dof is constrained when the coincident fine dof is
  // (nested walls); derived from level 0's mask by injection
  for (size_t l = 1; l < st->L.size(); l++) launch_inject_bc(st->L[l - 1].g, st->L[l].g, st->L[l - 1].bc, st->L[l].bc, 0);
  for (auto& A : st->L) launch_diag(A.g, A.W, A.bc, A.dinv, 0);
  for (auto& A : st->L) estimate_lmax(st, A);
  build_coarse(st);
  st->fp32_ready = false;
  if (st->nn && st->nn_level < (int)st->L.size()) {
    std::vector<float> f;
    double eta_ref;
    level_feature(st, st->nn_level, f, &eta_ref);
    nn_set_feature(st->nn.get(), f.data(), eta_ref);
  }
}

// ---------------------------------------------------------------------------------------
// multigrid V-cycle (templated on the arithmetic type of the cycle)
// ---------------------------------------------------------------------------------------
namespace {
template <typename T> struct LV;
template <> struct LV<double> {
  static double* W(Level& A) { return A.W; }
  static const uint32_t* Wbf(Level&) { return nullptr; }
  static double* dinv(Level& A) { return A.dinv; }
  static double* x(Level& A) { return A.x; }
  static double* b(Level& A) { return A.b; }
  static double* r(Level& A) { return A.r; }
  static double* d(Level& A) { return A.d; }
};
template <> struct LV<float> {
  static float* W(Level& A) { return A.W32; }
  static const uint32_t* Wbf(Level& A) { return A.Wbf; }
  static float* dinv(Level& A) { return A.dinv32; }
  static float* x(Level& A) { return A.x32; }
  static float* b(Level& A) { return A.b32; }
  static float* r(Level& A) { return A.r32; }
  static float* d(Level& A) { return A.d32; }
};
}  // namespace

static int g_smooth_degree = 2;
static int g_verbose = 0;

template <typename T>
// Chebyshev(degree) smoothing of K x = b with Jacobi scaling, eigenvalue window
// [0.12, 1.1] * lmax.  Each step after the first is one fused kernel (apply, residual and
// update); the iterate alternates between x and the level's r buffer, arranged so that the
// last step writes x.
static void chebyshev(Level& A, const T* b, T* x, int degree, bool zero_guess) {
  const long n = A.n();
  const double emax = 1.1 * A.lmax, emin = 0.12 * A.lmax;
  const double theta = 0.5 * (emax + emin), delta = 0.5 * (emax - emin);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  T* tmp = LV<T>::r(A);
  T* d = LV<T>::d(A);
  const T* dinv = LV<T>::dinv(A);
  const int steps = zero_guess ? degree - 1 : degree;  // fused steps
  T* cur = x;
  if (zero_guess) {
    // x0 = 0: d = dinv b / theta, x1 = d; written where the alternation ends in x
    cur = (steps % 2 == 1) ? tmp : x;
    vk_cheb_first<T>(n, (T)(1.0 / theta), dinv, b, d, cur, 0);
  }
  for (int k = 0; k < steps; k++) {
    T* nxt = (cur == x) ? tmp : x;
    if (!zero_guess && k == 0) {
      launch_cheb_step_K<T>(A.g, LV<T>::W(A), cur, nxt, b, dinv, d, (T)0, (T)(1.0 / theta), true, A.bc, 0,
                            LV<T>::Wbf(A));
    } else {
      const double rho_new = 1.0 / (2.0 * sigma - rho);
      launch_cheb_step_K<T>(A.g, LV<T>::W(A), cur, nxt, b, dinv, d, (T)(rho_new * rho), (T)(2.0 * rho_new / delta),
                            false, A.bc, 0, LV<T>::Wbf(A));
      rho = rho_new;
    }
    cur = nxt;
  }
  if (cur != x) vk_copy<T>(n, cur, x, 0);
}

template <typename T>
static void coarse_solve(uwg_stokes* st, Level& A, const T* b, T* x);

template <>
void coarse_solve<double>(uwg_stokes* st, Level& A, const double* b, double* x) {
  if (st->coarse.Ainv)
    launch_dense_mv(st->coarse.n, st->coarse.Ainv, b, x, 0);
  else
    chebyshev<double>(A, b, x, 24, true);
}
template <>
void coarse_solve<float>(uwg_stokes* st, Level& A, const float* b, float* x) {
  // FP64 coarsest solve inside the FP32 cycle
  const long n = A.n();
  vk_f2d(n, 1.0, b, A.b, 0);
  coarse_solve<double>(st, A, A.b, A.x);
  vk_d2f(n, 1.0, A.x, x, 0);
}

template <typename T>
static void capture_vector(uwg_stokes* st, long n, const T* v) {
  float* dst = st->capture_buf + (long)st->captured * n;
  if constexpr (std::is_same<T, float>::value)
    vk_copy<float>(n, v, dst, 0);
  else
    vk_d2f(n, 1.0, v, dst, 0);
  st->captured++;
}

template <typename T>
static void vcycle(uwg_stokes* st, int l, const T* b, T* x) {
  Level& A = st->L[l];
  if (l == (int)st->L.size() - 1) {
    coarse_solve<T>(st, A, b, x);
    return;
  }
  Level& C = st->L[l + 1];
  chebyshev<T>(A, b, x, g_smooth_degree, true);
  T* r = LV<T>::r(A);
  launch_residual_K<T>(A.g, LV<T>::W(A), x, b, r, A.bc, 0, LV<T>::Wbf(A));
  launch_restrict2<T>(A.g, C.g, r, LV<T>::b(C), C.bc, A.bc, 0);
  const bool capture = st->capture_level == l + 1 && st->captured < st->capture_max &&
                       (st->capture_seen++ % st->capture_every) == 0;
  if (capture && st->capture_kind == 0) capture_vector<T>(st, C.n(), LV<T>::b(C));
  bool use_nn = false;
  if constexpr (std::is_same<T, float>::value)
    use_nn = st->nn && !st->nn_disabled && l + 1 == st->nn_level;
  if (use_nn && st->nn_mode == 0) {
    // the learned network stands in for the coarse-grid cycle below this level; its output is
    // scaled by the residual-minimising step on this level
    if constexpr (std::is_same<T, float>::value) {
      nn_coarse_correction(st->nn.get(), C.g, C.b32, C.d32, C.bc, 0);
      launch_apply_K<float>(C.g, C.W32, C.d32, C.r32, C.bc, 0, C.Wbf);
      nn_line_search(C.n(), C.b32, C.r32, C.d32, C.x32, false, 0);
    }
  } else {
    vcycle<T>(st, l + 1, LV<T>::b(C), LV<T>::x(C));
    if (capture && st->capture_kind == 1) {
      // residual the coarse cycle leaves behind: its error K_C^{-1} r is what an additive
      // correction has to supply
      launch_residual_K<T>(C.g, LV<T>::W(C), LV<T>::x(C), LV<T>::b(C), LV<T>::r(C), C.bc, 0, LV<T>::Wbf(C));
      capture_vector<T>(st, C.n(), LV<T>::r(C));
    }
    if constexpr (std::is_same<T, float>::value) {
      if (use_nn && st->nn_mode == 1) {
        // learned correction of the coarse cycle's remaining error: x_C += NN(b_C - K_C x_C)
        // (b_C is not needed after the cycle: it receives K_C d for the step length)
        launch_residual_K<float>(C.g, C.W32, C.x32, C.b32, C.r32, C.bc, 0, C.Wbf);
        nn_coarse_correction(st->nn.get(), C.g, C.r32, C.d32, C.bc, 0);
        launch_apply_K<float>(C.g, C.W32, C.d32, C.b32, C.bc, 0, C.Wbf);
        nn_line_search(C.n(), C.r32, C.b32, C.d32, C.x32, true, 0);
      }
    }
  }
  launch_prolong_add<T>(A.g, C.g, LV<T>::x(C), x, A.bc, 0);
  chebyshev<T>(A, b, x, g_smooth_degree, false);
}

// ---------------------------------------------------------------------------------------
// Operators of the finest level (Q1/dQ0 or Q2/dPc1) and its multigrid cycle
// ---------------------------------------------------------------------------------------
namespace {
template <typename T> struct Q2V;
template <> struct Q2V<double> {
  static const double* D(Q2Level& A) { return A.D; }
  static const double* W(Q2Level& A) { return A.W; }
  static const double* dinv(Q2Level& A) { return A.dinv; }
  static const double* Sinv(Q2Level& A) { return A.Sinv; }
  static double* x(Q2Level& A) { return A.x; }
  static double* b(Q2Level& A) { return A.b; }
  static double* r(Q2Level& A) { return A.r; }
  static double* d(Q2Level& A) { return A.d; }
};
template <> struct Q2V<float> {
  static const float* D(Q2Level& A) { return A.D32; }
  static const float* W(Q2Level& A) { return A.W32; }
  static const float* dinv(Q2Level& A) { return A.dinv32; }
  static const float* Sinv(Q2Level& A) { return A.Sinv32; }
  static float* x(Q2Level& A) { return A.x32; }
  static float* b(Q2Level& A) { return A.b32; }
  static float* r(Q2Level& A) { return A.r32; }
  static float* d(Q2Level& A) { return A.d32; }
};
}  // namespace

// K on the finest level; bc selects the Dirichlet rows/columns (identity rows)
template <typename T>
static void top_K(uwg_stokes* st, const T* x, T* y, const uint8_t* bc) {
  if (st->q2)
    launch_q2_apply<T>(st->g, Q2V<T>::W(*st->q2), x, y, bc, 0, Q2V<T>::D(*st->q2));
  else
    launch_apply_K<T>(st->g, LV<T>::W(st->L[0]), x, y, bc, 0);
}
template <typename T>
static void top_G(uwg_stokes* st, const T* p, T* y, const uint8_t* bc) {
  if (st->q2)
    launch_q2_apply_G<T>(st->g, st->q2->Gr, p, bc, y, 0);
  else
    launch_apply_G<T>(st->g, st->Ge, p, bc, y, 0);
}
template <typename T>
static void top_GT(uwg_stokes* st, const T* u, T* q, const uint8_t* bc) {
  if (st->q2)
    launch_q2_apply_GT<T>(st->g, st->q2->Gr, u, bc, q, 0);
  else
    launch_apply_GT<T>(st->g, st->Ge, u, bc, q, 0);
}
static const uint8_t* top_bc(uwg_stokes* st) { return st->q2 ? st->q2->bc : st->L[0].bc; }
static const Geom& top_vgeom(uwg_stokes* st) { return st->q2 ? st->q2->gn : st->g; }
// Schur preconditioner z = -approxS^-1 r: diagonal (dQ0, diagS = -1/approxS) or element
// blocks (dPc1)
template <typename T>
static void top_schur(uwg_stokes* st, const T* diagS, const T* r, T* z) {
  if (st->q2)
    launch_q2_schur_prec<T>(st->g.E, st->g.dim + 1, Q2V<T>::Sinv(*st->q2), r, z, 0);
  else
    vk_xmy_scale<T>(st->np, diagS, r, z, 0);
}

// Schur block of the coupled preconditioner: the learned stencil while it is enabled for this
// solve (its coefficients follow the assembly), else the diagonal -approxS^-1
static bool g_snn_active = false;
static void ensure_snn_coefficients(uwg_stokes* st) {
  if (st->snn_stamp == st->assembly_stamp && st->snn_a) return;
  if (!st->snn_a) st->snn_a = dalloc<uint16_t>(27 * st->g.E);
  schur_net_coefficients(st->snn.get(), st->g, st->approxS, st->snn_a, 0);
  st->snn_stamp = st->assembly_stamp;
}
template <typename T>
static void schur_prec(uwg_stokes* st, const T* diagS, const T* r, T* z) {
  if (g_snn_active && st->snn && !st->nn_disabled)
    schur_stencil_apply<T>(st->g, st->snn_a, r, z, 0);
  else
    top_schur<T>(st, diagS, r, z);
}

// Chebyshev smoothing on the Q2 level: same window and recurrence as the Q1 smoother
template <typename T>
static void q2_chebyshev(uwg_stokes* st, const T* b, T* x, int degree, bool zero_guess) {
  Q2Level& A = *st->q2;
  const long n = A.n();
  const double emax = 1.1 * A.lmax, emin = 0.12 * A.lmax;
  const double theta = 0.5 * (emax + emin), delta = 0.5 * (emax - emin);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  T* r = Q2V<T>::r(A);
  T* d = Q2V<T>::d(A);
  const T* dinv = Q2V<T>::dinv(A);
  int k0 = 0;
  if (zero_guess) {
    vk_cheb_first<T>(n, (T)(1.0 / theta), dinv, b, d, x, 0);
    k0 = 1;
  }
  for (int k = k0; k < degree; k++) {
    top_K<T>(st, x, r, A.bc);
    vk_residual<T>(n, b, r, 0);
    if (!zero_guess && k == 0) {
      vk_cheb_update<T>(n, (T)0, (T)(1.0 / theta), dinv, r, d, x, 0);
    } else {
      const double rho_new = 1.0 / (2.0 * sigma - rho);
      vk_cheb_update<T>(n, (T)(rho_new * rho), (T)(2.0 * rho_new / delta), dinv, r, d, x, 0);
      rho = rho_new;
    }
  }
}

// V-cycle of the Q2 level: smooth, restrict the residual to the embedded Q1 space (the Q2
// vertices are the vertices of the once-refined Q1 grid, and the Q1 -> Q2 interpolation is the
// same averaging as Q1 grid refinement), the Q1 cycle, prolong, smooth
template <typename T>
static void q2_vcycle(uwg_stokes* st, const T* b, T* x) {
  Q2Level& A = *st->q2;
  Level& C = st->L[0];
  q2_chebyshev<T>(st, b, x, g_smooth_degree, true);
  T* r = Q2V<T>::r(A);
  top_K<T>(st, x, r, A.bc);
  vk_residual<T>(A.n(), b, r, 0);
  launch_restrict2<T>(A.gn, C.g, r, LV<T>::b(C), C.bc, A.bc, 0);
  vcycle<T>(st, 0, LV<T>::b(C), LV<T>::x(C));
  launch_prolong_add<T>(A.gn, C.g, LV<T>::x(C), x, A.bc, 0);
  q2_chebyshev<T>(st, b, x, g_smooth_degree, false);
}

template <typename T>
static void top_vcycle(uwg_stokes* st, const T* b, T* x) {
  if (st->q2)
    q2_vcycle<T>(st, b, x);
  else
    vcycle<T>(st, 0, b, x);
}
template <typename T>
static T* top_buf_b(uwg_stokes* st) {
  return st->q2 ? Q2V<T>::b(*st->q2) : LV<T>::b(st->L[0]);
}
template <typename T>
static T* top_buf_x(uwg_stokes* st) {
  return st->q2 ? Q2V<T>::x(*st->q2) : LV<T>::x(st->L[0]);
}

// ---------------------------------------------------------------------------------------
// FGMRES (right preconditioned, classical Gram-Schmidt), FP64 vectors on the device.
// Converged when |r| <= max(rtol |r0|, atol) with r0 the initial residual
// (KSPConvergedDefault with zero initial guess).
// ---------------------------------------------------------------------------------------
// The Arnoldi basis V may be stored in a narrower type TV than the preconditioned vectors Z,
// the work vector w and the iterate (T).
template <typename T, typename TV>
struct uwg::KSpaceT {
  long n = 0;
  int m = 0;
  std::vector<TV*> V;
  std::vector<T*> Z;
  T* w = nullptr;
  const TV** dV = nullptr;
  void init(long n_, int m_) {
    if (n_ == n && m_ == m) return;
    free_all();
    n = n_;
    m = m_;
    V.resize(m + 1);
    Z.resize(m);
    for (auto& v : V) v = dalloc<TV>(n);
    for (auto& z : Z) z = dalloc<T>(n);
    w = dalloc<T>(n);
    UWG_HIP(hipMalloc(&dV, sizeof(TV*) * (m + 1)));
    UWG_HIP(hipMemcpy(dV, V.data(), sizeof(TV*) * (m + 1), hipMemcpyHostToDevice));
  }
  void free_all() {
    for (auto& v : V) dfree(v);
    for (auto& z : Z) dfree(z);
    dfree(w);
    dfree(dV);
    V.clear();
    Z.clear();
    w = nullptr;
    dV = nullptr;
  }
  void release() {
    free_all();
    n = 0;
    m = 0;
  }
  ~KSpaceT() { free_all(); }
};
struct uwg::KrylovSpace : uwg::KSpaceT<double> {};

using OpFn = std::function<void(const double*, double*)>;
using OpFn32 = std::function<void(const float*, float*)>;

template <typename T>
static double dev_dot(uwg_stokes* st, long n, const T* a, const T* b) {
  const T* ptr[1] = {b};
  UWG_HIP(hipMemcpy(st->dptrs, ptr, sizeof(T*), hipMemcpyHostToDevice));
  vk_multi_dot<T>(n, 1, a, (const T* const*)st->dptrs, st->red_partial, st->red_out, 0);
  double r;
  UWG_HIP(hipMemcpy(&r, st->red_out, sizeof(double), hipMemcpyDeviceToHost));
  return r;
}

// Right-preconditioned FGMRES with classical Gram-Schmidt; vectors of type T, the Arnoldi
// basis of type TV (T, or float under an FP64 iteration), every reduction and the Hessenberg
// least-squares problem in double.  The preconditioner reads a basis vector.
template <typename T, typename TV = T>
static int fgmres(uwg_stokes* st, KSpaceT<T, TV>& K, const std::function<void(const T*, T*)>& A,
                  const std::function<void(const TV*, T*)>& M, const T* b, T* x, double rtol, double atol, int maxit,
                  int* its_out, double* rnorm_out, bool x0_zero = true,
                  const std::function<void(int, double)>* monitor = nullptr) {
  const long n = K.n;
  const int m = K.m;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hcol(m + 1);
  int its = 0;
  double r0 = -1.0, rn = 0.0;
  bool conv = false;
  while (true) {
    // r = b - A x (formed in the basis' first vector, or in w when the basis is narrower)
    T* r = nullptr;
    if constexpr (std::is_same<T, TV>::value)
      r = K.V[0];
    else
      r = K.w;
    if (x0_zero && its == 0) {
      vk_copy<T>(n, b, r, 0);
    } else {
      A(x, r);
      vk_axpby<T>(n, (T)1, b, (T)-1, r, 0);
    }
    double beta = sqrt(std::max(0.0, dev_dot<T>(st, n, r, r)));
    if (r0 < 0) r0 = beta;
    rn = beta;
    if (beta <= std::max(rtol * r0, atol) || its >= maxit) {
      conv = beta <= std::max(rtol * r0, atol);
      break;
    }
    if constexpr (std::is_same<T, TV>::value)
      vk_scale<T>(n, (T)(1.0 / beta), K.V[0], K.V[0], 0);
    else
      vk_d2f(n, 1.0 / beta, r, K.V[0], 0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m && its < maxit; j++) {
      M(K.V[j], K.Z[j]);
      A(K.Z[j], K.w);
      // Gram-Schmidt on the device; one copy of the Hessenberg column per iteration
      vk_gs_step<T, TV>(n, j + 1, K.dV, K.w, K.V[j + 1], st->dh, st->red_partial, 0);
      UWG_HIP(hipMemcpy(hcol.data(), st->dh, sizeof(double) * (j + 2), hipMemcpyDeviceToHost));
      double hn = sqrt(std::max(0.0, hcol[j + 1]));
      hcol[j + 1] = hn;
      // Givens
      for (int i = 0; i < j; i++) {
        double t = cs[i] * hcol[i] + sn[i] * hcol[i + 1];
        hcol[i + 1] = -sn[i] * hcol[i] + cs[i] * hcol[i + 1];
        hcol[i] = t;
      }
      double den = sqrt(hcol[j] * hcol[j] + hcol[j + 1] * hcol[j + 1]);
      cs[j] = den > 0 ? hcol[j] / den : 1.0;
      sn[j] = den > 0 ? hcol[j + 1] / den : 0.0;
      hcol[j] = den;
      hcol[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = hcol[i];
      its++;
      rn = fabs(g[j + 1]);
      if (monitor) (*monitor)(its, rn / r0);
      if (g_verbose) fprintf(stderr, "  fgmres<%s> n=%ld it %d |r| %.3e rel %.3e\n", sizeof(T) == 4 ? "f32" : "f64", n,
                             its, rn, rn / r0);
      if (rn <= std::max(rtol * r0, atol) || hn == 0.0) {
        j++;
        break;
      }
    }
    // back substitution and update x += Z y
    int k = j;
    for (int i = k - 1; i >= 0; i--) {
      double t = g[i];
      for (int c = i + 1; c < k; c++) t -= H[(size_t)i * m + c] * y[c];
      y[i] = t / H[(size_t)i * m + i];
    }
    UWG_HIP(hipMemcpy(st->dcoef, y.data(), sizeof(double) * k, hipMemcpyHostToDevice));
    std::vector<const T*> zp(K.Z.begin(), K.Z.begin() + k);
    UWG_HIP(hipMemcpy(st->dptrs, zp.data(), sizeof(T*) * k, hipMemcpyHostToDevice));
    vk_multi_axpy<T>(n, k, (const T* const*)st->dptrs, st->dcoef, x, 0);
    x0_zero = false;
    // with a narrower basis the estimate can run ahead of the true residual by the basis'
    // rounding: the loop head then decides on the true residual (one more apply of A)
    if (rn <= std::max(rtol * r0, atol) && std::is_same<T, TV>::value) {
      conv = true;
      break;
    }
    if (its >= maxit && std::is_same<T, TV>::value) break;
  }
  if (its_out) *its_out = its;
  if (rnorm_out) *rnorm_out = rn;
  return conv ? 0 : 1;
}

// ---------------------------------------------------------------------------------------
// Stokes handle and solves
// ---------------------------------------------------------------------------------------
extern "C" int uwg_stokes_create(uwg_field* vel, uwg_field* p, const unsigned char* bcmask, uwg_stokes** out) {
  if (vel->on_elements || vel->ncomp != vel->mesh->g.dim) return arg_error("velocityField must be a nodal vector field");
  const bool q2 = vel->mesh->order == 2;
  if (p->ncomp != 1 || p->mesh != vel->mesh || p->on_elements != (q2 ? 2 : 1))
    return arg_error(q2 ? "pressureField must be a scalar field on the dPc1 submesh of the Q2 mesh"
                        : "pressureField must be a scalar field on the dQ0 submesh");
  API_BEGIN
  uwg_stokes* st = new uwg_stokes();
  st->vel = vel;
  st->p = p;
  st->g = vel->mesh->g;
  const Geom& g = st->g;
  const int walls = g.dim == 3 ? 0x3f : 0x0f;
  st->bc0 = dalloc<uint8_t>(g.N);
  if (q2) {
    st->q2 = new Q2Level();
    Q2Level& A = *st->q2;
    A.gn = vel->mesh->gn;
    A.bc = dalloc<uint8_t>(A.gn.N);
    if (bcmask)
      UWG_HIP(hipMemcpy(A.bc, bcmask, A.gn.N, hipMemcpyHostToDevice));
    else
      launch_bc_mask_from_walls(A.gn, A.bc, walls, 0);
    // the embedded Q1 space: a Q1 vertex is the Q2 vertex (2i, 2j, 2k)
    launch_inject_bc(A.gn, g, A.bc, st->bc0, 0);
    const int NQ = g.dim == 3 ? 125 : 25, NEN = g.dim == 3 ? 27 : 9, NP = g.dim + 1;
    const long n = A.n();
    A.W = dalloc<double>(g.E * NQ);
    A.dinv = dalloc<double>(n);
    for (double** v : {&A.x, &A.b, &A.r, &A.d}) {
      *v = dalloc<double>(n);
      vk_fill<double>(n, 0.0, *v, 0);
    }
    std::vector<double> Gr;
    q2_element_G(g, Gr);
    A.Gr = dalloc<double>((long)Gr.size());
    UWG_HIP(hipMemcpy(A.Gr, Gr.data(), sizeof(double) * Gr.size(), hipMemcpyHostToDevice));
    A.S = dalloc<double>(g.E * NP * NP);
    A.Sinv = dalloc<double>(g.E * NP * NP);
    A.fe = dalloc<double>(g.E * NEN * g.dim);
    if (g.dim == 3) {
      std::vector<double> D;
      q2_mfma_matrix(g, D);
      std::vector<float> D32(D.begin(), D.end());
      A.D = dalloc<double>((long)D.size());
      A.D32 = dalloc<float>((long)D.size());
      UWG_HIP(hipMemcpy(A.D, D.data(), sizeof(double) * D.size(), hipMemcpyHostToDevice));
      UWG_HIP(hipMemcpy(A.D32, D32.data(), sizeof(float) * D32.size(), hipMemcpyHostToDevice));
    }
    st->nu = n;
    st->np = g.E * NP;
  } else {
    if (bcmask)
      UWG_HIP(hipMemcpy(st->bc0, bcmask, g.N, hipMemcpyHostToDevice));
    else
      launch_bc_mask_from_walls(g, st->bc0, walls, 0);
    st->nu = g.N * g.dim;
    st->np = g.E;
  }
  build_levels(st, 0);
  compute_Ge(g, st->Ge);
  const int NN = 1 << g.dim;
  st->fe = dalloc<double>(g.E * NN * g.dim);
  st->f = dalloc<double>(st->nu);
  st->ubc = dalloc<double>(st->nu);
  st->approxS = dalloc<double>(g.E);
  st->h = dalloc<double>(st->np);
  st->red_partial = dalloc<double>(32 * vk_reduce_blocks());
  st->red_out = dalloc<double>(8);
  st->dcoef = dalloc<double>(256);
  st->dh = dalloc<double>(256);
  UWG_HIP(hipMalloc(&st->dptrs, sizeof(double*) * 256));
  *out = st;
  return 0;
  API_END
}

extern "C" int uwg_stokes_destroy(uwg_stokes* st) {
  if (!st) return 0;
  for (auto& A : st->L)
    for (void* p : {(void*)A.W, (void*)A.W32, (void*)A.Wbf, (void*)A.bc, (void*)A.dinv, (void*)A.dinv32, (void*)A.x,
                    (void*)A.b, (void*)A.r, (void*)A.d, (void*)A.x32, (void*)A.b32, (void*)A.r32, (void*)A.d32})
      if (p != (void*)st->bc0) dfree(p);
  dfree(st->bc0);
  if (st->q2) {
    Q2Level& A = *st->q2;
    for (void* p : {(void*)A.bc, (void*)A.W, (void*)A.W32, (void*)A.dinv, (void*)A.dinv32, (void*)A.x, (void*)A.b,
                    (void*)A.r, (void*)A.d, (void*)A.x32, (void*)A.b32, (void*)A.r32, (void*)A.d32, (void*)A.Gr,
                    (void*)A.S, (void*)A.Sinv, (void*)A.Sinv32, (void*)A.fe, (void*)A.D, (void*)A.D32})
      dfree(p);
    delete st->q2;
  }
  delete st->kin;
  delete st->kout;
  delete st->kb;
  delete st->kf;
  delete st->kb32;
  dfree(st->capture_buf);
  dfree(st->snn_a);
  for (int k = 0; k < WS_COUNT; k++) dfree(st->ws_ptr[k]);
  for (void* p : {(void*)st->fe, (void*)st->f, (void*)st->ubc, (void*)st->approxS, (void*)st->h, (void*)st->eta_tab,
                  (void*)st->force_tab, (void*)st->eta_p, (void*)st->force_p, (void*)st->red_partial,
                  (void*)st->red_out, (void*)st->dcoef, (void*)st->dh, (void*)st->dptrs, (void*)st->coarse.Ainv})
    dfree(p);
  delete st;
  return 0;
}

extern "C" int uwg_stokes_set_materials(uwg_stokes* st, int nmat, const double* eta, const double* force) {
  if (nmat < 1) return arg_error("nmat must be >= 1");
  for (int m = 0; m < nmat; m++)
    if (!(eta[m] > 0.0) || !std::isfinite(eta[m])) return arg_error("viscosity must be positive and finite");
  API_BEGIN
  dfree(st->eta_tab);
  dfree(st->force_tab);
  st->eta_tab = dalloc<double>(nmat);
  st->force_tab = dalloc<double>(nmat * st->g.dim);
  UWG_HIP(hipMemcpy(st->eta_tab, eta, sizeof(double) * nmat, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(st->force_tab, force, sizeof(double) * nmat * st->g.dim, hipMemcpyHostToDevice));
  st->nmat = nmat;
  st->mode = 0;
  st->assembled_state = -1;
  return 0;
  API_END
}

extern "C" int uwg_stokes_set_particle_values(uwg_stokes* st, long P, const double* eta, const double* force) {
  API_BEGIN
  dfree(st->eta_p);
  dfree(st->force_p);
  st->eta_p = dalloc<double>(P);
  st->force_p = dalloc<double>(P * st->g.dim);
  UWG_HIP(hipMemcpy(st->eta_p, eta, sizeof(double) * P, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(st->force_p, force, sizeof(double) * P * st->g.dim, hipMemcpyHostToDevice));
  st->np_vals = P;
  st->mode = 1;
  st->assembled_state = -1;
  return 0;
  API_END
}

extern "C" int uwg_swarm_set_var(uwg_swarm* s, int slot, int ncomp, const double* data) {
  if (slot < 0 || slot >= 8) return arg_error("swarm variable slot must be 0..7");
  if (ncomp < 1 || ncomp > 3) return arg_error("swarm variables of 1 to 3 components");
  API_BEGIN
  if (s->var_rows[slot] != s->P || s->var_nc[slot] != ncomp) {
    dfree(s->var[slot]);
    s->var[slot] = dalloc<double>(s->P * ncomp);
    s->var_rows[slot] = s->P;
    s->var_nc[slot] = ncomp;
  }
  UWG_HIP(hipMemcpy(s->var[slot], data, sizeof(double) * s->P * ncomp, hipMemcpyHostToDevice));
  return 0;
  API_END
}

// Per-particle viscosity and body force from a compiled function program (fn_kernels.hip):
// prog[nins][5] = (op, dst, a, b, n).  Checked here so that every index the kernel follows
// stays inside its arrays.
extern "C" int uwg_stokes_eval_particle_functions(uwg_stokes* st, uwg_swarm* s, int nins, const int* prog, int nconst,
                                                  const double* consts, int nfields, uwg_field* const* fields,
                                                  int out_eta, int out_force) {
  const int dim = st->g.dim;
  if (nins < 1 || nins > 4096) return arg_error("function program must hold 1..4096 instructions");
  if (nfields < 0 || nfields > 8) return arg_error("at most 8 mesh fields per program");
  if (nconst < 0) return arg_error("bad constant count");
  for (int f = 0; f < nfields; f++)
    if (!fields[f] || fields[f]->on_elements || fields[f]->ncomp < 1 || fields[f]->ncomp > 3 ||
        fields[f]->mesh != s->mesh || fields[f]->mesh->order != 1)
      return arg_error("program fields must be nodal fields of 1-3 components on the swarm's (Q1) mesh");
  int nc_reg[48];
  for (int r = 0; r < 48; r++) nc_reg[r] = 0;
  for (int i = 0; i < nins; i++) {
    const int op = prog[5 * i], dst = prog[5 * i + 1], a = prog[5 * i + 2], b = prog[5 * i + 3], n = prog[5 * i + 4];
    auto reg_ok = [&](int r) { return r >= 0 && r < 48 && nc_reg[r] > 0; };
    if (dst < 0 || dst >= 48 || n < 1 || n > 3) return arg_error("function program: bad destination");
    bool ok = true;
    switch (op) {
      case 0: ok = a >= 0 && a + n <= nconst; break;                                   // CONST
      case 1: ok = n == dim; break;                                                    // COORD
      case 2: ok = a >= 0 && a < nfields && fields[a]->ncomp == n; break;              // FIELD
      case 3: ok = a >= 0 && a < 8 && s->var[a] && s->var_rows[a] == s->P && s->var_nc[a] == n; break;
      case 4: ok = s->has_mat && n == 1; break;                                        // MAT
      case 10: case 11: case 12: case 13: case 14: case 15: case 16: case 17: case 18: case 19: case 20:
        ok = reg_ok(a) && reg_ok(b) && (nc_reg[a] == n || nc_reg[a] == 1) && (nc_reg[b] == n || nc_reg[b] == 1);
        break;
      case 30: case 31: case 32: case 33: case 34: case 35: case 36: case 37:
        ok = reg_ok(a) && nc_reg[a] == n;
        break;
      case 40: ok = reg_ok(a) && reg_ok(b) && n == 1 && (nc_reg[b] == nc_reg[a] || nc_reg[b] == 1); break;
      case 41: ok = reg_ok(a) && n == 1 && b >= 0 && b < nc_reg[a]; break;             // INDEX
      case 42: ok = reg_ok(a) && reg_ok(b) && nc_reg[a] + nc_reg[b] == n; break;       // APPEND
      case 51: ok = reg_ok(a) && b >= 0 && b < nconst && n == 1; break;                // EQK
      case 52: ok = reg_ok(a) && reg_ok(b) && nc_reg[dst] == n && nc_reg[a] == n; break; // SELECT
      case 53: ok = reg_ok(a); break;                                                  // CHECK
      case 54: ok = reg_ok(a) && nc_reg[a] == n; break;                                // MOVE
      default: ok = false;
    }
    if (!ok) return arg_error("function program: invalid instruction " + std::to_string(i));
    if (op != 53) nc_reg[dst] = n;
  }
  if (out_eta >= 0 && (out_eta >= 48 || nc_reg[out_eta] != 1)) return arg_error("viscosity must be a scalar");
  if (out_force >= 0 && (out_force >= 48 || (nc_reg[out_force] != dim && nc_reg[out_force] != 1)))
    return arg_error("fn_bodyforce must be a vector of the mesh dimension");
  API_BEGIN
  const long P = s->P;
  if (st->np_vals != P || !st->eta_p) {
    dfree(st->eta_p);
    dfree(st->force_p);
    st->eta_p = dalloc<double>(P);
    st->force_p = dalloc<double>(P * dim);
    st->np_vals = P;
  }
  int* dprog = dalloc<int>(5L * nins);
  double* dconst = dalloc<double>(std::max(nconst, 1));
  int* derr = dalloc<int>(2);
  UWG_HIP(hipMemcpy(dprog, prog, sizeof(int) * 5 * nins, hipMemcpyHostToDevice));
  if (nconst) UWG_HIP(hipMemcpy(dconst, consts, sizeof(double) * nconst, hipMemcpyHostToDevice));
  UWG_HIP(hipMemset(derr, 0, sizeof(int) * 2));
  const double* fptr[8] = {};
  int fnc[8] = {};
  for (int f = 0; f < nfields; f++) {
    fptr[f] = fields[f]->d;
    fnc[f] = fields[f]->ncomp;
  }
  const double* vptr[8];
  int vnc[8];
  for (int v = 0; v < 8; v++) {
    vptr[v] = s->var[v];
    vnc[v] = s->var_nc[v];
  }
  launch_fn_eval(st->g, P, s->coords, s->mat, dprog, nins, dconst, fptr, fnc, nfields, vptr, vnc, 8, out_eta,
                 out_force, st->eta_p, st->force_p, derr, 0);
  int herr[2];
  UWG_HIP(hipMemcpy(herr, derr, sizeof(int) * 2, hipMemcpyDeviceToHost));
  dfree(dprog);
  dfree(dconst);
  dfree(derr);
  if (herr[0] == 1) return domain_error("a mesh field was evaluated outside the mesh");
  if (herr[0] == 2) return arg_error("branching.map / conditional: no mapping or clause for some particles");
  if (herr[0] == 3) return arg_error("function program: unknown operation");
  if (herr[1]) return arg_error("viscosity must be positive everywhere");
  st->mode = 1;
  st->assembled_state = -1;
  return 0;
  API_END
}

extern "C" int uwg_stokes_get_particle_values(uwg_stokes* st, long P, double* eta, double* force) {
  if (st->mode != 1 || !st->eta_p) return arg_error("no per-particle values");
  if (P != st->np_vals) return arg_error("particle count mismatch");
  API_BEGIN
  if (eta) UWG_HIP(hipMemcpy(eta, st->eta_p, sizeof(double) * P, hipMemcpyDeviceToHost));
  if (force) UWG_HIP(hipMemcpy(force, st->force_p, sizeof(double) * P * st->g.dim, hipMemcpyDeviceToHost));
  return 0;
  API_END
}

static void finish_assembly(uwg_stokes* st) {
  Level& A = st->L[0];
  launch_node_gather_fe(st->g, st->fe, A.bc, st->f, 0);
  UWG_HIP(hipMemset(st->h, 0, sizeof(double) * st->g.E));
  setup_hierarchy(st);
  st->assembled = true;
  st->assembly_stamp++;
}

// power iteration for lambda_max(D^-1 K) on the Q2 level
static void q2_estimate_lmax(uwg_stokes* st) {
  Q2Level& A = *st->q2;
  const long n = A.n();
  launch_fill_hash(n, A.x, 0);
  double lam = 0.0;
  for (int it = 0; it < 12; it++) {
    top_K<double>(st, A.x, A.r, A.bc);
    vk_xmy_scale<double>(n, A.dinv, A.r, A.d, 0);
    const double nd = dev_dot<double>(st, n, A.d, A.d);
    const double nx2 = dev_dot<double>(st, n, A.x, A.x);
    lam = sqrt(nd / nx2);
    vk_scale<double>(n, 1.0 / sqrt(nd), A.d, A.x, 0);
  }
  A.lmax = lam;
  vk_fill<double>(n, 0.0, A.d, 0);
}

// Q2/dPc1: body force, approxS blocks, the Q1 weights of the embedded space (the multigrid
// hierarchy below the Q2 level), Jacobi diagonal and spectral bound of the Q2 level
static void finish_assembly_q2(uwg_stokes* st) {
  Q2Level& A = *st->q2;
  const Geom& g = st->g;
  launch_q2_gather_f(g, A.fe, A.bc, st->f, 0);
  vk_fill<double>(st->np, 0.0, st->h, 0);
  int* bad = wsbuf<int>(st, WS_Q2_BAD, 1);
  UWG_HIP(hipMemsetAsync(bad, 0, sizeof(int), 0));
  launch_q2_approxS(g, A.W, A.Gr, A.S, A.Sinv, bad, 0);
  int hb = 0;
  UWG_HIP(hipMemcpy(&hb, bad, sizeof(int), hipMemcpyDeviceToHost));
  if (hb) throw std::runtime_error("singular approxS element block (non-positive viscosity or empty element?)");
  launch_q2_project_W(g, A.W, st->L[0].W, 0);
  setup_hierarchy(st);
  launch_q2_diag(g, A.W, A.bc, A.dinv, 0);
  q2_estimate_lmax(st);
  st->fp32_ready = false;
  st->assembled = true;
}

extern "C" int uwg_stokes_assemble(uwg_stokes* st, uwg_swarm* s) {
  if (s->mesh->g.E != st->g.E) return arg_error("swarm and Stokes system live on different meshes");
  if (st->mode == 0 && st->nmat == 0) return arg_error("no material tables / particle values set");
  if (st->mode == 0 && !s->has_mat && st->nmat > 1) return arg_error("swarm has no material index");
  if (st->mode == 1 && st->np_vals != s->P) return arg_error("particle values do not match the swarm size");
  API_BEGIN
  if (!s->has_integration || s->integration_state != s->geom_state) {
    int rc = uwg_swarm_integration_points(s, 1, 15);
    if (rc) return rc;
  }
  st->swarm = s;
  Level& A = st->L[0];
  if (st->q2) {
    launch_q2_weights(st->g, s->cell_ptr, s->cell_pid, s->xi_slot, s->w_slot, st->mode == 0 ? s->mat : nullptr,
                      st->nmat, st->eta_tab, st->force_tab, st->eta_p, st->force_p, nullptr, nullptr, st->q2->W,
                      st->q2->fe, 0);
    finish_assembly_q2(st);
  } else {
    launch_moments(st->g, s->cell_ptr, s->cell_pid, s->xi_slot, s->w_slot, st->mode == 0 ? s->mat : nullptr,
                   st->eta_tab, st->force_tab, st->eta_p, st->force_p, A.W, st->fe, st->approxS, st->Ge, st->nmat, 0);
    finish_assembly(st);
  }
  st->assembled_state = s->state;
  return 0;
  API_END
}

extern "C" int uwg_stokes_assemble_fields(uwg_stokes* st, const double* eta_g, const double* force_g) {
  if (!eta_g) return arg_error("viscosity values are required");
  API_BEGIN
  // Gauss-swarm path of the reference (no voronoi swarm): function values at the Gauss points of
  // every element (3^dim on Q1 meshes, 5^dim on Q2 meshes, PY/swarm/_integration_swarm.py:173-181),
  // point index qx + n qy + n^2 qz.  When the viscosities equal those of the previous assembly
  // the operators and the multigrid hierarchy are kept and only the body force is re-assembled.
  const Geom& g = st->g;
  const long nq = g.E * (st->q2 ? (g.dim == 3 ? 125 : 25) : (g.dim == 3 ? 27 : 9));
  const bool same_eta = st->assembled && st->assembled_state == -2 && (long)st->eta_g_prev.size() == nq &&
                        memcmp(st->eta_g_prev.data(), eta_g, sizeof(double) * nq) == 0;
  double* de = wsbuf<double>(st, WS_Q2_ETAG, nq);
  double* df = wsbuf<double>(st, WS_Q2_FORCEG, nq * g.dim);
  UWG_HIP(hipMemcpy(de, eta_g, sizeof(double) * nq, hipMemcpyHostToDevice));
  if (force_g)
    UWG_HIP(hipMemcpy(df, force_g, sizeof(double) * nq * g.dim, hipMemcpyHostToDevice));
  else
    vk_fill<double>(nq * g.dim, 0.0, df, 0);
  if (st->q2) {
    launch_q2_weights(g, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr, de, df,
                      st->q2->W, st->q2->fe, 0);
    if (same_eta)
      launch_q2_gather_f(g, st->q2->fe, st->q2->bc, st->f, 0);
    else
      finish_assembly_q2(st);
  } else {
    Level& A = st->L[0];
    launch_gauss_assemble(g, de, df, A.W, st->fe, st->approxS, st->Ge, 0);
    if (same_eta)
      launch_node_gather_fe(g, st->fe, A.bc, st->f, 0);
    else
      finish_assembly(st);
  }
  if (!same_eta) st->eta_g_prev.assign(eta_g, eta_g + nq);
  st->assembled_state = -2;
  return 0;
  API_END
}

extern "C" void uwg_solve_opts_default(uwg_solve_opts* o) {
  memset(o, 0, sizeof(*o));
  o->method = 0;
  o->precision = 0;
  o->learned = 0;
  o->outer_rtol = 1e-5;
  o->inner_rtol = 1e-6;
  o->mg_levels = 0;
  o->max_outer_its = 10000;
  o->max_inner_its = 10000;
  o->restart = 30;
  o->smooth_degree = 2;
  o->verbose = 0;
}

// lift Dirichlet values: rhs_u = f - K ubc (free dofs), rhs_p = h - G^T ubc
static void lifted_rhs(uwg_stokes* st, double* fu, double* hp, bool* has_bc) {
  const long n = st->nu, np = st->np;
  const Geom& gv = top_vgeom(st);
  const uint8_t* bc = top_bc(st);
  // ubc = vel on constrained dofs, 0 elsewhere
  int* flag = wsbuf<int>(st, WS_UBC_FLAG, 1);
  UWG_HIP(hipMemsetAsync(flag, 0, sizeof(int), 0));
  launch_bc_split(gv, bc, st->vel->d, st->ubc, flag, 0);
  int any = 0;
  UWG_HIP(hipMemcpy(&any, flag, sizeof(int), hipMemcpyDeviceToHost));
  *has_bc = any != 0;
  vk_copy<double>(n, st->f, fu, 0);
  vk_copy<double>(np, st->h, hp, 0);
  if (!any) return;
  // K applied to ubc with BC dofs treated as free (an all-free mask), BC rows then zeroed
  uint8_t* free_mask = wsbuf<uint8_t>(st, WS_FREEMASK, gv.N);
  UWG_HIP(hipMemsetAsync(free_mask, 0, gv.N, 0));
  double* t = wsbuf<double>(st, WS_LIFT_T, n);
  top_K<double>(st, st->ubc, t, free_mask);
  launch_zero_bc(gv, bc, t, 0);
  vk_axpy<double>(n, -1.0, t, fu, 0);
  double* q = wsbuf<double>(st, WS_LIFT_Q, np);
  top_GT<double>(st, st->ubc, q, free_mask);
  vk_axpy<double>(np, -1.0, q, hp, 0);
}

static void true_residuals(uwg_stokes* st, const double* u, const double* p, double* out3) {
  // u includes Dirichlet values; residual on free dofs of f - K u - G p and h - G^T u
  const long n = st->nu, np = st->np;
  const Geom& gv = top_vgeom(st);
  const uint8_t* bc = top_bc(st);
  bool has_bc;
  double* fu = wsbuf<double>(st, WS_RES_FU, n);
  double* hp = wsbuf<double>(st, WS_RES_HP, np);
  lifted_rhs(st, fu, hp, &has_bc);
  // uf = u with BC dofs zeroed (K with identity BC rows maps those to themselves)
  double* uf = wsbuf<double>(st, WS_RES_UF, n);
  vk_copy<double>(n, u, uf, 0);
  launch_zero_bc(gv, bc, uf, 0);
  double* t = wsbuf<double>(st, WS_RES_T, n);
  double* gp = wsbuf<double>(st, WS_RES_GP, n);
  top_K<double>(st, uf, t, bc);
  top_G<double>(st, p, gp, bc);
  vk_axpy<double>(n, 1.0, gp, t, 0);
  vk_axpby<double>(n, 1.0, fu, -1.0, t, 0);  // t = fu - K uf - G p  (BC rows: fu=0, K uf = 0)
  double* q = wsbuf<double>(st, WS_RES_Q, np);
  top_GT<double>(st, uf, q, bc);
  vk_axpby<double>(np, 1.0, hp, -1.0, q, 0);
  out3[0] = sqrt(dev_dot(st, n, t, t));
  out3[1] = sqrt(dev_dot(st, np, q, q));
  out3[2] = sqrt(dev_dot(st, n, fu, fu));
}

extern "C" int uwg_stokes_residuals(uwg_stokes* st, double* out3) {
  if (!st->assembled) return arg_error("system not assembled");
  API_BEGIN
  true_residuals(st, st->vel->d, st->p->d, out3);
  return 0;
  API_END
}

// Schur complement solve following BSSCR_DRIVER_auglag (auglag-driver-DGTGD.c:64-584):
//   u* = K^-1 f ; h^ = h - G^T u* ; FGMRES on S p = h^, S = -G^T K^-1 G, PC approxS ;
//   u = K^-1 (f - G p)
static int solve_schur(uwg_stokes* st, const uwg_solve_opts& o, uwg_solve_stats* S, const double* fu,
                       const double* hp, double* u, double* p) {
  const long n = st->nu, E = st->np;
  const uint8_t* bc = top_bc(st);
  if (!st->kin) st->kin = new KrylovSpace();
  if (!st->kout) st->kout = new KrylovSpace();
  KrylovSpace& Kin = *st->kin;
  KrylovSpace& Kout = *st->kout;
  Kin.init(n, o.restart);
  Kout.init(E, o.restart);
  int status = 0;
  OpFn Kop = [&](const double* x, double* y) { top_K<double>(st, x, y, bc); };
  OpFn Mv = [&](const double* x, double* y) {
    st->vcycles++;
    if (o.precision != 0) {
      float* b32 = top_buf_b<float>(st);
      float* x32 = top_buf_x<float>(st);
      vk_d2f(n, 1.0, x, b32, 0);
      top_vcycle<float>(st, b32, x32);
      vk_f2d(n, 1.0, x32, y, 0);
    } else {
      double* b64 = top_buf_b<double>(st);
      vk_copy<double>(n, x, b64, 0);
      top_vcycle<double>(st, b64, y);
    }
  };
  auto inner = [&](const double* rhs, double* x, int* its) {
    vk_fill<double>(n, 0.0, x, 0);
    double rn;
    int rc = fgmres(st, Kin, Kop, Mv, rhs, x, o.inner_rtol, 1e-300, o.max_inner_its, its, &rn, true);
    if (rc) status = UWG_WARN_NOT_CONVERGED;
  };
  double* ustar = wsbuf<double>(st, WS_SCH_USTAR, n);
  double* hhat = wsbuf<double>(st, WS_SCH_HHAT, E);
  double* t = wsbuf<double>(st, WS_SCH_T, n);
  double* z = wsbuf<double>(st, WS_SCH_Z, n);
  double* diagS = wsbuf<double>(st, WS_SCH_DS, E);
  int its = 0;
  inner(fu, ustar, &its);
  S->velocity_presolve_its = its;
  top_GT<double>(st, ustar, hhat, bc);
  vk_axpby<double>(E, 1.0, hp, -1.0, hhat, 0);  // h^ = h - G^T u*
  int pss_its = 0;
  OpFn Sop = [&](const double* v, double* y) {
    top_G<double>(st, v, t, bc);
    int ii = 0;
    inner(t, z, &ii);
    pss_its += ii;
    top_GT<double>(st, z, y, bc);
    vk_scale<double>(E, -1.0, y, y, 0);  // S v = -G^T K^-1 G v
  };
  // preconditioner: -approxS^-1 (approxS approximates G^T diag(K)^-1 G > 0)
  if (!st->q2) launch_neg_recip<double>(E, st->approxS, diagS, 0);
  OpFn Sprec = [&](const double* x, double* y) { top_schur<double>(st, diagS, x, y); };
  vk_fill<double>(E, 0.0, p, 0);
  double rn;
  int oits = 0;
  int rc = fgmres(st, Kout, Sop, Sprec, hhat, p, o.outer_rtol, 1e-300, o.max_outer_its, &oits, &rn, true);
  if (rc) status = UWG_WARN_NOT_CONVERGED;
  S->pressure_its = oits;
  S->velocity_pressuresolve_its = pss_its;
  // back solve
  top_G<double>(st, p, t, bc);
  vk_axpby<double>(n, 1.0, fu, -1.0, t, 0);
  inner(t, u, &its);
  S->velocity_backsolve_its = its;
  S->velocity_total_its = S->velocity_presolve_its + S->velocity_pressuresolve_its + S->velocity_backsolve_its;
  return status;
}

// Coupled solve: FGMRES on [[K, G],[G^T, 0]] with the block lower-triangular right
// preconditioner  z_u = V-cycle(r_u), z_p = -approxS^-1 (r_p - G^T z_u).
static int inner_solve_fp64(uwg_stokes* st, const double* b, double* x, double rtol, int* its);

static int solve_block(uwg_stokes* st, const uwg_solve_opts& o, uwg_solve_stats* S, const double* fu,
                       const double* hp, double* u, double* p) {
  const long n = st->nu, E = st->np, nt = n + E;
  const uint8_t* bc = top_bc(st);
  if (!st->kb) st->kb = new KrylovSpace();
  KrylovSpace& Kb = *st->kb;
  double* b = wsbuf<double>(st, WS_BLK_B, nt);
  double* x = wsbuf<double>(st, WS_BLK_X, nt);
  double* diagS = wsbuf<double>(st, WS_BLK_DS, E);
  double* q = wsbuf<double>(st, WS_BLK_Q, E);
  vk_copy<double>(n, fu, b, 0);
  vk_copy<double>(E, hp, b + n, 0);
  vk_fill<double>(nt, 0.0, x, 0);
  if (!st->q2) launch_neg_recip<double>(E, st->approxS, diagS, 0);
  double* gp = wsbuf<double>(st, WS_BLK_GP, n);
  OpFn Aop = [&](const double* xx, double* yy) {
    top_K<double>(st, xx, yy, bc);
    top_G<double>(st, xx + n, gp, bc);
    vk_axpy<double>(n, 1.0, gp, yy, 0);
    top_GT<double>(st, xx, yy + n, bc);
  };
  OpFn Mop = [&](const double* rr, double* zz) {
    st->vcycles++;
    if (o.precision != 0) {
      float* b32 = top_buf_b<float>(st);
      float* x32 = top_buf_x<float>(st);
      vk_d2f(n, 1.0, rr, b32, 0);
      top_vcycle<float>(st, b32, x32);
      vk_f2d(n, 1.0, x32, zz, 0);
    } else if (st->exact_vel_rtol > 0.0) {
      int vits = 0;  // K z_u = r_u solved (the iteration floor of the Schur-block choice)
      inner_solve_fp64(st, rr, zz, st->exact_vel_rtol, &vits);
      st->vcycles += vits - 1;
    } else {
      double* b64 = top_buf_b<double>(st);
      vk_copy<double>(n, rr, b64, 0);
      top_vcycle<double>(st, b64, zz);
    }
    top_GT<double>(st, zz, q, bc);
    vk_axpby<double>(E, 1.0, rr + n, -1.0, q, 0);  // q = r_p - G^T z_u
    schur_prec<double>(st, diagS, q, zz + n);
  };
  int its = 0;
  double rn;
  // learned-correction safeguard: the FP64 residual must keep falling; if 20 iterations
  // reduce it by less than 1 %, the learned component is switched off and the plain
  // preconditioner takes over (a shorter window trips on the plateaus after restarts; a
  // larger required drop trips in the slow tail of tight tolerances, where the plain
  // preconditioner is slower still)
  // (FGMRES admits a preconditioner that changes between iterations)
  // hist[k] is the relative residual after iteration k (hist[0] = 1, the start), so the test
  // compares the residual with the one nn_window iterations earlier
  //
  // second test: within a cycle the Arnoldi estimate never rises, so a rise from one iteration
  // to the next is the first estimate after a restart exceeding the last one before it.  With
  // the FP32 Arnoldi basis, long cycles under the learned block accumulate basis rounding that
  // the estimate does not see (128^3 sinker, restart 70: 1.28e-5 estimated, 2.09e-5 after the
  // restart; 96^3 convection: rises of at most 1.01x).  A rise above nn_rise means the error
  // left is that rounding rather than smooth error, which the diagonal Schur block removes in a
  // few iterations (sinker: 89 iterations instead of 172 at restart 70) and the learned
  // stencil, fitted to smooth error, removes slowly.
  std::vector<double> hist{1.0};
  std::function<void(int, double)> mon = [&](int k, double rel) {
    hist.push_back(rel);
    const int wdw = st->nn_window;
    const bool learned_on = st->nn || st->snn;
    if (!o.learned || !learned_on || st->nn_disabled || k < 1 || k >= (int)hist.size()) return;
    const bool stalled = k >= wdw && rel > (1.0 - st->nn_min_drop) * hist[k - wdw];
    const bool rose = st->nn_rise > 0.0 && rel > st->nn_rise * hist[k - 1];
    if (stalled || rose) {
      st->nn_disabled = true;
      S->fallback_used = 1;
      if (g_verbose)
        fprintf(stderr, " learned correction disabled at iteration %d (rel %.3e, %s)\n", k, rel,
                rose ? "rise after restart" : "stalled");
    }
  };
  st->nn_disabled = !o.learned;
  g_snn_active = o.learned && st->snn && !st->q2;
  if (g_snn_active) ensure_snn_coefficients(st);
  int rc;
  if (o.precision == 1 && g_basis_fp32) {
    // FP32 Arnoldi basis: the FP32 cycle reads the basis vector's velocity part as it is (no
    // conversion pass); Z, w, the iterate and every reduction stay FP64, and the restarts and
    // the stopping test use the FP64 true residual
    if (!st->kb32) st->kb32 = new KSpaceT<double, float>();
    KSpaceT<double, float>& Kb32 = *st->kb32;
    Kb32.init(nt, o.restart);
    Kb.release();  // the FP64 basis of the other precisions is not kept alongside
    double* rp = wsbuf<double>(st, WS_BLK_RP, E);
    std::function<void(const float*, double*)> Mop32 = [&](const float* rr, double* zz) {
      st->vcycles++;
      float* x32 = top_buf_x<float>(st);
      top_vcycle<float>(st, rr, x32);
      vk_f2d(n, 1.0, x32, zz, 0);
      top_GT<double>(st, zz, q, bc);
      vk_f2d(E, 1.0, rr + n, rp, 0);
      vk_axpby<double>(E, 1.0, rp, -1.0, q, 0);  // q = r_p - G^T z_u
      schur_prec<double>(st, diagS, q, zz + n);
    };
    rc = fgmres<double, float>(st, Kb32, Aop, Mop32, b, x, o.outer_rtol, 1e-300, o.max_outer_its, &its, &rn, true,
                               &mon);
  } else {
    if (st->kb32) st->kb32->release();
    Kb.init(nt, o.restart);
    rc = fgmres<double>(st, Kb, Aop, Mop, b, x, o.outer_rtol, 1e-300, o.max_outer_its, &its, &rn, true, &mon);
  }
  S->outer_its = its;
  vk_copy<double>(n, x, u, 0);
  vk_copy<double>(E, x + n, p, 0);
  return rc ? UWG_WARN_NOT_CONVERGED : 0;
}

// Mixed precision (precision 1): FP64 iterative refinement around an FP32 inner FGMRES.
//   r = b - A x            FP64 operators (K with FP64 weights, G, G^T)
//   A32 d = r / |r|        FGMRES with FP32 vectors, FP32 K/G/G^T applies, FP32 multigrid
//                          cycle and Schur scaling; dot products and the Hessenberg problem
//                          accumulate in FP64
//   x += |r| d
// The stopping test is always the FP64 true residual.
static int solve_block_mixed(uwg_stokes* st, const uwg_solve_opts& o, uwg_solve_stats* S, const double* fu,
                             const double* hp, double* u, double* p) {
  const long n = st->nu, E = st->np, nt = n + E;
  const uint8_t* bc = top_bc(st);
  if (!st->kf) st->kf = new KSpaceT<float>();
  KSpaceT<float>& Kf = *st->kf;
  Kf.init(nt, o.restart);
  double* b = wsbuf<double>(st, WS_MIX_B, nt);
  double* x = wsbuf<double>(st, WS_MIX_X, nt);
  double* r = wsbuf<double>(st, WS_MIX_R, nt);
  double* gp = wsbuf<double>(st, WS_MIX_GP, n);
  float* r32 = wsbuf<float>(st, WS_MIX_R32, nt);
  float* d32 = wsbuf<float>(st, WS_MIX_D32, nt);
  float* diagS32 = wsbuf<float>(st, WS_MIX_DS32, E);
  float* q32 = wsbuf<float>(st, WS_MIX_Q32, E);
  float* gp32 = wsbuf<float>(st, WS_MIX_GP32, n);
  vk_copy<double>(n, fu, b, 0);
  vk_copy<double>(E, hp, b + n, 0);
  vk_fill<double>(nt, 0.0, x, 0);
  if (!st->q2) launch_neg_recip<float>(E, st->approxS, diagS32, 0);
  auto A64 = [&](const double* xx, double* yy) {
    top_K<double>(st, xx, yy, bc);
    top_G<double>(st, xx + n, gp, bc);
    vk_axpy<double>(n, 1.0, gp, yy, 0);
    top_GT<double>(st, xx, yy + n, bc);
  };
  OpFn32 A32 = [&](const float* xx, float* yy) {
    top_K<float>(st, xx, yy, bc);
    top_G<float>(st, xx + n, gp32, bc);
    vk_axpy<float>(n, 1.0f, gp32, yy, 0);
    top_GT<float>(st, xx, yy + n, bc);
  };
  OpFn32 M32 = [&](const float* rr, float* zz) {
    st->vcycles++;
    top_vcycle<float>(st, rr, zz);
    top_GT<float>(st, zz, q32, bc);
    vk_axpby<float>(E, 1.0f, rr + n, -1.0f, q32, 0);  // q = r_p - G^T z_u
    top_schur<float>(st, diagS32, q32, zz + n);
  };
  const double bn = sqrt(dev_dot<double>(st, nt, b, b));
  int total = 0, sweeps = 0, status = 0;
  double rn = bn;
  while (true) {
    if (sweeps == 0) {
      vk_copy<double>(nt, b, r, 0);
    } else {
      A64(x, r);
      vk_axpby<double>(nt, 1.0, b, -1.0, r, 0);
      rn = sqrt(dev_dot<double>(st, nt, r, r));
    }
    if (g_verbose) fprintf(stderr, " refinement sweep %d: FP64 |r|/|b| = %.3e\n", sweeps, rn / bn);
    if (rn <= o.outer_rtol * bn || total >= o.max_outer_its) break;
    if (sweeps >= 12) {
      status = UWG_WARN_NOT_CONVERGED;
      break;
    }
    vk_d2f(nt, 1.0 / rn, r, r32, 0);
    vk_fill<float>(nt, 0.0f, d32, 0);
    // ask the FP32 solve for the remaining reduction, bounded by what FP32 can deliver
    double want = 0.5 * o.outer_rtol * bn / rn;
    static const double floor32 = getenv("UWG_FP32_FLOOR") ? atof(getenv("UWG_FP32_FLOOR")) : 1e-6;
    double inner = std::min(0.5, std::max(want, floor32));
    int its = 0;
    double irn;
    fgmres<float>(st, Kf, A32, M32, r32, d32, inner, 0.0, o.max_outer_its - total, &its, &irn, true);
    total += its;
    vk_f2d_axpy(nt, rn, d32, x, 0);
    sweeps++;
    if (its == 0) {
      status = UWG_WARN_NOT_CONVERGED;
      break;
    }
  }
  S->outer_its = total;
  S->pressure_its = sweeps;  // refinement sweeps
  vk_copy<double>(n, x, u, 0);
  vk_copy<double>(E, x + n, p, 0);
  if (rn > o.outer_rtol * bn) status = UWG_WARN_NOT_CONVERGED;
  return status;
}

extern "C" int uwg_stokes_solve(uwg_stokes* st, const uwg_solve_opts* opts, uwg_solve_stats* S) {
  uwg_solve_opts o;
  if (opts)
    o = *opts;
  else
    uwg_solve_opts_default(&o);
  uwg_solve_stats dummy;
  if (!S) S = &dummy;
  memset(S, 0, sizeof(*S));
  if (o.restart < 2 || o.restart > 200) return arg_error("restart must be in [2, 200]");
  if (o.method == 0 && o.learned) return arg_error("the learned correction is only available with method 1");
  if (o.learned && st->nn && o.precision != 1)
    return arg_error("the learned coarse correction runs inside the FP32 cycle (precision 1)");
  if (o.learned && o.precision == 2) return arg_error("the learned preconditioners are not provided with precision 2");
  if (st->exact_vel_rtol > 0.0 && (o.method != 1 || o.precision != 0))
    return arg_error("the solved velocity block is provided for the coupled FP64 solve (method 1, precision 0)");
  API_BEGIN
  double t0 = now();
  if (o.mg_levels > 0 && o.mg_levels != (int)st->L.size()) {
    st->levels_req = o.mg_levels;
    build_levels(st, o.mg_levels);  // keeps the level-0 weights
    if (st->nn && st->nn_level >= (int)st->L.size()) {
      st->nn.reset();  // its level is gone
      st->nn_level = -1;
    }
    if (st->assembled) setup_hierarchy(st);
  }
  if (o.learned && !st->nn && !st->snn) return arg_error("no learned model loaded");
  if (o.learned && st->q2) return arg_error("the learned correction is provided for Q1/dQ0 systems");
  if (st->swarm && st->assembled_state != st->swarm->state && st->assembled_state != -2) {
    int rc = uwg_stokes_assemble(st, st->swarm);
    if (rc) return rc;
  }
  if (!st->assembled) return arg_error("system not assembled");
  g_smooth_degree = std::max(1, o.smooth_degree);
  g_verbose = o.verbose || getenv("UWG_VERBOSE") != nullptr;
  if (o.precision != 0) ensure_fp32(st);
  UWG_HIP(hipDeviceSynchronize());
  double t1 = now();
  st->vcycles = 0;
  const long n = st->nu, E = st->np;
  bool has_bc;
  double* fu = wsbuf<double>(st, WS_FU, n);
  double* hp = wsbuf<double>(st, WS_HP, E);
  lifted_rhs(st, fu, hp, &has_bc);
  double* u = wsbuf<double>(st, WS_U, n);
  double* p = wsbuf<double>(st, WS_P, E);
  int status;
  if (o.method == 0)
    status = solve_schur(st, o, S, fu, hp, u, p);
  else if (o.precision == 2)
    status = solve_block_mixed(st, o, S, fu, hp, u, p);
  else
    status = solve_block(st, o, S, fu, hp, u, p);
  // u (free dofs) + Dirichlet values -> velocity field ; p -> pressure field
  if (has_bc) vk_axpy<double>(n, 1.0, st->ubc, u, 0);
  else {
    // keep Dirichlet entries of the velocity field (zero here): zero them in u
  }
  vk_copy<double>(n, u, st->vel->d, 0);
  vk_copy<double>(E, p, st->p->d, 0);
  UWG_HIP(hipDeviceSynchronize());
  double t2 = now();
  double r3[3];
  true_residuals(st, st->vel->d, st->p->d, r3);
  S->res_momentum = r3[0];
  S->res_continuity = r3[1];
  S->norm_f = r3[2];
  S->vcycles = st->vcycles;
  S->levels = (int)st->L.size();
  S->converged = status == 0;
  S->setup_time = t1 - t0;
  S->solve_time = t2 - t1;
  S->total_time = t2 - t0;
  if (!std::isfinite(r3[0]) || !std::isfinite(r3[1])) {
    set_error("A floating-point error has been detected during the solve.");
    return UWG_WARN_NONFINITE;
  }
  if (status) set_error("Krylov solve did not converge within the iteration limit");
  return status;
  API_END
}

// ---------------------------------------------------------------------------------------
// hooks
// ---------------------------------------------------------------------------------------
extern "C" int uwg_stokes_apply(uwg_stokes* st, int which, int level, const double* x, double* y) {
  if (!st->assembled) return arg_error("system not assembled");
  if (st->q2) {
    // level 0 is the Q2/dPc1 level, level l >= 1 the Q1 level l-1 below it
    if (level < 0 || level > (int)st->L.size()) return arg_error("bad level");
    if (level == 0) {
      if (which < 0 || which > 4) return arg_error("which must be 0..4");
      API_BEGIN
      const long nin = which == 1 ? st->np : st->nu, nout = which == 2 ? st->np : st->nu;
      double* dx = dalloc<double>(nin);
      double* dy = dalloc<double>(nout);
      UWG_HIP(hipMemcpy(dx, x, sizeof(double) * nin, hipMemcpyHostToDevice));
      if (which == 4) {
        // the FP32 K of the mixed cycle on x rounded to FP32, widened to FP64
        ensure_fp32(st);
        float* fx = dalloc<float>(nin);
        float* fy = dalloc<float>(nout);
        vk_d2f(nin, 1.0, dx, fx, 0);
        top_K<float>(st, fx, fy, st->q2->bc);
        vk_f2d(nout, 1.0, fy, dy, 0);
        dfree(fx);
        dfree(fy);
      } else if (which == 0)
        top_K<double>(st, dx, dy, st->q2->bc);
      else if (which == 1)
        top_G<double>(st, dx, dy, st->q2->bc);
      else if (which == 2)
        top_GT<double>(st, dx, dy, st->q2->bc);
      else {
        vk_copy<double>(nin, dx, st->q2->b, 0);
        q2_vcycle<double>(st, st->q2->b, dy);
      }
      UWG_HIP(hipMemcpy(y, dy, sizeof(double) * nout, hipMemcpyDeviceToHost));
      dfree(dx);
      dfree(dy);
      return 0;
      API_END
    }
    if (which != 0) return arg_error("only K (which 0) is available on the Q1 levels of a Q2 system");
    level -= 1;
  }
  if (level < 0 || level >= (int)st->L.size()) return arg_error("bad level");
  API_BEGIN
  Level& A = st->L[level];
  long nin, nout;
  if (which == 0 || (which >= 3 && which <= 7)) nin = nout = A.n();
  else if (which == 1) { nin = st->g.E; nout = st->L[0].n(); }
  else if (which == 2) { nin = st->L[0].n(); nout = st->g.E; }
  else return arg_error("which must be 0..7");
  double* dx = dalloc<double>(nin);
  double* dy = dalloc<double>(nout);
  UWG_HIP(hipMemcpy(dx, x, sizeof(double) * nin, hipMemcpyHostToDevice));
  if (which == 6) {
    // the multigrid residual kernel (apply with the r = b - K x epilogue), b = x
    launch_residual_K<double>(A.g, A.W, dx, dx, dy, A.bc, 0);
    UWG_HIP(hipMemcpy(y, dy, sizeof(double) * nout, hipMemcpyDeviceToHost));
    dfree(dx);
    dfree(dy);
    return 0;
  }
  if (which == 4 || which == 5 || which == 7) {
    // the FP32 cycle's K (4: with the cycle's weights, bfloat16 on 3D levels when enabled;
    // 5: FP32 weights; 7: the cycle's residual kernel r = b - K x with b = x) on x rounded to
    // FP32, result widened to FP64
    ensure_fp32(st);
    float* fx = dalloc<float>(nin);
    float* fy = dalloc<float>(nout);
    vk_d2f(nin, 1.0, dx, fx, 0);
    if (which == 7)
      launch_residual_K<float>(A.g, A.W32, fx, fx, fy, A.bc, 0, A.Wbf);
    else
      launch_apply_K<float>(A.g, A.W32, fx, fy, A.bc, 0, which == 4 ? A.Wbf : nullptr);
    vk_f2d(nout, 1.0, fy, dy, 0);
    UWG_HIP(hipMemcpy(y, dy, sizeof(double) * nout, hipMemcpyDeviceToHost));
    dfree(fx);
    dfree(fy);
    dfree(dx);
    dfree(dy);
    return 0;
  }
  if (which == 0)
    launch_apply_K<double>(A.g, A.W, dx, dy, A.bc, 0);
  else if (which == 1)
    launch_apply_G(st->g, st->Ge, dx, st->L[0].bc, dy, 0);
  else if (which == 2)
    launch_apply_GT(st->g, st->Ge, dx, st->L[0].bc, dy, 0);
  else {
    vk_copy<double>(nin, dx, st->L[0].b, 0);
    vcycle<double>(st, 0, st->L[0].b, dy);
  }
  UWG_HIP(hipMemcpy(y, dy, sizeof(double) * nout, hipMemcpyDeviceToHost));
  dfree(dx);
  dfree(dy);
  return 0;
  API_END
}

extern "C" int uwg_stokes_levels(uwg_stokes* st, int* nlev, int* res_out) {
  const int off = st->q2 ? 1 : 0;
  *nlev = (int)st->L.size() + off;
  if (res_out) {
    if (off)
      for (int d = 0; d < 3; d++) res_out[d] = st->g.res[d];
    for (size_t l = 0; l < st->L.size(); l++)
      for (int d = 0; d < 3; d++) res_out[(l + off) * 3 + d] = st->L[l].g.res[d];
  }
  return 0;
}

extern "C" int uwg_stokes_get_assembly(uwg_stokes* st, int level, double* W, double* fe, double* approxS) {
  if (!st->assembled) return arg_error("system not assembled");
  if (st->q2) {
    if (level < 0 || level > (int)st->L.size()) return arg_error("bad level");
    if (level == 0) {
      // W [E][5^dim], fe [E][3^dim][dim], approxS blocks [E][dim+1][dim+1]
      API_BEGIN
      const Geom& g = st->g;
      const long nq = g.E * (g.dim == 3 ? 125 : 25), nf = g.E * (g.dim == 3 ? 27 : 9) * g.dim;
      const long ns = g.E * (g.dim + 1) * (g.dim + 1);
      if (W) UWG_HIP(hipMemcpy(W, st->q2->W, sizeof(double) * nq, hipMemcpyDeviceToHost));
      if (fe) UWG_HIP(hipMemcpy(fe, st->q2->fe, sizeof(double) * nf, hipMemcpyDeviceToHost));
      if (approxS) UWG_HIP(hipMemcpy(approxS, st->q2->S, sizeof(double) * ns, hipMemcpyDeviceToHost));
      return 0;
      API_END
    }
    if (fe || approxS) return arg_error("fe and approxS exist on level 0 only");
    level -= 1;
  }
  if (level < 0 || level >= (int)st->L.size()) return arg_error("bad level");
  API_BEGIN
  Level& A = st->L[level];
  const long Q = A.g.dim == 3 ? 27 : 9;
  (void)Q;
  if (W) copy_W_to_host(A, W);
  if (level == 0) {
    const int NN = 1 << st->g.dim;
    if (fe) UWG_HIP(hipMemcpy(fe, st->fe, sizeof(double) * st->g.E * NN * st->g.dim, hipMemcpyDeviceToHost));
    if (approxS) UWG_HIP(hipMemcpy(approxS, st->approxS, sizeof(double) * st->g.E, hipMemcpyDeviceToHost));
  }
  return 0;
  API_END
}

extern "C" int uwg_stokes_get_rhs(uwg_stokes* st, double* f) {
  API_BEGIN
  UWG_HIP(hipMemcpy(f, st->f, sizeof(double) * st->nu, hipMemcpyDeviceToHost));
  return 0;
  API_END
}

extern "C" int uwg_set_q2_mfma(int on) {
  set_q2_mfma(on);
  return 0;
}

extern "C" int uwg_set_q2_apply_kernel(int kind) {
  if (kind < 0 || kind > 2) return arg_error("kind must be 0 (vector), 1 (matrix cores) or 2 (sum-factorised)");
  set_q2_apply_kernel(kind);
  return 0;
}

extern "C" int uwg_set_cycle_bf16(int on) {
  g_cycle_bf16 = on != 0;
  return 0;
}

extern "C" int uwg_set_krylov_basis_fp32(int on) {
  g_basis_fp32 = on != 0;
  return 0;
}

extern "C" int uwg_set_apply_chunk(int kz3, int kz2) {
  set_apply_chunk(kz3, kz2);
  return 0;
}

// STREAM-style copy probe: y = x over n doubles, timed with HIP events; the kernel is also the
// calibration point for hardware byte counters (known bytes read and written per launch).
extern "C" int uwg_stream_copy(long n, int reps, double* ms) {
  if (n < 1 || reps < 1) return arg_error("stream copy: n and reps must be positive");
  API_BEGIN
  double* x = dalloc<double>(n);
  double* y = dalloc<double>(n);
  vk_fill<double>(n, 1.0, x, 0);
  vk_copy<double>(n, x, y, 0);
  hipEvent_t e0, e1;
  UWG_HIP(hipEventCreate(&e0));
  UWG_HIP(hipEventCreate(&e1));
  UWG_HIP(hipEventRecord(e0, 0));
  for (int i = 0; i < reps; i++) vk_copy<double>(n, x, y, 0);
  UWG_HIP(hipEventRecord(e1, 0));
  UWG_HIP(hipEventSynchronize(e1));
  float t;
  UWG_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  UWG_HIP(hipEventDestroy(e0));
  UWG_HIP(hipEventDestroy(e1));
  dfree(x);
  dfree(y);
  return 0;
  API_END
}

extern "C" int uwg_stokes_time_apply(uwg_stokes* st, int precision, int level, int reps, double* ms) {
  if (!st->assembled) return arg_error("system not assembled");
  if (reps < 1) return arg_error("reps must be positive");
  if (st->q2) {
    if (level < 0 || level > (int)st->L.size()) return arg_error("bad level");
    if (level == 0) {
      API_BEGIN
      Q2Level& A = *st->q2;
      hipEvent_t e0, e1;
      UWG_HIP(hipEventCreate(&e0));
      UWG_HIP(hipEventCreate(&e1));
      if (precision == 1) {
        ensure_fp32(st);
        top_K<float>(st, A.x32, A.r32, A.bc);
        UWG_HIP(hipEventRecord(e0, 0));
        for (int i = 0; i < reps; i++) top_K<float>(st, A.x32, A.r32, A.bc);
        UWG_HIP(hipEventRecord(e1, 0));
      } else {
        top_K<double>(st, A.x, A.r, A.bc);
        UWG_HIP(hipEventRecord(e0, 0));
        for (int i = 0; i < reps; i++) top_K<double>(st, A.x, A.r, A.bc);
        UWG_HIP(hipEventRecord(e1, 0));
      }
      UWG_HIP(hipEventSynchronize(e1));
      float t;
      UWG_HIP(hipEventElapsedTime(&t, e0, e1));
      *ms = t / reps;
      UWG_HIP(hipEventDestroy(e0));
      UWG_HIP(hipEventDestroy(e1));
      return 0;
      API_END
    }
    level -= 1;
  }
  if (level < 0 || level >= (int)st->L.size()) return arg_error("bad level");
  API_BEGIN
  Level& A = st->L[level];
  hipEvent_t e0, e1;
  UWG_HIP(hipEventCreate(&e0));
  UWG_HIP(hipEventCreate(&e1));
  if (precision == 1 || precision == 2) {
    // 1: the FP32 cycle's apply (bfloat16 weights when the cycle uses them), 2: FP32 weights
    ensure_fp32(st);
    const uint32_t* wb = precision == 1 ? A.Wbf : nullptr;
    launch_apply_K<float>(A.g, A.W32, A.x32, A.r32, A.bc, 0, wb);
    UWG_HIP(hipEventRecord(e0, 0));
    for (int i = 0; i < reps; i++) launch_apply_K<float>(A.g, A.W32, A.x32, A.r32, A.bc, 0, wb);
    UWG_HIP(hipEventRecord(e1, 0));
  } else {
    launch_apply_K<double>(A.g, A.W, A.x, A.r, A.bc, 0);
    UWG_HIP(hipEventRecord(e0, 0));
    for (int i = 0; i < reps; i++) launch_apply_K<double>(A.g, A.W, A.x, A.r, A.bc, 0);
    UWG_HIP(hipEventRecord(e1, 0));
  }
  UWG_HIP(hipEventSynchronize(e1));
  float t;
  UWG_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  UWG_HIP(hipEventDestroy(e0));
  UWG_HIP(hipEventDestroy(e1));
  return 0;
  API_END
}

extern "C" int uwg_stokes_load_nn(uwg_stokes* st, int level, int nparams, const float* params, int channels, int depth) {
  if (st->q2) return arg_error("the learned correction is provided for Q1/dQ0 systems");
  if (level < 1 || level >= (int)st->L.size()) return arg_error("learned correction level out of range");
  API_BEGIN
  st->nn.reset(nn_create(st->L[level].g, channels, depth, params, nparams));
  st->nn_level = level;
  st->nn_disabled = false;
  if (st->assembled) {
    std::vector<float> f;
    double eta_ref;
    level_feature(st, level, f, &eta_ref);
    nn_set_feature(st->nn.get(), f.data(), eta_ref);
  }
  return 0;
  API_END
}

extern "C" int uwg_stokes_set_nn_mode(uwg_stokes* st, int mode) {
  if (mode != 0 && mode != 1) return arg_error("mode must be 0 (replace) or 1 (add)");
  st->nn_mode = mode;
  return 0;
}

extern "C" int uwg_stokes_set_velocity_block_solve(uwg_stokes* st, double rtol) {
  if (!(rtol >= 0.0 && rtol < 1.0)) return arg_error("rtol must be in [0, 1)");
  st->exact_vel_rtol = rtol;
  return 0;
}

extern "C" int uwg_stokes_set_nn_monitor(uwg_stokes* st, int window, double min_drop) {
  if (window < 1 || !(min_drop >= 0.0 && min_drop < 1.0)) return arg_error("window >= 1 and 0 <= min_drop < 1");
  st->nn_window = window;
  st->nn_min_drop = min_drop;
  return 0;
}

extern "C" int uwg_stokes_set_nn_restart_rise(uwg_stokes* st, double factor) {
  if (!(factor == 0.0 || (factor >= 1.0 && factor < 1e300))) return arg_error("factor must be 0 or >= 1");
  st->nn_rise = factor;
  return 0;
}

// ---------------------------------------------------------------------------------------
// learned Schur stencil: attach a trained network; generate training triples
// ---------------------------------------------------------------------------------------
extern "C" int uwg_stokes_load_schur_nn(uwg_stokes* st, int channels, int layers, long nparams, const float* params) {
  if (st->q2 || st->g.dim != 3) return arg_error("the learned Schur stencil is provided for 3D Q1/dQ0 systems");
  API_BEGIN
  st->snn.reset(schur_net_create(channels, layers, params, nparams));
  st->snn_stamp = -1;
  return 0;
  API_END
}

// x = K^-1 b: FGMRES(30) preconditioned by one FP64 V-cycle (the Schur path's inner solve)
static int inner_solve_fp64(uwg_stokes* st, const double* b, double* x, double rtol, int* its) {
  const long n = st->nu;
  const uint8_t* bc = top_bc(st);
  if (!st->kin) st->kin = new KrylovSpace();
  st->kin->init(n, 30);
  OpFn Kop = [&](const double* xx, double* yy) { top_K<double>(st, xx, yy, bc); };
  OpFn Mv = [&](const double* xx, double* yy) {
    double* b64 = top_buf_b<double>(st);
    vk_copy<double>(n, xx, b64, 0);
    top_vcycle<double>(st, b64, yy);
  };
  vk_fill<double>(n, 0.0, x, 0);
  double rn;
  return fgmres(st, *st->kin, Kop, Mv, b, x, rtol, 1e-300, 10000, its, &rn, true);
}

// Training triples from the plain solver's operators: e^ white noise (seeded) with the
// constant-pressure null mode removed, e = D^-1/2 e^ (D = approxS), r = S e = G^T K^-1 G e with
// K^-1 by the inner solve to rtol, r^ = D^-1/2 r; the sample (features(approxS), r^, e^) is
// appended to the network's training set.
extern "C" int uwg_snn_add_samples(uwg_snn* h, uwg_stokes* st, int nsamples, long seed, double rtol) {
  if (!st->assembled) return arg_error("system not assembled");
  if (st->q2 || st->g.dim != 3) return arg_error("the learned Schur stencil is provided for 3D Q1/dQ0 systems");
  if (nsamples < 1) return arg_error("nsamples must be >= 1");
  API_BEGIN
  const long E = st->np, n = st->nu;
  const uint8_t* bc = top_bc(st);
  std::vector<double> S(E), sq(E), eh(E), e(E), r(E);
  UWG_HIP(hipMemcpy(S.data(), st->approxS, sizeof(double) * E, hipMemcpyDeviceToHost));
  double nn2 = 0.0;
  for (long i = 0; i < E; i++) {
    sq[i] = sqrt(S[i]);
    nn2 += S[i];
  }
  double* de = dalloc<double>(E);
  double* deh = dalloc<double>(E);
  double* dr = dalloc<double>(E);
  double* t = dalloc<double>(n);
  double* z = dalloc<double>(n);
  std::mt19937_64 gen((unsigned long long)seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int k = 0; k < nsamples; k++) {
    double proj = 0.0;
    for (long i = 0; i < E; i++) {
      eh[i] = nd(gen);
      proj += eh[i] * sq[i];
    }
    proj /= nn2;  // remove the component along D^1/2 1 (the constant pressure)
    for (long i = 0; i < E; i++) {
      eh[i] -= proj * sq[i];
      e[i] = eh[i] / sq[i];
    }
    UWG_HIP(hipMemcpy(de, e.data(), sizeof(double) * E, hipMemcpyHostToDevice));
    top_G<double>(st, de, t, bc);
    int its = 0;
    inner_solve_fp64(st, t, z, rtol, &its);
    top_GT<double>(st, z, dr, bc);
    UWG_HIP(hipMemcpy(r.data(), dr, sizeof(double) * E, hipMemcpyDeviceToHost));
    for (long i = 0; i < E; i++) r[i] /= sq[i];
    UWG_HIP(hipMemcpy(dr, r.data(), sizeof(double) * E, hipMemcpyHostToDevice));
    UWG_HIP(hipMemcpy(deh, eh.data(), sizeof(double) * E, hipMemcpyHostToDevice));
    schur_net_append(h->m, st->g, st->approxS, dr, deh);
  }
  for (void* p : {(void*)de, (void*)deh, (void*)dr, (void*)t, (void*)z}) dfree(p);
  return 0;
  API_END
}

extern "C" int uwg_stokes_capture(uwg_stokes* st, int level, int kind, int max_samples, int every) {
  if (level < 1 || level >= (int)st->L.size()) return arg_error("capture level out of range");
  if (kind != 0 && kind != 1) return arg_error("capture kind must be 0 (right-hand side) or 1 (residual)");
  if (kind == 1 && level == (int)st->L.size() - 1) return arg_error("the coarsest level leaves no residual");
  API_BEGIN
  dfree(st->capture_buf);
  st->capture_buf = max_samples > 0 ? dalloc<float>((long)max_samples * st->L[level].n()) : nullptr;
  st->capture_level = max_samples > 0 ? level : -1;
  st->capture_kind = kind;
  st->capture_max = max_samples;
  st->capture_every = std::max(1, every);
  st->captured = 0;
  st->capture_seen = 0;
  return 0;
  API_END
}

extern "C" int uwg_stokes_get_captured(uwg_stokes* st, float* out, int* count) {
  API_BEGIN
  *count = st->captured;
  if (out && st->captured > 0)
    UWG_HIP(hipMemcpy(out, st->capture_buf, sizeof(float) * (long)st->captured * st->L[st->capture_level].n(),
                      hipMemcpyDeviceToHost));
  return 0;
  API_END
}

// K_l x = b on one level: FGMRES preconditioned by V-cycles starting at that level (FP64)
extern "C" int uwg_stokes_level_solve(uwg_stokes* st, int level, const double* b, double* x, double rtol, int* its) {
  if (!st->assembled) return arg_error("system not assembled");
  if (level < 0 || level >= (int)st->L.size()) return arg_error("bad level");
  API_BEGIN
  Level& A = st->L[level];
  const long n = A.n();
  KrylovSpace K;
  K.init(n, 30);
  double* db = dalloc<double>(n);
  double* dx = dalloc<double>(n);
  double* wb = dalloc<double>(n);
  UWG_HIP(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice));
  vk_fill<double>(n, 0.0, dx, 0);
  OpFn Aop = [&](const double* xx, double* yy) { launch_apply_K<double>(A.g, A.W, xx, yy, A.bc, 0); };
  OpFn Mop = [&](const double* rr, double* zz) {
    vk_copy<double>(n, rr, wb, 0);
    vcycle<double>(st, level, wb, zz);
  };
  int k = 0;
  double rn;
  int rc = fgmres<double>(st, K, Aop, Mop, db, dx, rtol, 0.0, 2000, &k, &rn, true);
  UWG_HIP(hipMemcpy(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (its) *its = k;
  for (void* p : {(void*)db, (void*)dx, (void*)wb}) dfree(p);
  return rc ? UWG_WARN_NOT_CONVERGED : 0;
  API_END
}

extern "C" int uwg_stokes_level_feature(uwg_stokes* st, int level, float* feat, double* eta_ref) {
  if (!st->assembled) return arg_error("system not assembled");
  if (level < 0 || level >= (int)st->L.size()) return arg_error("bad level");
  API_BEGIN
  std::vector<float> f;
  double er;
  level_feature(st, level, f, &er);
  if (eta_ref) *eta_ref = er;
  memcpy(feat, f.data(), sizeof(float) * f.size());
  return 0;
  API_END
}

extern "C" int uwg_stokes_level_mask(uwg_stokes* st, int level, unsigned char* mask) {
  if (level < 0 || level >= (int)st->L.size()) return arg_error("bad level");
  API_BEGIN
  UWG_HIP(hipMemcpy(mask, st->L[level].bc, st->L[level].g.N, hipMemcpyDeviceToHost));
  return 0;
  API_END
}

// ---------------------------------------------------------------------------------------
// SUPG advection-diffusion (AdvectionDiffusion(method="SUPG"), PY/systems/_advectiondiffusion.py
// :838-1010; AD/ = LU/Underworld/Utils/AdvectionDiffusion/src/)
// ---------------------------------------------------------------------------------------
struct uwg_advdiff {
  uwg_field* phi = nullptr;
  uwg_field* phidot = nullptr;
  uwg_field* vel = nullptr;
  Geom g;
  uint8_t* bc = nullptr;      // bit 0: Dirichlet vertex
  double* kap = nullptr;      // [E][3^dim] diffusivity at the Gauss points
  double* src = nullptr;      // [E][3^dim] source, or null
  double* res_e = nullptr;    // [E][2^dim]
  double* me = nullptr;       // [E][2^dim]
  double* M = nullptr;        // lumped mass [N]
  double* phi_v = nullptr;    // solution vectors [N] (free entries)
  double* pdot_v = nullptr;
  double* partial = nullptr;
  double kappa_max = 0.0;
  bool have_coef = false;
  double gamma = 0.5;         // AdvDiffMulticorrector defaults (Multicorrector.c:101-102)
  int iterations = 2;
  int upwind = 0;             // "Exact" (Residual.cpp:150)
  double min_sep = 0.0, sep[3] = {1.0, 1.0, 1.0};
};

extern "C" int uwg_advdiff_create(uwg_field* phi, uwg_field* phidot, uwg_field* vel, const unsigned char* bcmask,
                                  uwg_advdiff** out) {
  if (!phi || !phidot || !vel) return arg_error("phiField, phiDotField and velocityField are required");
  if (phi->on_elements || phi->ncomp != 1) return arg_error("phiField must be a nodal scalar field");
  if (phidot->on_elements || phidot->ncomp != 1 || phidot->mesh != phi->mesh)
    return arg_error("phiDotField must be a nodal scalar field on the phiField's mesh");
  if (phi->mesh->order != 1) return arg_error("SUPG advection-diffusion is provided for Q1 meshes");
  if (vel->on_elements || vel->mesh != phi->mesh || vel->ncomp != phi->mesh->g.dim)
    return arg_error("velocityField must be a nodal vector field on the phiField's mesh");
  API_BEGIN
  uwg_advdiff* a = new uwg_advdiff();
  a->phi = phi;
  a->phidot = phidot;
  a->vel = vel;
  a->g = phi->mesh->g;
  const Geom& g = a->g;
  const int NN = 1 << g.dim, Q = g.dim == 3 ? 27 : 9;
  a->bc = dalloc<uint8_t>(g.N);
  if (bcmask)
    UWG_HIP(hipMemcpy(a->bc, bcmask, g.N, hipMemcpyHostToDevice));
  else
    UWG_HIP(hipMemset(a->bc, 0, g.N));
  a->kap = dalloc<double>(g.E * Q);
  a->res_e = dalloc<double>(g.E * NN);
  a->me = dalloc<double>(g.E * NN);
  a->M = dalloc<double>(g.N);
  a->phi_v = dalloc<double>(g.N);
  a->pdot_v = dalloc<double>(g.N);
  a->partial = dalloc<double>(512);
  launch_supg_mass(g, a->me, a->M, 0);
  // minimum vertex separation and its per-axis values (Mesh_HexType_GetMinimumSeparation,
  // Mesh_HexType.c:177-272, taken from the first element attaining the minimum)
  std::vector<double> ax[3];
  for (int d = 0; d < g.dim; d++) {
    ax[d].resize(g.res[d]);
    for (int i = 0; i < g.res[d]; i++)
      ax[d][i] = node_coord_1d(i + 1, g.res[d] + 1, g.lo[d], g.hi[d]) - node_coord_1d(i, g.res[d] + 1, g.lo[d], g.hi[d]);
  }
  double best = HUGE_VAL;
  for (int k = 0; k < (g.dim == 3 ? g.res[2] : 1); k++)
    for (int j = 0; j < g.res[1]; j++)
      for (int i = 0; i < g.res[0]; i++) {
        const double s0 = ax[0][i], s1 = ax[1][j], s2 = g.dim == 3 ? ax[2][k] : HUGE_VAL;
        double cur = s0;
        cur = s1 < cur ? s1 : cur;
        if (g.dim == 3) cur = s2 < cur ? s2 : cur;
        if (cur < best) {
          best = cur;
          a->sep[0] = s0;
          a->sep[1] = s1;
          a->sep[2] = g.dim == 3 ? s2 : 1.0;
        }
      }
  a->min_sep = best;
  *out = a;
  return 0;
  API_END
}

extern "C" int uwg_advdiff_destroy(uwg_advdiff* a) {
  if (!a) return 0;
  for (void* p : {(void*)a->bc, (void*)a->kap, (void*)a->src, (void*)a->res_e, (void*)a->me, (void*)a->M,
                  (void*)a->phi_v, (void*)a->pdot_v, (void*)a->partial})
    dfree(p);
  delete a;
  return 0;
}

extern "C" int uwg_advdiff_set_coefficients(uwg_advdiff* a, const double* kappa_g, const double* source_g) {
  if (!kappa_g) return arg_error("diffusivity values are required");
  const Geom& g = a->g;
  const long nq = g.E * (g.dim == 3 ? 27 : 9);
  double kmax = -1.0;  // AdvDiffResidualForceTerm_GetMaxDiffusivity starts below any diffusivity
  for (long q = 0; q < nq; q++) {
    if (std::isnan(kappa_g[q])) return arg_error("diffusivity is NaN");
    if (kappa_g[q] > kmax) kmax = kappa_g[q];
  }
  API_BEGIN
  UWG_HIP(hipMemcpy(a->kap, kappa_g, sizeof(double) * nq, hipMemcpyHostToDevice));
  if (source_g) {
    if (!a->src) a->src = dalloc<double>(nq);
    UWG_HIP(hipMemcpy(a->src, source_g, sizeof(double) * nq, hipMemcpyHostToDevice));
  } else {
    dfree(a->src);
    a->src = nullptr;
  }
  a->kappa_max = kmax;
  a->have_coef = true;
  return 0;
  API_END
}

extern "C" int uwg_advdiff_set_scheme(uwg_advdiff* a, double gamma, int iterations, int upwind) {
  if (!(gamma >= 0.0 && gamma <= 1.0)) return arg_error("gamma must lie in [0, 1]");
  if (iterations < 1) return arg_error("at least one corrector iteration is required");
  if (upwind < 0 || upwind > 2) return arg_error("upwind must be 0 (exact), 1 (doubly asymptotic) or 2 (critical)");
  a->gamma = gamma;
  a->iterations = iterations;
  a->upwind = upwind;
  return 0;
}

// one time step (AdvDiffMulticorrector_Solve, Multicorrector.c:158-200): predictor on the
// solution vectors, then the corrector passes; each pass assembles the residual from the nodal
// fields and refreshes them from the vectors when it ends
extern "C" int uwg_advdiff_integrate(uwg_advdiff* a, double dt) {
  if (!a->have_coef) return arg_error("diffusivity / source not set");
  if (!std::isfinite(dt)) return arg_error("dt must be finite");
  API_BEGIN
  const Geom& g = a->g;
  launch_supg_predict(g.N, a->bc, a->phi->d, a->phidot->d, dt * (1.0 - a->gamma), a->phi_v, a->pdot_v, 0);
  for (int it = 0; it < a->iterations; it++) {
    launch_supg_elem(g, a->phi->d, a->phidot->d, a->vel->d, a->kap, a->src, a->upwind, a->res_e, 0);
    launch_supg_correct(g, a->bc, a->res_e, a->M, dt * a->gamma, a->phi_v, a->pdot_v, a->phi->d, a->phidot->d, 0);
  }
  return 0;
  API_END
}

// AdvectionDiffusionSLE_CalculateDt (Timestep.c:36-110): courant factor 0.5
extern "C" int uwg_advdiff_max_dt(uwg_advdiff* a, double* dt) {
  if (!a->have_coef) return arg_error("diffusivity / source not set");
  API_BEGIN
  const int nblk = 512;
  launch_supg_adv_dt(a->g, a->vel->d, a->sep, a->partial, nblk, 0);
  std::vector<double> h(nblk);
  UWG_HIP(hipMemcpy(h.data(), a->partial, sizeof(double) * nblk, hipMemcpyDeviceToHost));
  double adv = HUGE_VAL;
  for (double v : h) adv = v < adv ? v : adv;
  adv = 0.5 * adv;
  const double dif = 0.5 * a->min_sep * a->min_sep / a->kappa_max;
  *dt = adv < dif ? adv : dif;
  return 0;
  API_END
}

// ---------------------------------------------------------------------------------------
// population control: PopulationControl(swarm, deleteThreshold, splitThreshold, maxDeletions,
// maxSplits).repopulate() (PY/swarm/_population_control.py:13-140), the non-aggressive path of
// _PCDVC_Calculate3D / 2D (LU/PICellerator/Utils/src/PCDVC.c:313-700, 703-1085).
// The DVC weights of every cell come from the device pass; the edits are the reference's, in
// its order: cells in order; in a cell, particles whose weight exceeds split*2^dim are split
// (first maxSplits in table order; the copy is appended to the swarm and to the cell at the
// centroid of the parent's Voronoi cell, FeMesh_CoordLocalToGlobal), then particles below
// delete*2^dim are deleted (by descending swarm index, at most maxDeletions) with the swarm's
// swap-with-last removal (Swarm_DeleteParticle, SwarmClass.c:1016-1082).  src[i] receives the
// original index whose data particle i now carries (for the caller's per-particle variables).
// ---------------------------------------------------------------------------------------
extern "C" int uwg_swarm_population_control(uwg_swarm* s, double delete_threshold, double split_threshold,
                                            int max_deletions, int max_splits, long src_cap, long* src,
                                            long* p_new) {
  if (!(delete_threshold >= 0.0 && delete_threshold <= 1.0) || !(split_threshold >= 0.0 && split_threshold <= 1.0))
    return arg_error("thresholds must lie in [0, 1]");
  if (delete_threshold > split_threshold) return arg_error("'deleteThreshold' must be less than 'splitThreshold'");
  if (max_deletions < 0 || max_splits < 0) return arg_error("maxDeletions and maxSplits must be non-negative");
  const Geom& g = s->mesh->g;
  const int dim = g.dim;
  const long E = g.E;
  int rc = uwg_swarm_integration_points(s, 1, 15);
  if (rc) return rc;
  API_BEGIN
  const long P0 = s->P;
  std::vector<int> ptr(E + 1), pid(P0);
  std::vector<double> w(P0), cen((size_t)P0 * dim);
  UWG_HIP(hipMemcpy(ptr.data(), s->cell_ptr, sizeof(int) * (E + 1), hipMemcpyDeviceToHost));
  if (P0) {
    UWG_HIP(hipMemcpy(pid.data(), s->cell_pid, sizeof(int) * P0, hipMemcpyDeviceToHost));
    UWG_HIP(hipMemcpy(w.data(), s->w_slot, sizeof(double) * P0, hipMemcpyDeviceToHost));
    UWG_HIP(hipMemcpy(cen.data(), s->xi_slot, sizeof(double) * P0 * dim, hipMemcpyDeviceToHost));
  }
  const double vol = (double)(1 << dim);
  const double maxW = split_threshold * vol, minW = delete_threshold * vol;
  // any deletions at all?  (they need the full bookkeeping of current indices)
  bool any_del = false;
  if (max_deletions > 0)
    for (long k = 0; k < P0 && !any_del; k++) any_del = w[k] < minW;
  std::vector<long> rsrc(P0);                  // data source of the particle at each current index
  for (long k = 0; k < P0; k++) rsrc[k] = k;
  std::vector<int> rnew(P0, -1);               // index into newc for particles created here
  std::vector<double> newc;                    // coordinates of created particles
  std::vector<std::vector<int>> mod(E);        // tables of edited cells (others: the CSR slice)
  std::vector<char> edited(E, 0);
  std::vector<int> cell_of;                    // current index -> cell (deletions only)
  if (any_del) {
    cell_of.assign(P0, 0);
    for (long e = 0; e < E; e++)
      for (int k = ptr[e]; k < ptr[e + 1]; k++) cell_of[pid[k]] = (int)e;
  }
  auto table = [&](long e) -> std::vector<int>& {
    if (!edited[e]) {
      mod[e].assign(pid.begin() + ptr[e], pid.begin() + ptr[e + 1]);
      edited[e] = 1;
    }
    return mod[e];
  };
  long P = P0;
  for (long e = 0; e < E; e++) {
    const int s0 = ptr[e], n0 = ptr[e + 1] - ptr[e];
    std::vector<int> splits;
    std::vector<std::pair<long, int>> dels;  // (index on the swarm, position in the cell)
    for (int i = 0; i < n0; i++) {
      if (w[s0 + i] > maxW) splits.push_back(i);
      if (w[s0 + i] < minW) dels.push_back({edited[e] ? (long)mod[e][i] : (long)pid[s0 + i], i});
    }
    const int ns = std::min((int)splits.size(), max_splits);
    const int nd = std::min((int)dels.size(), max_deletions);
    if (ns == 0 && nd == 0) continue;
    std::vector<int>& T = table(e);
    for (int k = 0; k < ns; k++) {
      const int i = splits[k];
      const long parent = T[i];
      double xi[3] = {0.0, 0.0, 0.0}, v[24], N[8], x[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < dim; d++) xi[d] = cen[(size_t)(s0 + i) * dim + d];
      if (dim == 3) {
        element_verts<3>(g.res, g.lo, g.hi, e, v);
        shape<3>(xi, N);
      } else {
        element_verts<2>(g.res, g.lo, g.hi, e, v);
        shape<2>(xi, N);
      }
      for (int a = 0; a < (dim == 3 ? 8 : 4); a++)
        for (int d = 0; d < dim; d++) x[d] += N[a] * v[a * dim + d];
      rsrc.push_back(rsrc[parent]);
      rnew.push_back((int)(newc.size() / dim));
      for (int d = 0; d < dim; d++) newc.push_back(x[d]);
      if (any_del) cell_of.push_back((int)e);
      T.push_back((int)P);
      P++;
    }
    if (nd > 0) {
      std::sort(dels.begin(), dels.end(), [](const std::pair<long, int>& a, const std::pair<long, int>& b) {
        return a.first > b.first;
      });
      for (int k = 0; k < nd; k++) {
        const long idx = dels[k].first;
        // remove from its cell: the cell's last entry takes its place
        std::vector<int>& Tc = table(cell_of[idx]);
        for (size_t q = 0; q < Tc.size(); q++)
          if (Tc[q] == idx) {
            Tc[q] = Tc.back();
            Tc.pop_back();
            break;
          }
        const long last = P - 1;
        if (idx != last) {
          std::vector<int>& Tl = table(cell_of[last]);
          for (size_t q = 0; q < Tl.size(); q++)
            if (Tl[q] == last) {
              Tl[q] = (int)idx;
              break;
            }
          rsrc[idx] = rsrc[last];
          rnew[idx] = rnew[last];
          cell_of[idx] = cell_of[last];
        }
        rsrc.pop_back();
        rnew.pop_back();
        cell_of.pop_back();
        P--;
      }
    }
  }
  if (P > src_cap) return arg_error("population control: output capacity too small");
  for (long k = 0; k < P; k++) src[k] = rsrc[k];
  *p_new = P;
  bool changed = P != P0;
  for (long e = 0; e < E && !changed; e++) changed = edited[e];
  if (!changed) return 0;
  // new cell tables, owners
  std::vector<int> nptr(E + 1, 0), npid(P), nown(P);
  for (long e = 0; e < E; e++) {
    const int c = edited[e] ? (int)mod[e].size() : ptr[e + 1] - ptr[e];
    nptr[e + 1] = nptr[e] + c;
    for (int q = 0; q < c; q++) {
      const int i = edited[e] ? mod[e][q] : pid[ptr[e] + q];
      npid[nptr[e] + q] = i;
      nown[i] = (int)e;
    }
  }
  // device: gather coordinates and material index by source, then the created coordinates
  std::vector<long> hsrc(rsrc.begin(), rsrc.begin() + P);
  std::vector<int> cidx;
  std::vector<double> cval;
  for (long k = 0; k < P; k++)
    if (rnew[k] >= 0) {
      cidx.push_back((int)k);
      for (int d = 0; d < dim; d++) cval.push_back(newc[(size_t)rnew[k] * dim + d]);
    }
  long* dsrc = dalloc<long>(P);
  UWG_HIP(hipMemcpy(dsrc, hsrc.data(), sizeof(long) * P, hipMemcpyHostToDevice));
  double* ncoords = dalloc<double>(P * dim);
  int* nmat = dalloc<int>(P);
  launch_gather_particles(P, dim, dsrc, s->coords, s->mat, ncoords, nmat, 0);
  if (!cidx.empty()) {
    int* dci = dalloc<int>((long)cidx.size());
    double* dcv = dalloc<double>((long)cval.size());
    UWG_HIP(hipMemcpy(dci, cidx.data(), sizeof(int) * cidx.size(), hipMemcpyHostToDevice));
    UWG_HIP(hipMemcpy(dcv, cval.data(), sizeof(double) * cval.size(), hipMemcpyHostToDevice));
    launch_scatter_coords((long)cidx.size(), dim, dci, dcv, ncoords, 0);
    UWG_HIP(hipDeviceSynchronize());
    dfree(dci);
    dfree(dcv);
  }
  swarm_reserve(s, P);
  UWG_HIP(hipMemcpy(s->coords, ncoords, sizeof(double) * P * dim, hipMemcpyDeviceToDevice));
  UWG_HIP(hipMemcpy(s->mat, nmat, sizeof(int) * P, hipMemcpyDeviceToDevice));
  UWG_HIP(hipMemcpy(s->owner, nown.data(), sizeof(int) * P, hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(s->cell_ptr, nptr.data(), sizeof(int) * (E + 1), hipMemcpyHostToDevice));
  UWG_HIP(hipMemcpy(s->cell_pid, npid.data(), sizeof(int) * P, hipMemcpyHostToDevice));
  UWG_HIP(hipDeviceSynchronize());
  dfree(dsrc);
  dfree(ncoords);
  dfree(nmat);
  s->P = P;
  for (int v = 0; v < 8; v++) s->var_rows[v] = 0;  // swarm variables must be uploaded again
  s->state++;
  s->geom_state++;
  s->has_integration = false;
  return 0;
  API_END
}
