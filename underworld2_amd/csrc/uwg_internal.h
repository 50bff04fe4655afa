// Internal declarations shared by the HIP translation units of libuwg.
// Layout conventions (all device memory, row-major):
//   node index  n = i + nx*(j + ny*k), nx = res[0]+1 (x fastest)
//   element     e = i + rx*(j + ry*k)
//   velocity-like vectors: [N][dim] interleaved (node-major, matches MeshVariable.data)
//   pressure-like vectors: [E]
//   per-element quadrature weights W: [Q][E] (structure of arrays), Q = 3^dim
//   Dirichlet mask: uint8 per node, bit i set when component i is constrained
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace uwg {

// last-error text, retrievable through uwg_last_error()
void set_error(const std::string& msg);

#define UWG_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) {                                                             \
      char _b[512];                                                                     \
      snprintf(_b, sizeof(_b), "HIP error %s at %s:%d: %s", hipGetErrorName(_e),        \
               __FILE__, __LINE__, hipGetErrorString(_e));                              \
      throw std::runtime_error(_b);                                                     \
    }                                                                                   \
  } while (0)

struct Geom {
  int dim;
  int res[3];      // elements per dim (res[2] = 1 in 2D)
  double lo[3], hi[3];
  long N, E;       // nodes, elements
  int nn[3];       // nodes per dim
  long Es;         // stride between the per-point arrays of the quadrature weights W[q][Es]:
                   // E padded so that the 3^dim arrays never sit a power of two apart
                   // (identical strides alias the same cache sets and memory channels)
};

inline Geom make_geom(int dim, const int* res, const double* lo, const double* hi) {
  Geom g{};
  g.dim = dim;
  for (int d = 0; d < 3; d++) {
    g.res[d] = d < dim ? res[d] : 1;
    g.lo[d] = d < dim ? lo[d] : 0.0;
    g.hi[d] = d < dim ? hi[d] : 1.0;
    g.nn[d] = d < dim ? res[d] + 1 : 1;
  }
  g.N = (long)g.nn[0] * g.nn[1] * g.nn[2];
  g.E = (long)g.res[0] * g.res[1] * g.res[2];
  g.Es = ((g.E + 63) / 64) * 64 + 64;
  return g;
}

// ---------------------------------------------------------------------------------------
// launch wrappers: swarm_kernels.hip (compiled with -ffp-contract=off, bit-exact paths)
// ---------------------------------------------------------------------------------------
void launch_layout_random(const Geom& g, int ppc, long seed, double* coords, hipStream_t s);
struct SobolDirs {
  unsigned int v[3][30];
};
void launch_layout_sobol(const Geom& g, int ppc, long seed, double* coords, hipStream_t s);
// out[i] = sum of flags[0..i-1] (relocate.hip, hipcub)
void scan_exclusive_flags(const int* flags, long long* out, long n, hipStream_t s);
void launch_mesh_coords(const Geom& g, double* out, hipStream_t s);
void launch_cell_of(const Geom& g, long P, const double* coords, int* owner, hipStream_t s);
void launch_coincident_map(const Geom& g, long P, const int* slot_cell, const int* cell_pid,
                           const double* coords, double* xi_slot, hipStream_t s);
void launch_slot_cell(long E, const int* cell_ptr, int* slot_cell, hipStream_t s);
// g: the element grid; order 1: vel on the element grid's nodes, order 2: on the Q2 vertex grid
void launch_advect_rk2(const Geom& g, long P, const double* vel, double* coords, double dt,
                       int* d_counts /*[2]: bad, fallback*/, hipStream_t s, int order = 1);
void launch_interpolate(const Geom& g, long P, const double* vel, const double* x, double* v,
                        int* ok, hipStream_t s);
void launch_node_speed_max(const Geom& g, const double* vel, double* partial, int nblk, hipStream_t s);
void launch_gather_particles(long P, int dim, const long* src, const double* coords, const int* mat, double* ncoords,
                             int* nmat, hipStream_t s);
void launch_scatter_coords(long n, int dim, const int* idx, const double* val, double* coords, hipStream_t s);

// ---------------------------------------------------------------------------------------
// launch wrappers: fem_kernels.hip (solver / assembly kernels)
// ---------------------------------------------------------------------------------------
struct MatTable {   // material tables for the moment kernel
  int nmat;
  const double* eta;    // [nmat]
  const double* force;  // [nmat][dim]
};
void launch_moments(const Geom& g, const int* cell_ptr, const int* cell_pid, const double* xi_slot,
                    const double* w_slot, const int* mat, const double* eta_tab, const double* force_tab,
                    const double* eta_p, const double* force_p, double* W, double* fe, double* approxS,
                    const double* Ge /*host, 2^d*d values*/, int nmat, hipStream_t s);
void launch_node_gather_fe(const Geom& g, const double* fe, const uint8_t* bc, double* f, hipStream_t s);
void launch_gauss_assemble(const Geom& g, const double* eta_g, const double* force_g, double* W, double* fe,
                           double* approxS, const double* Ge, hipStream_t s);
void launch_restrict_W(const Geom& fine, const Geom& coarse, const double* Wf, double* Wc, hipStream_t s);
void launch_diag(const Geom& g, const double* W, const uint8_t* bc, double* dinv, hipStream_t s);
// K applies.  Wbf (FP32 instance, 3D only): the weights as bfloat16 pairs [14][Es] (see
// launch_pack_w3_bf16); when given, the apply reads them instead of W.
template <typename T>
void launch_apply_K(const Geom& g, const T* W, const T* x, T* y, const uint8_t* bc, hipStream_t s,
                    const uint32_t* Wbf = nullptr);
template <typename T>
void launch_residual_K(const Geom& g, const T* W, const T* x, const T* b, T* r, const uint8_t* bc, hipStream_t s,
                       const uint32_t* Wbf = nullptr);
template <typename T>
void launch_cheb_step_K(const Geom& g, const T* W, const T* x, T* xn, const T* b, const T* dinv, T* d, T a, T c,
                        bool first, const uint8_t* bc, hipStream_t s, const uint32_t* Wbf = nullptr);
// FP64 weights [27][Es] -> bfloat16 pairs [14][Es]: word p = bf16(W_2p) | bf16(W_2p+1) << 16
void launch_pack_w3_bf16(long Es, const double* W, uint32_t* Wb, hipStream_t s);
template <typename T>
void launch_apply_G(const Geom& g, const double* Ge, const T* p, const uint8_t* bc, T* y, hipStream_t s);
template <typename T>
void launch_apply_GT(const Geom& g, const double* Ge, const T* u, const uint8_t* bc, T* q, hipStream_t s);
template <typename T>
void launch_prolong_add(const Geom& fine, const Geom& coarse, const T* xc, T* xf, const uint8_t* bcf,
                        hipStream_t s);
template <typename T>
void launch_restrict2(const Geom& fine, const Geom& coarse, const T* rf, T* rc, const uint8_t* bcc,
                      const uint8_t* bcf, hipStream_t s);
void set_apply_chunk(int kz3, int kz2);
long dvc_scratch_bytes(int dim, int num, int max_np, int cap);
bool dvc_wave_supported(int dim, int num);
long dvc_wave_bound_shorts(int cap);
// fn_kernels.hip: per-particle function programs and general field interpolation
void launch_fn_eval(const Geom& g, long P, const double* coords, const int* mat, const void* prog, int nins,
                    const double* consts, const double* const* fields, const int* field_nc, int nfields,
                    const double* const* vars, const int* var_nc, int nvars, int out_eta, int out_force, double* eta,
                    double* force, int* err, hipStream_t s);
void launch_field_interpolate(const Geom& g, long P, const double* f, int nc, const double* x, double* v, int* ok,
                              hipStream_t s);
void launch_dvc_wave(int dim, int num, long E, const int* cell_ptr, double* xi_slot, double* w_slot, short* gbound,
                     int nblk, int cap, int* flag, hipStream_t s);
void launch_dvc_list(int dim, int num, long E0, long nE, const int* elist, const int* cell_ptr,
                     double* xi_slot, double* w_slot, void* scratch, int max_np, int cap, int* d_flag,
                     hipStream_t s);
// relocation (relocate.hip): rebuild per-cell tables after owners change; returns the number
// of particles outside the domain
struct RelocWork;
RelocWork* reloc_create(long P, long E);
void reloc_destroy(RelocWork* w);
long relocate(const Geom& g, long P, const double* coords, int* owner, int* cell_ptr, int* cell_pid,
              RelocWork* w, hipStream_t s);
void launch_bc_mask_from_walls(const Geom& g, uint8_t* bc, int wallbits, hipStream_t s);
void launch_inject_bc(const Geom& f, const Geom& c, const uint8_t* fbc, uint8_t* cbc, hipStream_t s);
void launch_fill_hash(long n, double* x, hipStream_t s);
void launch_bc_split(const Geom& g, const uint8_t* bc, const double* v, double* ubc, int* flag, hipStream_t s);
void launch_zero_bc(const Geom& g, const uint8_t* bc, double* v, hipStream_t s);
template <typename T> void launch_neg_recip(long n, const double* sv, T* out, hipStream_t s);

// q2_kernels.hip: the Q2/dPc1 element pair.  g is the element grid; vertex vectors live on
// the (2 res + 1)^dim grid, pressure vectors are [E][dim+1], weights W are [E][5^dim].
void launch_q2_weights(const Geom& g, const int* cell_ptr, const int* cell_pid, const double* xi_slot,
                       const double* w_slot, const int* mat, int nmat, const double* eta_tab, const double* force_tab,
                       const double* eta_p, const double* force_p, const double* eta_g, const double* force_g,
                       double* W, double* fe, hipStream_t s);
void launch_q2_gather_f(const Geom& g, const double* fe, const uint8_t* bc, double* f, hipStream_t s);
template <typename T>
void launch_q2_apply(const Geom& g, const T* W, const T* x, T* y, const uint8_t* bc, hipStream_t s,
                     const T* Dm = nullptr);  // Dm: [32][384] derivative matrix -> matrix-core path (3D)
void q2_mfma_matrix(const Geom& g, std::vector<double>& D);
void set_q2_mfma(int on);
void set_q2_apply_kernel(int kind);
void launch_q2_diag(const Geom& g, const double* W, const uint8_t* bc, double* dinv, hipStream_t s);
template <typename T>
void launch_q2_apply_G(const Geom& g, const double* Gr, const T* p, const uint8_t* bc, T* y, hipStream_t s);
template <typename T>
void launch_q2_apply_GT(const Geom& g, const double* Gr, const T* u, const uint8_t* bc, T* q, hipStream_t s);
void launch_q2_approxS(const Geom& g, const double* W, const double* Gr, double* S, double* Sinv, int* bad,
                       hipStream_t s);
template <typename T>
void launch_q2_schur_prec(long E, int np, const T* Sinv, const T* r, T* z, hipStream_t s);
void launch_q2_project_W(const Geom& g, const double* W, double* W1, hipStream_t s);
void q2_element_G(const Geom& g, std::vector<double>& Gr);
void launch_q2_interpolate(const Geom& g, long P, const double* f, int nc, const double* x, double* v, int* ok,
                           hipStream_t s);

// advdiff_kernels.hip: SUPG advection-diffusion of a nodal scalar (Q1 mesh)
void launch_supg_elem(const Geom& g, const double* phi, const double* phidot, const double* vel, const double* kap,
                      const double* src, int upwind, double* res_e, hipStream_t s);
void launch_supg_mass(const Geom& g, double* me, double* M, hipStream_t s);
void launch_supg_predict(long N, const uint8_t* bc, const double* phi, const double* phidot, double a, double* phi_v,
                         double* pdot_v, hipStream_t s);
void launch_supg_correct(const Geom& g, const uint8_t* bc, const double* res_e, const double* M, double dtg,
                         double* phi_v, double* pdot_v, double* phi, double* phidot, hipStream_t s);
void launch_supg_adv_dt(const Geom& g, const double* vel, const double* sep, double* partial, int nblk,
                        hipStream_t s);

// vector kernels (T = double or float)
template <typename T> void vk_fill(long n, T v, T* x, hipStream_t s);
template <typename T> void vk_copy(long n, const T* x, T* y, hipStream_t s);
template <typename T> void vk_axpy(long n, T a, const T* x, T* y, hipStream_t s);           // y += a x
template <typename T> void vk_axpby(long n, T a, const T* x, T b, T* y, hipStream_t s);     // y = a x + b y
template <typename T> void vk_scale(long n, T a, const T* x, T* y, hipStream_t s);          // y = a x (y may alias x)
template <typename T> void vk_xmy_scale(long n, const T* d, const T* x, T* y, hipStream_t s); // y = d .* x
void vk_d2f(long n, double scale, const double* x, float* y, hipStream_t s);
void vk_scale_inv_sqrt(long n, const double* ss, const double* x, double* y, hipStream_t s);  // y = x / sqrt(ss[0])
void vk_f2d_axpy(long n, double scale, const float* x, double* y, hipStream_t s);  // y += scale*x
void vk_f2d(long n, double scale, const float* x, double* y, hipStream_t s);        // y = scale*x
void vk_d2f_vec(long n, const double* x, float* y, hipStream_t s);
// y = sum_j c[j] * X[j]  (+ y if accumulate)   X: array of device pointers (device memory)
template <typename T>
void vk_multi_axpy(long n, int k, const T* const* X, const double* c, T* y, hipStream_t s);
// partial sums for dot products: out[j] = <x, Y[j]> for j < k ; deterministic
template <typename T>
void vk_multi_dot(long n, int k, const T* x, const T* const* Y, double* partial, double* out, hipStream_t s);
// one classical Gram-Schmidt step with device-resident coefficients: h[0..k) = V^T w,
// w -= sum h_j V_j, h[k] = |w|^2 (bit-identical to a separate dot), v_next = w / |w| (skipped
// when |w| = 0).  The basis V (and v_next) may be stored in a narrower type TV than w; every
// product and sum is in double.  partial holds 32 x vk_reduce_blocks() doubles
template <typename TW, typename TV>
void vk_gs_step(long n, int k, const TV* const* V, TW* w, TV* v_next, double* h, double* partial, hipStream_t s);
int vk_reduce_blocks();
// chebyshev helper: d = a*d + b*(dinv .* r) ; x += d  (fused)
template <typename T>
void vk_cheb_update(long n, T a, T b, const T* dinv, const T* r, T* d, T* x, hipStream_t s);
template <typename T>
void vk_cheb_first(long n, T c, const T* dinv, const T* b, T* d, T* x, hipStream_t s);
// r = b - r  (r holds A x on input)
template <typename T> void vk_residual(long n, const T* b, T* r, hipStream_t s);
// dense coarse solve y = Ainv x for a small dense inverse (double)
void launch_dense_mv(int n, const double* A, const double* x, double* y, hipStream_t s);
// dense inverse of the coarsest level's operator (identity Dirichlet rows), n = N dim; work:
// a second n x n buffer
void launch_coarse_inverse(const Geom& g, const double* W, const uint8_t* bc, double* Ainv, double* work, int* flag,
                           hipStream_t s);
// max-reduction partials -> value
double finish_max(const double* d_partial, int nblk);

}  // namespace uwg
