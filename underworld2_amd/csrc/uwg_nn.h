// Learned coarse-grid correction: a small convolutional network evaluated with the
// hand-written HIP kernels of nn_kernels.hip.
#pragma once
#include "uwg_internal.h"

namespace uwg {

struct NNModel;
NNModel* nn_create(const Geom& g, int channels, int depth, const float* params, int nparams);
void nn_destroy(NNModel* m);
// per-node input feature (log10 of viscosity over its geometric mean eta_ref) of the level the
// model acts on; outputs are scaled by rms(residual) / eta_ref
void nn_set_feature(NNModel* m, const float* feat_host, double eta_ref);
// x = NN(feature, r) on the nodes of level g (FP32); Dirichlet dofs of x are zero
// x = (accumulate ? x : 0) + omega d with omega = (r . Kd)/(Kd . Kd) clamped to [0, 2]
void nn_line_search(long n, const float* r, const float* kd, const float* d, float* x, bool accumulate,
                    hipStream_t s);
void nn_coarse_correction(NNModel* m, const Geom& g, const float* r, float* x, const uint8_t* bc, hipStream_t s);
long nn_param_count(NNModel* m);

struct NNDeleter {
  void operator()(NNModel* m) const { nn_destroy(m); }
};

// Learned Schur-complement stencil on the element grid (nn_kernels.hip)
struct SchurNet;
SchurNet* schur_net_create(int channels, int layers, const float* params, long nparams);
void schur_net_destroy(SchurNet* m);
// feat[9][E] from approxS[E] (device arrays)
void schur_net_features(const Geom& g, const double* approxS, float* feat, hipStream_t s);
// preconditioner coefficients a[27][E] (bfloat16 bits): z = sum_s a_s q_e+s approximates -S^-1 q
void schur_net_coefficients(SchurNet* m, const Geom& g, const double* approxS, uint16_t* a, hipStream_t s);
template <typename T>
void schur_stencil_apply(const Geom& g, const uint16_t* a, const T* q, T* z, hipStream_t s);
// append one training sample: features of approxS, r^ = D^-1/2 S e, e^ = D^1/2 e (device, FP64)
void schur_net_append(SchurNet* m, const Geom& g, const double* approxS, const double* rhat, const double* ehat);

struct SchurNetDeleter {
  void operator()(SchurNet* m) const { schur_net_destroy(m); }
};

}  // namespace uwg

// training handle of the C ABI
struct uwg_snn {
  uwg::SchurNet* m;
};
