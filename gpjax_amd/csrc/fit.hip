// One optimiser step of gpjax.fit on the device, in one launch.
//
// gpjax/fit.py:133-180 trains on unconstrained coordinates: the parameters go
// through the inverse bijection once (fit.py:136-138), every step maps them
// back (softplus for "positive" / "non_negative", identity for "real":
// DEFAULT_BIJECTION, gpjax/parameters.py:140-146), takes value_and_grad of the
// objective and applies an optax update (fit.py:160-162).  With optax's adam
// defaults (b1 = 0.9, b2 = 0.999, eps = 1e-8, eps_root = 0; Kingma & Ba 2015)
// the update of one coordinate u with loss gradient g (w.r.t. the constrained
// value p = f(u)) is
//   g_u = g f'(u)                       (f' = sigmoid for softplus, 1 otherwise)
//   m <- b1 m + (1 - b1) g_u,   v <- b2 v + (1 - b2) g_u^2,   t <- t + 1
//   u <- u - lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)
// and the next step's constrained value is p = f(u).  This kernel does all of
// it for every coordinate of every trainable parameter, writes p in place (the
// caller's parameter tensors are views of `value`) and records the step's loss
// in history[t - 1].  One workgroup: the parameter count of a GP model is
// small (lengthscales, variance, noise, mean, inducing inputs), and the step
// counter is read by every lane before lane 0 advances it.
#include "gpk_internal.h"

namespace gpk {
namespace {

__device__ __forceinline__ double softplus_d(double u) {
  // log(1 + e^u) without overflow: max(u, 0) + log1p(e^-|u|)
  return fmax(u, 0.0) + log1p(exp(-fabs(u)));
}

__device__ __forceinline__ double sigmoid_d(double u) {
  if (u >= 0.0) return 1.0 / (1.0 + exp(-u));
  const double e = exp(u);
  return e / (1.0 + e);
}

constexpr int ADAM_THREADS = 256;

__global__ __launch_bounds__(ADAM_THREADS) void adam_step_kernel(
    int64_t count, const double* __restrict__ grad, const uint8_t* __restrict__ softplus_mask,
    double* __restrict__ raw, double* __restrict__ m, double* __restrict__ v,
    double* __restrict__ step, double lr, double b1, double b2, double eps,
    double* __restrict__ value, const double* __restrict__ loss, double* __restrict__ history,
    int64_t history_len) {
  const double t = step[0] + 1.0;
  const double c1 = 1.0 - pow(b1, t);
  const double c2 = 1.0 - pow(b2, t);
  for (int64_t i = threadIdx.x; i < count; i += ADAM_THREADS) {
    const double u = raw[i];
    const bool sp = softplus_mask[i] != 0;
    const double g = sp ? grad[i] * sigmoid_d(u) : grad[i];
    const double mi = b1 * m[i] + (1.0 - b1) * g;
    const double vi = b2 * v[i] + (1.0 - b2) * (g * g);
    m[i] = mi;
    v[i] = vi;
    const double un = u - lr * (mi / c1) / (sqrt(vi / c2) + eps);
    raw[i] = un;
    value[i] = sp ? softplus_d(un) : un;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t k = (int64_t)t - 1;
    if (history && loss && k >= 0 && k < history_len) history[k] = loss[0];
    step[0] = t;
  }
}

}  // namespace
}  // namespace gpk

using namespace gpk;

extern "C" int gpk_adam_step(int64_t count, const double* grad, const uint8_t* softplus_mask,
                             double* raw, double* m, double* v, double* step, double lr,
                             double b1, double b2, double eps, double* value, const double* loss,
                             double* history, int64_t history_len, void* stream) {
  if (count < 1 || !grad || !softplus_mask || !raw || !m || !v || !step || !value) {
    set_error("adam_step: need count >= 1 and non-null grad/mask/raw/m/v/step/value");
    return GPK_ERR_BAD_ARG;
  }
  if (!(lr > 0.0) || !(b1 >= 0.0 && b1 < 1.0) || !(b2 >= 0.0 && b2 < 1.0) || !(eps >= 0.0)) {
    set_error("adam_step: need lr > 0, 0 <= b1, b2 < 1, eps >= 0");
    return GPK_ERR_BAD_ARG;
  }
  if (history && history_len < 0) {
    set_error("adam_step: negative history length");
    return GPK_ERR_BAD_ARG;
  }
  hipLaunchKernelGGL(adam_step_kernel, dim3(1), dim3(ADAM_THREADS), 0, (hipStream_t)stream,
                     count, grad, softplus_mask, raw, m, v, step, lr, b1, b2, eps, value, loss,
                     history, history_len);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}
