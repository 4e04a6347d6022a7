/* qcml.h -- C ABI of the learned initial-guess network (no reference counterpart: the reference
 * seeds its SCF with the core guess or SAD only, ChargeDensity/Imp/Seed.C:64-77).
 *
 * A 3-layer perceptron  X -> SiLU(X W1 + b1) -> SiLU(H1 W2 + b2) -> H2 W3 + b3  with BF16 weights
 * and FP32 accumulation, trained and evaluated by the library's own HIP kernels (bf16 MFMA GEMMs,
 * bias/activation, MSE, Adam).  FP32 master weights and Adam moments stay on the device; the BF16
 * copies the GEMMs read are refreshed after every update.  Row-major everywhere: X is B x nin, W1 is
 * nin x width, W2 width x width, W3 width x nout, outputs B x nout.  One net = one host thread at a
 * time; every call is asynchronous on `stream` exactly as given (null = the default stream, so a
 * caller on the default stream is ordered with the net) unless it returns host data.
 */
#ifndef QCML_H
#define QCML_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qcml_net qcml_net;

/* Allocates the net for batches up to max_batch; weights from a seeded He-uniform draw, biases 0. */
int qcml_create(int32_t nin, int32_t width, int32_t nout, int32_t max_batch, uint64_t seed, int32_t device,
                qcml_net** out);

/* Parameter count and the FP32 master parameters in the order W1, b1, W2, b2, W3, b3 (host). */
int64_t qcml_nparams(const qcml_net* net);
int qcml_get_params(qcml_net* net, float* host);
int qcml_set_params(qcml_net* net, const float* host);  /* also refreshes the BF16 copies, resets Adam */

/* Y = net(X) for B <= max_batch rows; device pointers. */
int qcml_forward(qcml_net* net, const float* dX, int32_t B, float* dY, void* stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean squared error against dT.  If loss
 * is non-null the call synchronises and stores the batch loss (before the update) there. */
int qcml_train_step(qcml_net* net, const float* dX, const float* dT, int32_t B, float lr, float* loss, void* stream);

/* Gradient of the mean squared error w.r.t. every parameter (same order as qcml_get_params), to the
 * host, without updating; synchronous.  loss (nullable) receives the loss. */
int qcml_gradients(qcml_net* net, const float* dX, const float* dT, int32_t B, float* host_grad, float* loss);

/* The GEMM the net runs, exposed for tests: C (M x N, FP32, row stride ldc) = alpha op(A) op(B) + beta C
 * with A(m,k) at A[m*sam + k*sak] and B(k,n) at B[k*sbk + n*sbn]; a_bf16/b_bf16 select BF16 (raw
 * 16-bit) or FP32 operands (rounded to BF16 on load).  Device pointers. */
int qcml_gemm(int32_t M, int32_t N, int32_t K, const void* A, int32_t a_bf16, int64_t sam, int64_t sak,
              const void* B, int32_t b_bf16, int64_t sbk, int64_t sbn, float* C, int64_t ldc, float alpha,
              float beta, void* stream);

void qcml_destroy(qcml_net* net);

#ifdef __cplusplus
}
#endif

#endif
