// mlp.h - layout and host entry points of the learned proposal's network.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#define MLP_IN 12
#define MLP_H 64
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_OFF_W1 + MLP_H * MLP_IN)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_H)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_H)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_H)
#define MLP_NPARAMS (MLP_OFF_B3 + 1)

#define MLP_MAX_EPOCHS 64
struct MlpOrderKeys {
  uint32_t sample[4], epoch[4 * MLP_MAX_EPOCHS];
  int half_n, half_m;
};

// floats of block-partial scratch a batch of n samples needs
size_t mlp_part_floats(int n);
// out[i] = sigmoid(net(X[i]))
hipError_t mlp_forward(const float *dX, int n, const float *dP, float *dOut, hipStream_t st);
// mean BCE gradient and loss over samples idx[b0 .. b0+n) (rows b0.. when idx is null)
hipError_t mlp_gradient(const float *dX, const float *dY, const int *dIdx, int b0, int n, const float *dP,
                        float *dPart, float *dGrad, float *dLoss, hipStream_t st);
// mean BCE loss alone over rows 0..n-1 (dPart: (n + 255) / 256 floats)
hipError_t mlp_loss(const float *dX, const float *dY, int n, const float *dP, float *dPart, float *dLoss,
                    hipStream_t st);
// the same gradient followed by one Adam step (beta 0.9 / 0.999, eps 1e-8)
hipError_t mlp_train_step(const float *dX, const float *dY, const int *dIdx, int b0, int n, float *dP, float *dPart,
                          float *dGrad, float *dLoss, float *dM, float *dV, float lr, int step, hipStream_t st);
// Training order of `epochs` epochs over a sample of m of the n rows (keyed
// by seed): out[e * m + i] = the i-th row epoch e visits.  n <= 2^30,
// epochs <= MLP_MAX_EPOCHS.
hipError_t mlp_order(uint64_t seed, long long n, long long m, int epochs, int *dOut, hipStream_t st);
// the same order computed on the host (false for arguments mlp_order refuses)
bool mlp_order_host(uint64_t seed, long long n, long long m, int epochs, int *out);
// Y[i] = hits[i] * inv_den[set of i], sets starting at setbase[] (ascending)
hipError_t mlp_labels(const unsigned *dHits, const float *dInvDen, const long long *dSetBase, int nsets, long long n,
                      float *dY, hipStream_t st);
