// mlp.hip - the learned proposal's network: 12 FP32 features -> 64 -> 64 -> 1
// (ReLU, ReLU, sigmoid), trained with binary cross-entropy on soft labels by
// Adam.  Hand-written HIP kernels; every reduction runs in a fixed order, so a
// training run is reproducible bit for bit.
//
// Parameter layout (MLP_NPARAMS floats):
//   W1 [64][12] at 0, b1 [64] at 768, W2 [64][64] at 832, b2 [64] at 4928,
//   w3 [64] at 4992, b3 at 5056.
//
// Training step = k_mlp_part (one workgroup per 64 samples: forward and
// backward with the activations in LDS, then that block's gradient sums) +
// k_mlp_reduce (sums the block partials in block order, divides by the batch
// size and, in training, applies the Adam update in the same pass).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "mlp.h"

#define MB 64          // samples per workgroup
#define HS (MLP_H + 1) // padded LDS row: lanes of one wave hit distinct banks
#define XS (MLP_IN + 1)

__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }
__device__ __forceinline__ float sigmoidf_(float z) { return 1.f / (1.f + __expf(-z)); }

// The logit of one sample (features x), the parameters read at wave-uniform
// addresses (every lane takes the same weight at the same step: scalar loads,
// no LDS copy), hidden layer 1 in registers; sums in the order of the
// parameter layout.
__device__ __forceinline__ float mlp_logit(const float *__restrict__ P, const float *x) {
  float a1[MLP_H];
#pragma unroll
  for (int j = 0; j < MLP_H; j++) {
    float s = P[MLP_OFF_B1 + j];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) s += P[MLP_OFF_W1 + j * MLP_IN + i] * x[i];
    a1[j] = relu(s);
  }
  float z = P[MLP_OFF_B3];
#pragma unroll 2
  for (int k = 0; k < MLP_H; k++) {
    float s = P[MLP_OFF_B2 + k];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) s += P[MLP_OFF_W2 + k * MLP_H + j] * a1[j];
    z += P[MLP_OFF_W3 + k] * relu(s);
  }
  return z;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Inference on the matrix cores.  v_mfma_f32_16x16x4f32 takes A (16 x 4) and
// B (4 x 16) and accumulates D (16 x 16): lane l holds A[l % 16][l / 16],
// B[l / 16][l % 16] and, in register r, D[4 (l / 16) + r][l % 16].  Each wave
// takes 16 samples at a time and computes the layers transposed, units x
// samples (D = W . H), so the 4 x 4 values a lane holds of layer 1's output
// (units 16m + 4 (l / 16) + r of sample l % 16) are already layer 2's B
// operand once contraction step (m, r) lets its index k = l / 16 stand for
// input unit 16m + 4k + r; A then takes W2's columns in that same order.  The
// weights stay in registers for every tile of the wave.  The logit is each
// lane's 16 terms of w3 . h2, then the sums over the four lane groups (xor 16,
// xor 32): a fixed order, so the output does not depend on the launch shape.
__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ X, int n, const float *__restrict__ P,
                                                     float *__restrict__ out) {
  const int l = threadIdx.x & 63, c = l & 15, g = l >> 4;
  float w1[4][3], b1[4][4], w2[4][4][4], b2[4][4], w3[4][4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
#pragma unroll
    for (int s = 0; s < 3; s++) w1[m][s] = P[MLP_OFF_W1 + (16 * m + c) * MLP_IN + 4 * s + g];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      b1[m][r] = P[MLP_OFF_B1 + 16 * m + 4 * g + r];
      b2[m][r] = P[MLP_OFF_B2 + 16 * m + 4 * g + r];
      w3[m][r] = P[MLP_OFF_W3 + 16 * m + 4 * g + r];
    }
  }
#pragma unroll
  for (int o = 0; o < 4; o++)
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int r = 0; r < 4; r++) w2[o][m][r] = P[MLP_OFF_W2 + (16 * o + c) * MLP_H + 16 * m + 4 * g + r];
  const float b3 = P[MLP_OFF_B3];
  const int ntile = (n + 15) >> 4;
  for (int t = blockIdx.x * 4 + (threadIdx.x >> 6); t < ntile; t += gridDim.x * 4) {
    const int smp = t * 16 + c;
    const bool live = smp < n;
    float xb[3];
#pragma unroll
    for (int s = 0; s < 3; s++) xb[s] = live ? X[(size_t)smp * MLP_IN + 4 * s + g] : 0.f;
    f32x4 h1[4], h2[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
      f32x4 a = {b1[m][0], b1[m][1], b1[m][2], b1[m][3]};
#pragma unroll
      for (int s = 0; s < 3; s++) a = __builtin_amdgcn_mfma_f32_16x16x4f32(w1[m][s], xb[s], a, 0, 0, 0);
      h1[m] = a;
      h2[m] = f32x4{b2[m][0], b2[m][1], b2[m][2], b2[m][3]};
    }
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float hb = relu(h1[m][r]);
#pragma unroll
        for (int o = 0; o < 4; o++) h2[o] = __builtin_amdgcn_mfma_f32_16x16x4f32(w2[o][m][r], hb, h2[o], 0, 0, 0);
      }
    float p = 0.f;
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
      for (int r = 0; r < 4; r++) p += w3[o][r] * relu(h2[o][r]);
    p += __shfl_xor(p, 16);
    p += __shfl_xor(p, 32);
    if (g == 0 && live) out[smp] = sigmoidf_(b3 + p);
  }
}

// One workgroup = MB samples (rows idx[b0 + b] of X/Y when idx is given, else
// rows b0 + b), 256 threads = 4 waves; wave q owns hidden units 16q..16q+15 in
// the forward and backward passes (lane = sample, 16 sums in registers while
// the lane's activations are read once each), then the block's gradient sums
// over its samples in sample order: W2's 64 x 64 as a 4 x 4 tile per thread,
// the rest one output per thread at a time.  part[blk][o] receives the sums.
__global__ __launch_bounds__(256) void k_mlp_part(const float *X, const float *Y, const int *idx, int b0, int n,
                                                  const float *__restrict__ P, float *part) {
  __shared__ float H1[MB * HS], H2[MB * HS], D1[MB * HS], D2[MB * HS];
  __shared__ float Xs[MB * XS];
  __shared__ float ZQ[4 * MB];
  __shared__ float DZ[MB], LS[MB];
  const int tid = threadIdx.x;
  const int lane = tid & 63, q = tid >> 6;
  const int blk = blockIdx.x;
  const int b = lane;                        // sample within the block
  const int gb = blk * MB + b;               // sample within the batch
  const bool live = gb < n;
  const int row = live ? (idx ? idx[b0 + gb] : b0 + gb) : 0;
  if (q == 0) {
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) Xs[b * XS + i] = live ? X[(size_t)row * MLP_IN + i] : 0.f;
  }
  __syncthreads();
  float acc[16];
  // layer 1: this wave's 16 units of h1
  {
    float x[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) x[i] = Xs[b * XS + i];
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
      const int j = q * 16 + jj;
      float s = P[MLP_OFF_B1 + j];
#pragma unroll
      for (int i = 0; i < MLP_IN; i++) s += P[MLP_OFF_W1 + j * MLP_IN + i] * x[i];
      H1[b * HS + j] = live ? relu(s) : 0.f;
    }
  }
  __syncthreads();
  // layer 2: 16 units of h2 (each b2 + the sum over j in order) and their
  // share of the logit
#pragma unroll
  for (int kk = 0; kk < 16; kk++) acc[kk] = P[MLP_OFF_B2 + q * 16 + kk];
  for (int j = 0; j < MLP_H; j++) {
    const float h = H1[b * HS + j];
#pragma unroll
    for (int kk = 0; kk < 16; kk++) acc[kk] += P[MLP_OFF_W2 + (q * 16 + kk) * MLP_H + j] * h;
  }
  float zq = 0.f;
#pragma unroll
  for (int kk = 0; kk < 16; kk++) {
    const int k = q * 16 + kk;
    const float h = live ? relu(acc[kk]) : 0.f;
    H2[b * HS + k] = h;
    zq += P[MLP_OFF_W3 + k] * h;
  }
  ZQ[q * MB + b] = zq;
  __syncthreads();
  if (q == 0) {
    const float z = P[MLP_OFF_B3] + ((ZQ[b] + ZQ[MB + b]) + (ZQ[2 * MB + b] + ZQ[3 * MB + b]));
    const float y = live ? Y[row] : 0.f;
    // BCE(sigmoid(z), y) = softplus(z) - y z ; d/dz = sigmoid(z) - y
    const float sp = z > 0.f ? z + log1pf(__expf(-z)) : log1pf(__expf(z));
    LS[b] = live ? sp - y * z : 0.f;
    DZ[b] = live ? sigmoidf_(z) - y : 0.f;
  }
  __syncthreads();
  // backward to layer 1: d2[k] = [h2_k > 0] dz w3[k] (wave 0 keeps it in D2),
  // d1[j] = [h1_j > 0] sum_k W2[k][j] d2[k], the sum over k in order
  {
    const float dz = DZ[b];
#pragma unroll
    for (int jj = 0; jj < 16; jj++) acc[jj] = 0.f;
    for (int k = 0; k < MLP_H; k++) {
      const float d2 = H2[b * HS + k] > 0.f ? dz * P[MLP_OFF_W3 + k] : 0.f;
      if (q == 0) D2[b * HS + k] = d2;
#pragma unroll
      for (int jj = 0; jj < 16; jj++) acc[jj] += P[MLP_OFF_W2 + k * MLP_H + q * 16 + jj] * d2;
    }
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
      const int j = q * 16 + jj;
      D1[b * HS + j] = H1[b * HS + j] > 0.f ? acc[jj] : 0.f;
    }
  }
  __syncthreads();
  float *out = part + (size_t)blk * (MLP_NPARAMS + 1);
  // W2[k][j] = sum_t d2[t][k] h1[t][j]: rows 4 (tid / 16).., columns 4 (tid % 16)..
  {
    const int k0 = (tid >> 4) * 4, j0 = (tid & 15) * 4;
    float g[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) g[u][v] = 0.f;
    for (int t = 0; t < MB; t++) {
      float d[4], h[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        d[u] = D2[t * HS + k0 + u];
        h[u] = H1[t * HS + j0 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) g[u][v] += d[u] * h[v];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) out[MLP_OFF_W2 + (k0 + u) * MLP_H + j0 + v] = g[u][v];
  }
  // the other outputs: W1 and b1, then b2.. and the loss (index NPARAMS)
  constexpr int kRest = MLP_OFF_W2 + (MLP_NPARAMS + 1 - MLP_OFF_B2);
  for (int r = tid; r < kRest; r += 256) {
    const int o = r < MLP_OFF_W2 ? r : r - MLP_OFF_W2 + MLP_OFF_B2;
    float s = 0.f;
    if (o < MLP_OFF_B1) {  // W1[j][i]
      const int j = o / MLP_IN, i = o % MLP_IN;
      for (int t = 0; t < MB; t++) s += D1[t * HS + j] * Xs[t * XS + i];
    } else if (o < MLP_OFF_W2) {  // b1[j]
      const int j = o - MLP_OFF_B1;
      for (int t = 0; t < MB; t++) s += D1[t * HS + j];
    } else if (o < MLP_OFF_W3) {  // b2[k]
      const int k = o - MLP_OFF_B2;
      for (int t = 0; t < MB; t++) s += D2[t * HS + k];
    } else if (o < MLP_OFF_B3) {  // w3[k]
      const int k = o - MLP_OFF_W3;
      for (int t = 0; t < MB; t++) s += DZ[t] * H2[t * HS + k];
    } else if (o == MLP_OFF_B3) {
      for (int t = 0; t < MB; t++) s += DZ[t];
    } else {  // o == MLP_NPARAMS: loss
      for (int t = 0; t < MB; t++) s += LS[t];
    }
    out[o] = s;
  }
}

// Loss only (no gradient): one thread per sample of rows 0..n-1, the block's
// 256 losses summed in sample order into part[blk].
__global__ __launch_bounds__(256) void k_mlp_loss_part(const float *__restrict__ X, const float *__restrict__ Y, int n,
                                                       const float *__restrict__ P, float *part) {
  __shared__ float L[256];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (b < n) {
    float x[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) x[i] = X[(size_t)b * MLP_IN + i];
    const float z = mlp_logit(P, x);
    // BCE(sigmoid(z), y) = softplus(z) - y z, as k_mlp_part
    const float y = Y[b];
    const float sp = z > 0.f ? z + log1pf(__expf(-z)) : log1pf(__expf(z));
    l = sp - y * z;
  }
  L[threadIdx.x] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int t = 0; t < 256; t++) s += L[t];
    part[blockIdx.x] = s;
  }
}
__global__ void k_mlp_loss_reduce(const float *part, int nblk, int n, float *loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.f;
  for (int k = 0; k < nblk; k++) s += part[k];
  *loss = s / (float)n;
}

// grad[q] = (sum over blocks of part[blk][q]) / n in block order; loss likewise.
// With adam != 0 the same thread then applies Adam to parameter q.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *part, int nblk, int n, float *grad, float *loss,
                                                    int adam, float *params, float *m, float *v, float lr, float b1,
                                                    float b2, float eps, int step) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > MLP_NPARAMS) return;
  float s = 0.f;
  for (int k = 0; k < nblk; k++) s += part[(size_t)k * (MLP_NPARAMS + 1) + q];
  const float g = s / (float)n;
  if (q == MLP_NPARAMS) {
    *loss = g;
    return;
  }
  grad[q] = g;
  if (!adam) return;
  const float mq = b1 * m[q] + (1.f - b1) * g;
  const float vq = b2 * v[q] + (1.f - b2) * g * g;
  m[q] = mq;
  v[q] = vq;
  const float mh = mq / (1.f - powf(b1, (float)step));
  const float vh = vq / (1.f - powf(b2, (float)step));
  params[q] -= lr * mh / (sqrtf(vh) + eps);
}

// Y[i] = hits[i] / den[set of i] (soft labels from the logged edge counts)
__global__ void k_mlp_labels(const unsigned *hits, const float *inv_den, const long long *setbase, int nsets,
                             long long n, float *Y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = nsets - 1;  // last set with setbase <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (setbase[mid] <= i) lo = mid; else hi = mid - 1;
  }
  Y[i] = (float)hits[i] * inv_den[lo];
}

// Training order.  Epoch e visits the m sampled rows in the order
// perm_e(0), perm_e(1), ..., and sampled row j is row perm_s(j) of the n
// logged rows, where perm_s (over [0, n)) and perm_e (over [0, m)) are keyed
// pseudo-random bijections: a four-round balanced Feistel network on the
// smallest even bit width covering the range, with cycle walking back into
// the range (Black & Rogaway, "Ciphers with arbitrary finite domains", 2002).
// The round function is MurmurHash3's 32-bit finaliser of the half and the
// round key.
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__host__ __device__ uint32_t feistel_perm(uint32_t x, uint32_t range, int half, const uint32_t *key) {
  const uint32_t mask = (1u << half) - 1u;
  do {
    uint32_t l = x >> half, r = x & mask;
    for (int k = 0; k < 4; k++) {
      const uint32_t t = l ^ (fmix32(r ^ key[k]) & mask);
      l = r;
      r = t;
    }
    x = (l << half) | r;
  } while (x >= range);
  return x;
}
__global__ void k_mlp_order(MlpOrderKeys keys, uint32_t n, uint32_t m, int epochs, int *out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)m * epochs) return;
  const int e = (int)(i / m);
  const uint32_t j = feistel_perm((uint32_t)(i - (long long)e * m), m, keys.half_m, keys.epoch + 4 * e);
  out[i] = (int)feistel_perm(j, n, keys.half_n, keys.sample);
}

// ---- host helpers (device pointers; stream-ordered) -------------------------
static int half_bits(uint32_t range) {
  int b = 2;
  while (b < 32 && ((uint64_t)1 << b) < range) b += 2;
  return b / 2;
}
static uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static bool order_keys(uint64_t seed, long long n, long long m, int epochs, MlpOrderKeys &k) {
  if (n <= 0 || m <= 0 || m > n || epochs <= 0 || epochs > MLP_MAX_EPOCHS || n > (1LL << 30) ||
      m * epochs > 0x7fffffffLL)
    return false;
  uint64_t s = seed ^ 0xADA3ULL;
  for (int i = 0; i < 4; i++) k.sample[i] = (uint32_t)splitmix64(s);
  for (int i = 0; i < 4 * epochs; i++) k.epoch[i] = (uint32_t)splitmix64(s);
  k.half_n = half_bits((uint32_t)n);
  k.half_m = half_bits((uint32_t)m);
  return true;
}

hipError_t mlp_order(uint64_t seed, long long n, long long m, int epochs, int *dOut, hipStream_t st) {
  MlpOrderKeys k;
  if (!order_keys(seed, n, m, epochs, k)) return hipErrorInvalidValue;
  const long long total = m * epochs;
  k_mlp_order<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(k, (uint32_t)n, (uint32_t)m, epochs, dOut);
  return hipGetLastError();
}

bool mlp_order_host(uint64_t seed, long long n, long long m, int epochs, int *out) {
  MlpOrderKeys k;
  if (!order_keys(seed, n, m, epochs, k)) return false;
  for (long long i = 0; i < m * epochs; i++) {
    const int e = (int)(i / m);
    const uint32_t j = feistel_perm((uint32_t)(i - (long long)e * m), (uint32_t)m, k.half_m, k.epoch + 4 * e);
    out[i] = (int)feistel_perm(j, (uint32_t)n, k.half_n, k.sample);
  }
  return true;
}
size_t mlp_part_floats(int n) { return (size_t)((n + MB - 1) / MB) * (MLP_NPARAMS + 1); }

hipError_t mlp_forward(const float *dX, int n, const float *dP, float *dOut, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  // a wave per 16 samples up to 2048 waves (two workgroups per CU at the
  // kernel's two waves per SIMD), each then striding over the tiles
  const int ntile = (n + 15) / 16;
  k_mlp_forward<<<std::min((ntile + 3) / 4, 512), 256, 0, st>>>(dX, n, dP, dOut);
  return hipGetLastError();
}

hipError_t mlp_gradient(const float *dX, const float *dY, const int *dIdx, int b0, int n, const float *dP,
                        float *dPart, float *dGrad, float *dLoss, hipStream_t st) {
  if (n <= 0) return hipErrorInvalidValue;
  const int nblk = (n + MB - 1) / MB;
  k_mlp_part<<<nblk, 256, 0, st>>>(dX, dY, dIdx, b0, n, dP, dPart);
  k_mlp_reduce<<<(MLP_NPARAMS + 1 + 255) / 256, 256, 0, st>>>(dPart, nblk, n, dGrad, dLoss, 0, nullptr, nullptr,
                                                               nullptr, 0.f, 0.f, 0.f, 0.f, 0);
  return hipGetLastError();
}

hipError_t mlp_loss(const float *dX, const float *dY, int n, const float *dP, float *dPart, float *dLoss,
                    hipStream_t st) {
  if (n <= 0) return hipErrorInvalidValue;
  const int nblk = (n + 255) / 256;
  k_mlp_loss_part<<<nblk, 256, 0, st>>>(dX, dY, n, dP, dPart);
  k_mlp_loss_reduce<<<1, 64, 0, st>>>(dPart, nblk, n, dLoss);
  return hipGetLastError();
}

hipError_t mlp_train_step(const float *dX, const float *dY, const int *dIdx, int b0, int n, float *dP, float *dPart,
                          float *dGrad, float *dLoss, float *dM, float *dV, float lr, int step, hipStream_t st) {
  if (n <= 0) return hipErrorInvalidValue;
  const int nblk = (n + MB - 1) / MB;
  k_mlp_part<<<nblk, 256, 0, st>>>(dX, dY, dIdx, b0, n, dP, dPart);
  k_mlp_reduce<<<(MLP_NPARAMS + 1 + 255) / 256, 256, 0, st>>>(dPart, nblk, n, dGrad, dLoss, 1, dP, dM, dV, lr, 0.9f,
                                                               0.999f, 1e-8f, step);
  return hipGetLastError();
}

hipError_t mlp_labels(const unsigned *dHits, const float *dInvDen, const long long *dSetBase, int nsets, long long n,
                      float *dY, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_mlp_labels<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(dHits, dInvDen, dSetBase, nsets, n, dY);
  return hipGetLastError();
}
