// Learned correction to the Kronecker preconditioner: an MLP per channel block,
//   x = v_c / |v|  (2 n^2 reals)  ->  GELU(W1 x + b1 + e_c)  ->  GELU(W2 h1 + b2)  ->  W3 h2 + b3,
//   P(v)_c = K_c^-1 (v_c + sigma |v| y_c)   (x = v_c / |v| with the norm of the whole vector),
// BF16 weights and activations into MFMA, FP32 accumulation, FP32 master weights + Adam.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "pp_common.h"
#include "pp_mixed.h"

namespace pp {

struct Learned {
  int nchan = 0, n = 0, in = 0, hidden = 0;
  float sigma = 1.0f;        // output scale
  long long p_count = 0;     // FP32 parameters
  // offsets into the flat parameter vector
  long long o_w1 = 0, o_b1 = 0, o_emb = 0, o_w2 = 0, o_b2 = 0, o_w3 = 0, o_b3 = 0;
  DeviceBuffer<float> params, grads, adam_m, adam_v;
  DeviceBuffer<uint16_t> w1h, w2h, w3h;   // BF16 copies of the weight matrices
  int adam_t = 0;
  // activations for a batch of `cap` rows
  int cap = 0;
  DeviceBuffer<uint16_t> xh, h1h, h2h, gh, dz2h, dz1h;
  DeviceBuffer<float> z1, z2, y, gf, dzf, part;
  DeviceBuffer<double2> loss_part;
  DeviceBuffer<int> iota;
  DeviceBuffer<double> nrm;          // per-block norms, [nchan] = global norm
  DeviceBuffer<double> n2p;          // partial sums of squares of the inference input
  DeviceBuffer<double2> tmp64;
  int splits_max = 32;
  void ensure(int rows);
  void refresh_bf16(hipStream_t s);
  // y (rows x in, FP32) = network(x) for BF16 inputs already in xh; keeps activations
  void forward(int rows, const int* chan, hipStream_t s);
  // loss = sum |y - t|^2 / rows and the gradient of it into grads (forward must have run)
  double backward(int rows, const int* chan, const float* t, hipStream_t s, bool want_loss);
  void adam(float lr, float beta1, float beta2, float eps, hipStream_t s);
};

// out = K^-1 (in + sigma |in| net(in_c / |in|)) per channel block c (FP32 vectors); sz (may be
// NULL): also (S (x) S) out, formed by the Kronecker solve (Kron::apply32_add)
void learned_apply(Learned& L, const Kron& K, double2 a, const float2* in, float2* out, hipStream_t s,
                   float2* sz = nullptr);

// C[M x N] (+)= sum_k A(m, k) B(k, n), A(m, k) = a[m*sam + k*sak], B(k, n) = b[k*sbk + n*sbn]
// (BF16 inputs, FP32 result).  Splits K into at most `splits` slices, slice z written to
// C + z*M*N; returns the number of slices written.
int gemm_bf16(int M, int N, int K, const uint16_t* a, long long sam, long long sak, const uint16_t* b,
               long long sbk, long long sbn, float* c, int splits, hipStream_t s);

}  // namespace pp

struct pp_learned {
  pp::Learned L;
};
