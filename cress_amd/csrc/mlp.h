// Learned cold-rain microphysics: a 9 -> 128 -> 128 -> 128 -> 7 GELU
// network standing in for the local source terms of s_srcblk
// (srcblk.f90:488-663, called at coldblk.f90:354-358), trained on pairs the
// plain scheme produces.  BF16 weights and activations into the matrix
// cores, FP32 accumulation and FP32 master weights.
//
// Features (per cell, past level, as srcblk reads them):
//   0 T = (ptbr + ptp') pi   1 p = pbr + p'   2 rho = rbr
//   3..8 qv, qc, qr, qi, qs, qg
// Targets: the change srcblk makes to the future level divided by the
// microphysics step 2 dtb:
//   0 dT/dt = pi d(ptp)/dt   1..6 dqv/dt .. dqg/dt
#pragma once
#include "common.h"

namespace cress {

constexpr int MLP_NIN = 9, MLP_NOUT = 7, MLP_H = 128;
constexpr int MLP_INP = 16;    // inputs padded to one 16-wide k step
constexpr int MLP_OUTP = 16;   // outputs padded to one 16-row block

// FP32 parameter block, row-major W[in][out]:
//   W1 16x128 | b1 128 | W2 128x128 | b2 128 | W3 128x128 | b3 128 | W4 128x16 | b4 16
constexpr int MLP_OW1 = 0;
constexpr int MLP_OB1 = MLP_OW1 + MLP_INP * MLP_H;
constexpr int MLP_OW2 = MLP_OB1 + MLP_H;
constexpr int MLP_OB2 = MLP_OW2 + MLP_H * MLP_H;
constexpr int MLP_OW3 = MLP_OB2 + MLP_H;
constexpr int MLP_OB3 = MLP_OW3 + MLP_H * MLP_H;
constexpr int MLP_OW4 = MLP_OB3 + MLP_H;
constexpr int MLP_OB4 = MLP_OW4 + MLP_H * MLP_OUTP;
constexpr int MLP_NPARAM = MLP_OB4 + MLP_OUTP;

// Normalisation and envelope block (FP32):
//   in_mu[16] in_rsd[16] out_mu[16] out_sd[16] env_lo[16] env_hi[16]
constexpr int MLP_NNORM = 96;

// Device view of a network for inference.
struct MlpNet {
  const unsigned short *w;   // BF16 copy of the parameter block
  const float *p;            // FP32 parameter block (biases are read from here)
  const float *norm;         // normalisation and envelope
};

// Per-step buffers of the microphysics with the network.
struct MlpStep {
  int *flag;                 // per cell: 0 plain srcblk, 1 network, 2 no active process
  int *list;                 // cells the network handles (linear indices)
  unsigned short *x;         // their normalised features, BF16 [count][16]
  unsigned *count;           // entries in list
  float *out;                // per cell, 7 blocks of n floats: the network's tendencies
};

// Training pairs captured from the plain scheme: FP32 [cap][9] and [cap][7].
struct MlpCapture {
  float *in, *tgt;
  unsigned long long *count;
  long cap;
};

__device__ __forceinline__ float mlp_gelu(float z) {
  // tanh form: 0.5 z (1 + tanh(c (z + 0.044715 z^3))), tanh u = 1 - 2/(1 + e^{2u})
  const float u = 0.7978845608f * (z + 0.044715f * z * z * z);
  const float t = 1.0f - 2.0f / (1.0f + __expf(2.0f * u));
  return 0.5f * z * (1.0f + t);
}

// The same GELU with the hardware exp2 and reciprocal (about 1 ulp each) in
// place of the library exp and the IEEE division: for the inference
// kernel, whose activations are rounded to BF16 anyway.
// 0.5 z (1 + tanh u) = z / (1 + e^{-2u}), e^{-2u} = 2^{-2u log2(e)}.
__device__ __forceinline__ float mlp_gelu_fast(float z) {
  const float u = z * __builtin_fmaf(0.7978845608f * 0.044715f, z * z, 0.7978845608f);
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-2.8853900818f * u));
}

__device__ __forceinline__ float mlp_dgelu(float z) {
  const float u = 0.7978845608f * (z + 0.044715f * z * z * z);
  const float t = 1.0f - 2.0f / (1.0f + __expf(2.0f * u));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * 0.7978845608f * (1.0f + 3.0f * 0.044715f * z * z);
}

// Round to nearest even BF16 (NaN stays NaN).
__host__ __device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__host__ __device__ __forceinline__ float bf2f(unsigned short h) {
  const unsigned u = (unsigned)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// Generic BF16 GEMM on the matrix cores: C[M x N] = A[M x K] B[K x N] with
// element strides (so transposes are free) and a fused epilogue.
enum MlpEpi { EPI_STORE = 0, EPI_ATOMIC = 1, EPI_BIAS_GELU = 2, EPI_BIAS = 3, EPI_DGELU = 4 };
struct GemmArgs {
  const unsigned short *A;
  long a_m, a_k;
  const unsigned short *B;
  long b_k, b_n;
  int M, N, K, kchunk, epi;
  float *C;            // FP32 result, row stride c_m
  long c_m;
  const float *bias;   // EPI_BIAS*, per column
  unsigned short *H;   // EPI_BIAS_GELU: BF16 gelu(C); EPI_DGELU: BF16 of the result
  long h_m;
  const float *Z;      // EPI_DGELU: pre-activation, row stride c_m
};

__global__ void k_gemm_bf16(GemmArgs g);
__global__ void k_mlp_infer(MlpNet net, MlpStep s, long n);
// workgroups (of MLP_INFER_THREADS threads) of k_mlp_infer for up to n cells
unsigned mlp_infer_blocks(long n);
constexpr int MLP_INFER_THREADS = 512;
__global__ void k_mlp_gather(const float *in, const float *tgt, long nsamp, const float *norm, unsigned seed,
                             int identity, int batch, unsigned short *x, float *t);
__global__ void k_mlp_loss(const float *y, const float *t, int batch, float scale, float *dy, unsigned short *dyb,
                           double *loss);
constexpr int MLP_CS_LANES = 4, MLP_CS_ROWS = 128;   // k_mlp_colsum's row lanes and rows per workgroup
__global__ void k_mlp_colsum(const float *d, long rows, int cols, long stride, float *out);
__global__ void k_mlp_adam(float *p, const float *gr, float *m, float *v, unsigned short *w, int n, float lr,
                           float b1, float b2, float eps, float c1, float c2, const unsigned char *trainable);
__global__ void k_mlp_fill(float *p, float v, long n);
__global__ void k_mlp_cast(const float *p, unsigned short *w, int n);
__global__ void k_mlp_stats(const float *a, long rows, int cols, double *sum, double *sumsq, unsigned *mn,
                            unsigned *mx);
__global__ void k_u32_set(unsigned *p, unsigned v);

}  // namespace cress

// Host-side network object behind the C ABI's cress_mlp (mlp.hip): the
// parameters and Adam state, the captured training pairs, and the training
// workspace for batches of up to bmax samples.  One HIP stream.
struct cress_mlp_impl {
  float *p, *g, *m, *v;     // FP32 parameters, gradient, Adam moments [MLP_NPARAM]
  unsigned short *w;        // BF16 copy of p
  unsigned char *trainable; // 0 on the padding rows/columns
  float *norm;              // [MLP_NNORM]
  float *din, *dtgt;        // captured pairs
  unsigned long long *dcount;
  long cap;
  int bmax;
  unsigned short *X, *H1, *H2, *H3, *dYb, *dZb;
  float *Z1, *Z2, *Z3, *Y, *T, *dY, *D;
  double *loss;             // per training step
  long nloss;
  int adam_t;
  hipStream_t st;
};
