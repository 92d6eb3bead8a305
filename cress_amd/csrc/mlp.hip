// Kernels of the learned microphysics (declarations and layouts in mlp.h):
// the matrix-core GEMM and element-wise kernels of training, and the fused
// inference kernel the large step calls.
//
// Matrix cores: V_MFMA_F32_16X16X16_BF16 (one wavefront, 16x16 output
// tile, K = 16).  Lane l = 16 q + r (q = l / 16, r = l % 16) holds
//   A fragment: A[row r][k 4q .. 4q+3]
//   B fragment: B[k 4q .. 4q+3][col r]
//   accumulator: C[row 4q + v][col r], v = 0..3.
// In the inference kernel the samples are the columns: the accumulator of
// one 16-row block of a layer's output is then exactly the B fragment of
// one k step of the next layer, so activations never leave registers.
#include <hip/hip_runtime.h>

#include "mlp.h"

namespace cress {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4_t mfma16(bf16x4_t a, bf16x4_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Four floats to a BF16 fragment, round to nearest even (v_cvt_pk_bf16_f32).
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bf16x4_t pack4(float a, float b, float c, float d) {
  bf16x2_t lo, hi;
  lo[0] = (__bf16)a;
  lo[1] = (__bf16)b;
  hi[0] = (__bf16)c;
  hi[1] = (__bf16)d;
  const unsigned ul = __builtin_bit_cast(unsigned, lo), uh = __builtin_bit_cast(unsigned, hi);
  bf16x4_t o;
  o[0] = (short)(ul & 0xffffu);
  o[1] = (short)(ul >> 16);
  o[2] = (short)(uh & 0xffffu);
  o[3] = (short)(uh >> 16);
  return o;
}

__device__ __forceinline__ bf16x4_t ld4(const unsigned short *p) {
  // four consecutive BF16 values, 8-byte aligned
  const uint2 u = *(const uint2 *)p;
  bf16x4_t o;
  o[0] = (short)(u.x & 0xffffu);
  o[1] = (short)(u.x >> 16);
  o[2] = (short)(u.y & 0xffffu);
  o[3] = (short)(u.y >> 16);
  return o;
}

// ---------------------------------------------------------------------------
// GEMM: 64x64 output tile per workgroup of four wavefronts (each a 32x32
// quadrant of four 16x16 tiles), K in slabs of 32 staged through LDS.
// blockIdx.z splits K into chunks of g.kchunk (EPI_ATOMIC adds the parts).
constexpr int GT = 64, GK = 32, GPAD = 8;

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ unsigned short As[GT][GK + GPAD];
  __shared__ unsigned short Bt[GT][GK + GPAD];
  const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
  const int kb0 = blockIdx.z * g.kchunk;
  const int kb1 = min(g.K, kb0 + g.kchunk);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r = lane & 15, q = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4_t acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int kb = kb0; kb < kb1; kb += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      int rr, cc;
      if (g.a_k == 1) { rr = e / GK; cc = e % GK; } else { rr = e % GT; cc = e / GT; }
      const int m = m0 + rr, k = kb + cc;
      As[rr][cc] = (m < g.M && k < kb1) ? g.A[(long)m * g.a_m + (long)k * g.a_k] : (unsigned short)0;
    }
    for (int e = tid; e < GT * GK; e += 256) {
      int nn, kk;
      if (g.b_n == 1) { nn = e % GT; kk = e / GT; } else { nn = e / GK; kk = e % GK; }
      const int n = n0 + nn, k = kb + kk;
      Bt[nn][kk] = (n < g.N && k < kb1) ? g.B[(long)k * g.b_k + (long)n * g.b_n] : (unsigned short)0;
    }
    __syncthreads();
    for (int ks = 0; ks < GK; ks += 16) {
      bf16x4_t af[2], bfr[2];
      for (int a = 0; a < 2; a++) af[a] = ld4(&As[wm + a * 16 + r][ks + 4 * q]);
      for (int b = 0; b < 2; b++) bfr[b] = ld4(&Bt[wn + b * 16 + r][ks + 4 * q]);
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = mfma16(af[a], bfr[b], acc[a][b]);
    }
    __syncthreads();
  }
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      const int n = n0 + wn + b * 16 + r;
      if (n >= g.N) continue;
      for (int v = 0; v < 4; v++) {
        const int m = m0 + wm + a * 16 + 4 * q + v;
        if (m >= g.M) continue;
        const float x = acc[a][b][v];
        const long o = (long)m * g.c_m + n;
        switch (g.epi) {
          case EPI_STORE: g.C[o] = x; break;
          case EPI_ATOMIC: atomicAdd(&g.C[o], x); break;
          case EPI_BIAS_GELU: {
            const float z = x + g.bias[n];
            g.C[o] = z;
            g.H[(long)m * g.h_m + n] = f2bf(mlp_gelu(z));
            break;
          }
          case EPI_BIAS: g.C[o] = x + g.bias[n]; break;
          default: {   // EPI_DGELU
            const float d = x * mlp_dgelu(g.Z[o]);
            g.C[o] = d;
            g.H[(long)m * g.h_m + n] = f2bf(d);
            break;
          }
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Fused inference over the step's list of network cells.  A workgroup of
// MLP_WAVES wavefronts shares one copy of the weights in LDS (BF16, transposed to
// [out][in] with padded rows so the A fragments are 8-byte reads without
// bank conflicts); each wavefront takes MLP_G groups of 16 cells at a time
// (grid-stride), so every weight fragment read from LDS feeds MLP_G matrix
// operations, and the activations never leave registers.
constexpr int IW1 = MLP_INP + 8, IWH = MLP_H + 8;
constexpr int MLP_G = 2;            // 16-cell groups per wavefront pass
constexpr int MLP_WAVES = MLP_INFER_THREADS / 64;   // wavefronts per workgroup

__global__ void __launch_bounds__(64 * MLP_WAVES) k_mlp_infer(MlpNet net, MlpStep s, long n) {
  __shared__ __align__(16) unsigned short w1[MLP_H * IW1];
  __shared__ __align__(16) unsigned short w2[MLP_H * IWH];
  __shared__ __align__(16) unsigned short w3[MLP_H * IWH];
  __shared__ __align__(16) unsigned short w4[MLP_OUTP * IWH];
  __shared__ float bias[3 * MLP_H + MLP_OUTP];
  __shared__ float omu[MLP_OUTP], osd[MLP_OUTP];
  const int tid = threadIdx.x, nth = 64 * MLP_WAVES;
  for (int e = tid; e < MLP_INP * MLP_H; e += nth) {      // W1[in][out] -> w1[out][in]
    const int i = e / MLP_H, o = e % MLP_H;
    w1[o * IW1 + i] = net.w[MLP_OW1 + e];
  }
  for (int e = tid; e < MLP_H * MLP_H; e += nth) {
    const int i = e / MLP_H, o = e % MLP_H;
    w2[o * IWH + i] = net.w[MLP_OW2 + e];
    w3[o * IWH + i] = net.w[MLP_OW3 + e];
  }
  for (int e = tid; e < MLP_H * MLP_OUTP; e += nth) {
    const int i = e / MLP_OUTP, o = e % MLP_OUTP;
    w4[o * IWH + i] = net.w[MLP_OW4 + e];
  }
  for (int e = tid; e < MLP_H; e += nth) {
    bias[e] = net.p[MLP_OB1 + e];
    bias[MLP_H + e] = net.p[MLP_OB2 + e];
    bias[2 * MLP_H + e] = net.p[MLP_OB3 + e];
  }
  if (tid < MLP_OUTP) {
    bias[3 * MLP_H + tid] = net.p[MLP_OB4 + tid];
    omu[tid] = net.norm[32 + tid];
    osd[tid] = net.norm[48 + tid];
  }
  __syncthreads();
  const unsigned cnt = *s.count;
  const unsigned npass = (cnt + 16u * MLP_G - 1u) / (16u * MLP_G);
  const int wave = tid >> 6, lane = tid & 63, r = lane & 15, q = lane >> 4;
  const unsigned nwaves = gridDim.x * (unsigned)MLP_WAVES;
  for (unsigned pass = blockIdx.x * (unsigned)MLP_WAVES + (unsigned)wave; pass < npass; pass += nwaves) {
    bf16x4_t h[MLP_G][8];
    f32x4_t acc[MLP_G][8];
    // layer 1: 128 x 16 weights by 16 x 16 cells per group
    for (int gi = 0; gi < MLP_G; gi++) {
      const unsigned smp = (pass * MLP_G + gi) * 16u + (unsigned)r;
      bf16x4_t xb = bf16x4_t{0, 0, 0, 0};
      if (smp < cnt) xb = ld4(s.x + (long)smp * MLP_INP + 4 * q);
      for (int mb = 0; mb < 8; mb++)
        acc[gi][mb] = mfma16(ld4(&w1[(mb * 16 + r) * IW1 + 4 * q]), xb, f32x4_t{0.f, 0.f, 0.f, 0.f});
    }
    for (int gi = 0; gi < MLP_G; gi++)
      for (int mb = 0; mb < 8; mb++) {
        const float *bq = bias + mb * 16 + 4 * q;
        h[gi][mb] = pack4(mlp_gelu_fast(acc[gi][mb][0] + bq[0]), mlp_gelu_fast(acc[gi][mb][1] + bq[1]),
                          mlp_gelu_fast(acc[gi][mb][2] + bq[2]), mlp_gelu_fast(acc[gi][mb][3] + bq[3]));
      }
    // layers 2 and 3: 128 x 128 weights by 128 x 16 cells per group; each
    // weight fragment serves every group
    for (int layer = 0; layer < 2; layer++) {
      const unsigned short *w = layer == 0 ? w2 : w3;
      const float *bb = bias + (layer + 1) * MLP_H;
      for (int gi = 0; gi < MLP_G; gi++)
        for (int mb = 0; mb < 8; mb++) acc[gi][mb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
      for (int kb = 0; kb < 8; kb++)
        for (int mb = 0; mb < 8; mb++) {
          const bf16x4_t a = ld4(&w[(mb * 16 + r) * IWH + kb * 16 + 4 * q]);
          for (int gi = 0; gi < MLP_G; gi++) acc[gi][mb] = mfma16(a, h[gi][kb], acc[gi][mb]);
        }
      for (int gi = 0; gi < MLP_G; gi++)
        for (int mb = 0; mb < 8; mb++) {
          const float *bq = bb + mb * 16 + 4 * q;
          h[gi][mb] = pack4(mlp_gelu_fast(acc[gi][mb][0] + bq[0]), mlp_gelu_fast(acc[gi][mb][1] + bq[1]),
                            mlp_gelu_fast(acc[gi][mb][2] + bq[2]), mlp_gelu_fast(acc[gi][mb][3] + bq[3]));
        }
    }
    // layer 4: 16 x 128 weights by 128 x 16 cells; rows 0..6 are the outputs
    f32x4_t y[MLP_G];
    for (int gi = 0; gi < MLP_G; gi++) y[gi] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < 8; kb++) {
      const bf16x4_t a = ld4(&w4[r * IWH + kb * 16 + 4 * q]);
      for (int gi = 0; gi < MLP_G; gi++) y[gi] = mfma16(a, h[gi][kb], y[gi]);
    }
    for (int gi = 0; gi < MLP_G; gi++) {
      const unsigned smp = (pass * MLP_G + gi) * 16u + (unsigned)r;
      if (smp < cnt && q < 2) {
        const long cell = s.list[smp];
        for (int v = 0; v < 4; v++) {
          const int o = 4 * q + v;
          if (o < MLP_NOUT) s.out[(long)o * n + cell] = (y[gi][v] + bias[3 * MLP_H + o]) * osd[o] + omu[o];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Training helpers.

// Normalised batch from the captured pairs at pseudo-random rows (drawn
// with replacement from a counter hash of seed and batch position), or
// rows 0..batch-1 in order when identity is set.
__global__ void k_mlp_gather(const float *in, const float *tgt, long nsamp, const float *norm, unsigned seed,
                             int identity, int batch, unsigned short *x, float *t) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  unsigned long long h = ((unsigned long long)seed << 32) ^ (unsigned long long)b;
  h += 0x9e3779b97f4a7c15ull;
  h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ull;
  h = (h ^ (h >> 27)) * 0x94d049bb133111ebull;
  h ^= h >> 31;
  const long row = identity ? (long)b : (long)(h % (unsigned long long)nsamp);
  for (int f = 0; f < MLP_INP; f++) {
    const float v = f < MLP_NIN ? (in[row * MLP_NIN + f] - norm[f]) * norm[16 + f] : 0.0f;
    x[(long)b * MLP_INP + f] = f2bf(v);
  }
  for (int o = 0; o < MLP_NOUT; o++) t[(long)b * MLP_NOUT + o] = (tgt[row * MLP_NOUT + o] - norm[32 + o]) / norm[48 + o];
}

// Mean-squared-error gradient dY = scale (Y - T) on the 7 outputs (zero on
// the padding), FP32 and BF16, and the summed squared error in FP64.
__global__ void k_mlp_loss(const float *y, const float *t, int batch, float scale, float *dy, unsigned short *dyb,
                           double *loss) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  double se = 0.0;
  if (b < batch) {
    for (int o = 0; o < MLP_OUTP; o++) {
      float d = 0.0f;
      if (o < MLP_NOUT) {
        const float e = y[(long)b * MLP_OUTP + o] - t[(long)b * MLP_NOUT + o];
        se += (double)e * (double)e;
        d = scale * e;
      }
      dy[(long)b * MLP_OUTP + o] = d;
      dyb[(long)b * MLP_OUTP + o] = f2bf(d);
    }
  }
  // wavefront sum, one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) se += __shfl_down(se, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, se);
}

// Column sums of a rows x cols matrix (row stride `stride`) added into out.
// A workgroup of 64 x MLP_CS_LANES threads takes 64 columns (x, coalesced
// along a row) and MLP_CS_ROWS rows, each of the MLP_CS_LANES row lanes
// summing every MLP_CS_LANES-th row; the lanes are combined in LDS and one
// atomic per column adds the workgroup's sum.
__global__ void __launch_bounds__(64 * MLP_CS_LANES) k_mlp_colsum(const float *d, long rows, int cols, long stride,
                                                                   float *out) {
  __shared__ float part[MLP_CS_LANES][64];
  const int c = blockIdx.x * 64 + threadIdx.x;
  const long r0 = (long)blockIdx.y * MLP_CS_ROWS;
  const long r1 = min(rows, r0 + MLP_CS_ROWS);
  float s = 0.0f;
  if (c < cols)
    for (long r = r0 + threadIdx.y; r < r1; r += MLP_CS_LANES) s += d[r * stride + c];
  part[threadIdx.y][threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.y == 0 && c < cols) {
    float t = part[0][threadIdx.x];
    for (int y = 1; y < MLP_CS_LANES; y++) t += part[y][threadIdx.x];
    atomicAdd(&out[c], t);
  }
}

// Adam (Kingma & Ba 2015) on the FP32 master block; c1 = 1/(1-b1^t),
// c2 = 1/(1-b2^t); refreshes the BF16 copy.  trainable[i] = 0 keeps the
// padding rows/columns at zero.
__global__ void k_mlp_adam(float *p, const float *gr, float *m, float *v, unsigned short *w, int n, float lr,
                           float b1, float b2, float eps, float c1, float c2, const unsigned char *trainable) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (trainable[i]) {
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
  w[i] = f2bf(p[i]);
}

__global__ void k_mlp_fill(float *p, float v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_mlp_cast(const float *p, unsigned short *w, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = f2bf(p[i]);
}

__device__ __forceinline__ unsigned ord_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Per-column sum, sum of squares (FP64) and min/max (order-preserving bits)
// of a rows x cols FP32 matrix; one thread per row, wavefront-reduced.
__global__ void k_mlp_stats(const float *a, long rows, int cols, double *sum, double *sumsq, unsigned *mn,
                            unsigned *mx) {
  const long rw = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = rw < rows;
  for (int c = 0; c < cols; c++) {
    const float x = in ? a[rw * cols + c] : 0.0f;
    double s = in ? (double)x : 0.0, s2 = in ? (double)x * (double)x : 0.0;
    unsigned lo = in ? ord_bits(x) : 0xffffffffu, hi = in ? ord_bits(x) : 0u;
    for (int off = 32; off > 0; off >>= 1) {
      s += __shfl_down(s, off);
      s2 += __shfl_down(s2, off);
      lo = min(lo, (unsigned)__shfl_down((int)lo, off));
      hi = max(hi, (unsigned)__shfl_down((int)hi, off));
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&sum[c], s);
      atomicAdd(&sumsq[c], s2);
      atomicMin(&mn[c], lo);
      atomicMax(&mx[c], hi);
    }
  }
}

__global__ void k_u32_set(unsigned *p, unsigned v) {
  if (threadIdx.x == 0) *p = v;
}

// Workgroups of k_mlp_infer for up to n cells: one per CU (its LDS holds one
// copy of the weights) up to the chip, fewer when n is small.
unsigned mlp_infer_blocks(long n) {
  const long passes = (n + 16L * MLP_G - 1) / (16L * MLP_G);
  const long b = (passes + MLP_WAVES - 1) / MLP_WAVES;
  return (unsigned)(b < 1 ? 1 : (b > 256 ? 256 : b));
}

}  // namespace cress

// ===========================================================================
// C ABI of the network (include/cress_hip.h).
// ===========================================================================
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "cress_hip.h"

struct cress_mlp : public cress_mlp_impl {};

using namespace cress;

namespace {

template <class T>
int dmalloc(T **p, size_t count) {
  void *q = nullptr;
  if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return CRESS_ENOMEM;
  *p = (T *)q;
  return CRESS_OK;
}

void gemm(cress_mlp *m, const unsigned short *A, long a_m, long a_k, const unsigned short *B, long b_k, long b_n,
          int M, int N, int K, int epi, float *C, long c_m, const float *bias = nullptr,
          unsigned short *H = nullptr, long h_m = 0, const float *Z = nullptr, int kchunk = 0) {
  GemmArgs g;
  g.A = A; g.a_m = a_m; g.a_k = a_k;
  g.B = B; g.b_k = b_k; g.b_n = b_n;
  g.M = M; g.N = N; g.K = K;
  g.kchunk = kchunk > 0 ? kchunk : K;
  g.epi = epi;
  g.C = C; g.c_m = c_m; g.bias = bias; g.H = H; g.h_m = h_m; g.Z = Z;
  const dim3 grid((M + GT - 1) / GT, (N + GT - 1) / GT, (K + g.kchunk - 1) / g.kchunk);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, m->st, g);
}

void fill(cress_mlp *m, float *p, float v, long n) {
  hipLaunchKernelGGL(k_mlp_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->st, p, v, n);
}

// Forward and backward of one batch already in m->X (BF16, normalised) and
// m->T (normalised targets); gradient into m->g, the batch's summed
// squared error into *loss (device).
void fwd_bwd(cress_mlp *m, int B, double *loss) {
  const int H = MLP_H;
  const unsigned short *w = m->w;
  gemm(m, m->X, MLP_INP, 1, w + MLP_OW1, H, 1, B, H, MLP_INP, EPI_BIAS_GELU, m->Z1, H, m->p + MLP_OB1, m->H1, H);
  gemm(m, m->H1, H, 1, w + MLP_OW2, H, 1, B, H, H, EPI_BIAS_GELU, m->Z2, H, m->p + MLP_OB2, m->H2, H);
  gemm(m, m->H2, H, 1, w + MLP_OW3, H, 1, B, H, H, EPI_BIAS_GELU, m->Z3, H, m->p + MLP_OB3, m->H3, H);
  gemm(m, m->H3, H, 1, w + MLP_OW4, MLP_OUTP, 1, B, MLP_OUTP, H, EPI_BIAS, m->Y, MLP_OUTP, m->p + MLP_OB4);
  const float scale = 2.0f / ((float)B * (float)MLP_NOUT);
  hipLaunchKernelGGL(k_mlp_loss, dim3((B + 255) / 256), dim3(256), 0, m->st, m->Y, m->T, B, scale, m->dY, m->dYb,
                     loss);
  fill(m, m->g, 0.0f, MLP_NPARAM);
  const int kc = 2048;                                   // split of the batch in the weight gradients
  const dim3 cs_block(64, MLP_CS_LANES);
  const unsigned cs_rows = (unsigned)((B + MLP_CS_ROWS - 1) / MLP_CS_ROWS);
  // output layer
  gemm(m, m->H3, 1, H, m->dYb, MLP_OUTP, 1, H, MLP_OUTP, B, EPI_ATOMIC, m->g + MLP_OW4, MLP_OUTP, nullptr, nullptr, 0,
       nullptr, kc);
  hipLaunchKernelGGL(k_mlp_colsum, dim3((MLP_OUTP + 63) / 64, cs_rows), cs_block, 0, m->st, m->dY, (long)B,
                     MLP_OUTP, (long)MLP_OUTP, m->g + MLP_OB4);
  // hidden layers: dH = dZ_next W_next^T, dZ = dH * gelu'(Z), dW = H_prev^T dZ
  struct L { int ow, ob; const float *Z; const unsigned short *Hprev; int kin; int on; };
  const L layers[3] = {{MLP_OW3, MLP_OB3, m->Z3, m->H2, H, MLP_OW4},
                       {MLP_OW2, MLP_OB2, m->Z2, m->H1, H, MLP_OW3},
                       {MLP_OW1, MLP_OB1, m->Z1, m->X, MLP_INP, MLP_OW2}};
  const unsigned short *dnext = m->dYb;
  int nnext = MLP_OUTP;
  for (int l = 0; l < 3; l++) {
    const L &c = layers[l];
    unsigned short *dz = m->dZb + (size_t)(l & 1) * (size_t)m->bmax * H;
    // dH = dnext [B x nnext] * W_next^T [nnext x H]  (W_next stored [H][nnext])
    gemm(m, dnext, nnext, 1, w + c.on, 1, nnext, B, H, nnext, EPI_DGELU, m->D, H, nullptr, dz, H, c.Z);
    gemm(m, c.Hprev, 1, c.kin, dz, H, 1, c.kin, H, B, EPI_ATOMIC, m->g + c.ow, H, nullptr, nullptr, 0, nullptr, kc);
    hipLaunchKernelGGL(k_mlp_colsum, dim3((H + 63) / 64, cs_rows), cs_block, 0, m->st, m->D, (long)B, H, (long)H,
                       m->g + c.ob);
    dnext = dz;
    nnext = H;
  }
}

int upload_rows(cress_mlp *m, const float *in, const float *tgt, int n) {
  // raw rows to temporary device arrays, then normalised in order into the
  // batch workspace
  float *din = nullptr, *dt = nullptr;
  int rc = dmalloc(&din, (size_t)n * MLP_NIN);
  if (!rc) rc = dmalloc(&dt, (size_t)n * MLP_NOUT);
  if (!rc && hipMemcpyAsync(din, in, (size_t)n * MLP_NIN * 4, hipMemcpyHostToDevice, m->st) != hipSuccess)
    rc = CRESS_EHIP;
  if (!rc && tgt && hipMemcpyAsync(dt, tgt, (size_t)n * MLP_NOUT * 4, hipMemcpyHostToDevice, m->st) != hipSuccess)
    rc = CRESS_EHIP;
  if (!rc && !tgt) fill(m, dt, 0.0f, (long)n * MLP_NOUT);
  if (!rc)
    hipLaunchKernelGGL(k_mlp_gather, dim3((n + 255) / 256), dim3(256), 0, m->st, din, dt, (long)n, m->norm,
                       0u, 1, n, m->X, m->T);
  if (hipStreamSynchronize(m->st) != hipSuccess && !rc) rc = CRESS_EHIP;
  hipFree(din);
  hipFree(dt);
  return rc;
}

unsigned long long splitmix(unsigned long long &s) {
  unsigned long long z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int cress_mlp_nparam(void) { return MLP_NPARAM; }
int cress_mlp_nnorm(void) { return MLP_NNORM; }

int cress_mlp_destroy(cress_mlp *m) {
  if (!m) return CRESS_OK;
  if (m->st) hipStreamSynchronize(m->st);
  void *ps[] = {m->p, m->g, m->m, m->v, m->w, m->trainable, m->norm, m->din, m->dtgt, m->dcount, m->X, m->H1,
                m->H2, m->H3, m->dYb, m->dZb, m->Z1, m->Z2, m->Z3, m->Y, m->T, m->dY, m->D, m->loss};
  for (void *q : ps)
    if (q) hipFree(q);
  if (m->st) hipStreamDestroy(m->st);
  delete m;
  return CRESS_OK;
}

int cress_mlp_create(long capacity, int max_batch, cress_mlp **out) {
  if (!out || capacity < 1 || max_batch < 16 || max_batch > (1 << 22)) return CRESS_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  cress_mlp *m = new cress_mlp();
  memset(static_cast<cress_mlp_impl *>(m), 0, sizeof(cress_mlp_impl));
  m->cap = capacity;
  m->bmax = max_batch;
  m->nloss = 1 << 16;
  const size_t B = (size_t)max_batch;
  int rc = hipStreamCreate(&m->st) == hipSuccess ? CRESS_OK : CRESS_EHIP;
  if (!rc) rc = dmalloc(&m->p, MLP_NPARAM);
  if (!rc) rc = dmalloc(&m->g, MLP_NPARAM);
  if (!rc) rc = dmalloc(&m->m, MLP_NPARAM);
  if (!rc) rc = dmalloc(&m->v, MLP_NPARAM);
  if (!rc) rc = dmalloc(&m->w, MLP_NPARAM);
  if (!rc) rc = dmalloc(&m->trainable, MLP_NPARAM);
  if (!rc) rc = dmalloc(&m->norm, MLP_NNORM);
  if (!rc) rc = dmalloc(&m->din, (size_t)capacity * MLP_NIN);
  if (!rc) rc = dmalloc(&m->dtgt, (size_t)capacity * MLP_NOUT);
  if (!rc) rc = dmalloc(&m->dcount, 1);
  if (!rc) rc = dmalloc(&m->X, B * MLP_INP);
  if (!rc) rc = dmalloc(&m->H1, B * MLP_H);
  if (!rc) rc = dmalloc(&m->H2, B * MLP_H);
  if (!rc) rc = dmalloc(&m->H3, B * MLP_H);
  if (!rc) rc = dmalloc(&m->dYb, B * MLP_OUTP);
  if (!rc) rc = dmalloc(&m->dZb, 2 * B * MLP_H);
  if (!rc) rc = dmalloc(&m->Z1, B * MLP_H);
  if (!rc) rc = dmalloc(&m->Z2, B * MLP_H);
  if (!rc) rc = dmalloc(&m->Z3, B * MLP_H);
  if (!rc) rc = dmalloc(&m->Y, B * MLP_OUTP);
  if (!rc) rc = dmalloc(&m->T, B * MLP_NOUT);
  if (!rc) rc = dmalloc(&m->dY, B * MLP_OUTP);
  if (!rc) rc = dmalloc(&m->D, B * MLP_H);
  if (!rc) rc = dmalloc(&m->loss, (size_t)m->nloss);
  if (rc) {
    cress_mlp_destroy(m);
    return rc;
  }
  // trainable mask: W1 rows >= 9 (padded inputs), W4 columns >= 7 and b4[>=7]
  // (padded outputs) stay zero
  std::vector<unsigned char> tr(MLP_NPARAM, 1);
  for (int i = MLP_NIN; i < MLP_INP; i++)
    for (int o = 0; o < MLP_H; o++) tr[MLP_OW1 + i * MLP_H + o] = 0;
  for (int i = 0; i < MLP_H; i++)
    for (int o = MLP_NOUT; o < MLP_OUTP; o++) tr[MLP_OW4 + i * MLP_OUTP + o] = 0;
  for (int o = MLP_NOUT; o < MLP_OUTP; o++) tr[MLP_OB4 + o] = 0;
  std::vector<float> nrm(MLP_NNORM, 0.0f);
  for (int c = 0; c < MLP_NIN; c++) {
    nrm[16 + c] = 1.0f;
    nrm[64 + c] = -INFINITY;
    nrm[80 + c] = INFINITY;
  }
  for (int o = 0; o < MLP_OUTP; o++) nrm[48 + o] = 1.0f;
  if (hipMemcpy(m->trainable, tr.data(), MLP_NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->norm, nrm.data(), MLP_NNORM * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->dcount, 0, 8) != hipSuccess || hipMemset(m->p, 0, MLP_NPARAM * 4) != hipSuccess ||
      hipMemset(m->w, 0, MLP_NPARAM * 2) != hipSuccess) {
    cress_mlp_destroy(m);
    return CRESS_EHIP;
  }
  *out = m;
  return CRESS_OK;
}

int cress_mlp_set(cress_mlp *m, const float *params, const float *norm) {
  if (!m) return CRESS_EINVAL;
  if (params) {
    if (hipMemcpy(m->p, params, MLP_NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess) return CRESS_EHIP;
    hipLaunchKernelGGL(k_mlp_cast, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->st, m->p, m->w, MLP_NPARAM);
    if (hipMemsetAsync(m->m, 0, MLP_NPARAM * 4, m->st) != hipSuccess ||
        hipMemsetAsync(m->v, 0, MLP_NPARAM * 4, m->st) != hipSuccess)
      return CRESS_EHIP;
    m->adam_t = 0;
  }
  if (norm && hipMemcpyAsync(m->norm, norm, MLP_NNORM * 4, hipMemcpyHostToDevice, m->st) != hipSuccess)
    return CRESS_EHIP;
  return hipStreamSynchronize(m->st) == hipSuccess ? CRESS_OK : CRESS_EHIP;
}

int cress_mlp_get(cress_mlp *m, float *params, float *norm) {
  if (!m) return CRESS_EINVAL;
  if (hipStreamSynchronize(m->st) != hipSuccess) return CRESS_EHIP;
  if (params && hipMemcpy(params, m->p, MLP_NPARAM * 4, hipMemcpyDeviceToHost) != hipSuccess) return CRESS_EHIP;
  if (norm && hipMemcpy(norm, m->norm, MLP_NNORM * 4, hipMemcpyDeviceToHost) != hipSuccess) return CRESS_EHIP;
  return CRESS_OK;
}

// Random initial parameters: weights uniform in +-sqrt(6/(fan_in+fan_out))
// (Glorot & Bengio 2010), biases zero, padding zero.
int cress_mlp_init(cress_mlp *m, unsigned long long seed) {
  if (!m) return CRESS_EINVAL;
  std::vector<float> p(MLP_NPARAM, 0.0f);
  unsigned long long s = seed;
  auto layer = [&](int off, int nin, int nout, int ld, int used_in, int used_out) {
    const double a = sqrt(6.0 / (double)(used_in + used_out));
    for (int i = 0; i < nin; i++)
      for (int o = 0; o < nout; o++) {
        const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
        p[off + i * ld + o] = (i < used_in && o < used_out) ? (float)(a * (2.0 * u - 1.0)) : 0.0f;
      }
  };
  layer(MLP_OW1, MLP_INP, MLP_H, MLP_H, MLP_NIN, MLP_H);
  layer(MLP_OW2, MLP_H, MLP_H, MLP_H, MLP_H, MLP_H);
  layer(MLP_OW3, MLP_H, MLP_H, MLP_H, MLP_H, MLP_H);
  layer(MLP_OW4, MLP_H, MLP_OUTP, MLP_OUTP, MLP_H, MLP_NOUT);
  return cress_mlp_set(m, p.data(), nullptr);
}

int cress_mlp_samples(cress_mlp *m, long *count) {
  if (!m || !count) return CRESS_EINVAL;
  unsigned long long c = 0;
  if (hipStreamSynchronize(m->st) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(&c, m->dcount, 8, hipMemcpyDeviceToHost) != hipSuccess)
    return CRESS_EHIP;
  *count = (long)(c < (unsigned long long)m->cap ? c : (unsigned long long)m->cap);
  return CRESS_OK;
}

int cress_mlp_clear_samples(cress_mlp *m) {
  if (!m) return CRESS_EINVAL;
  return hipMemset(m->dcount, 0, 8) == hipSuccess ? CRESS_OK : CRESS_EHIP;
}

int cress_mlp_put_samples(cress_mlp *m, const float *in, const float *tgt, long n) {
  if (!m || !in || !tgt || n < 0) return CRESS_EINVAL;
  long c = 0;
  int rc = cress_mlp_samples(m, &c);
  if (rc) return rc;
  if (c + n > m->cap) return CRESS_ETOOBIG;
  if (hipMemcpy(m->din + c * MLP_NIN, in, (size_t)n * MLP_NIN * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->dtgt + c * MLP_NOUT, tgt, (size_t)n * MLP_NOUT * 4, hipMemcpyHostToDevice) != hipSuccess)
    return CRESS_EHIP;
  const unsigned long long nc = (unsigned long long)(c + n);
  return hipMemcpy(m->dcount, &nc, 8, hipMemcpyHostToDevice) == hipSuccess ? CRESS_OK : CRESS_EHIP;
}

int cress_mlp_get_samples(cress_mlp *m, long first, long n, float *in, float *tgt) {
  if (!m || first < 0 || n < 0) return CRESS_EINVAL;
  long c = 0;
  int rc = cress_mlp_samples(m, &c);
  if (rc) return rc;
  if (first + n > c) return CRESS_ERANGE;
  if (in && hipMemcpy(in, m->din + first * MLP_NIN, (size_t)n * MLP_NIN * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return CRESS_EHIP;
  if (tgt && hipMemcpy(tgt, m->dtgt + first * MLP_NOUT, (size_t)n * MLP_NOUT * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return CRESS_EHIP;
  return CRESS_OK;
}

// Normalisation (mean, 1/standard deviation of each feature; mean and
// standard deviation of each target) and the envelope [min - margin*range,
// max + margin*range] of each feature, over the captured pairs, on the GPU.
int cress_mlp_fit_norm(cress_mlp *m, float margin) {
  if (!m || !(margin >= 0.0f)) return CRESS_EINVAL;
  long c = 0;
  int rc = cress_mlp_samples(m, &c);
  if (rc) return rc;
  if (c < 2) return CRESS_EINVAL;
  double *sums = nullptr;
  unsigned *mm = nullptr;
  rc = dmalloc(&sums, 4 * 16);
  if (!rc) rc = dmalloc(&mm, 4 * 16);
  std::vector<double> hs(64, 0.0);
  std::vector<unsigned> hm(64, 0u);
  for (int c2 = 0; c2 < 16; c2++) {
    hm[c2] = 0xffffffffu;
    hm[32 + c2] = 0xffffffffu;
  }
  if (!rc && (hipMemcpy(sums, hs.data(), 64 * 8, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(mm, hm.data(), 64 * 4, hipMemcpyHostToDevice) != hipSuccess))
    rc = CRESS_EHIP;
  if (!rc) {
    const dim3 gr((unsigned)((c + 255) / 256));
    hipLaunchKernelGGL(k_mlp_stats, gr, dim3(256), 0, m->st, m->din, c, MLP_NIN, sums, sums + 16, mm, mm + 16);
    hipLaunchKernelGGL(k_mlp_stats, gr, dim3(256), 0, m->st, m->dtgt, c, MLP_NOUT, sums + 32, sums + 48, mm + 32,
                       mm + 48);
    if (hipStreamSynchronize(m->st) != hipSuccess ||
        hipMemcpy(hs.data(), sums, 64 * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hm.data(), mm, 64 * 4, hipMemcpyDeviceToHost) != hipSuccess)
      rc = CRESS_EHIP;
  }
  hipFree(sums);
  hipFree(mm);
  if (rc) return rc;
  auto unord = [](unsigned u) {
    const unsigned b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &b, 4);
    return f;
  };
  std::vector<float> nrm(MLP_NNORM, 0.0f);
  for (int f = 0; f < MLP_NIN; f++) {
    const double mu = hs[f] / (double)c, var = hs[16 + f] / (double)c - mu * mu;
    const double sd = sqrt(var > 0.0 ? var : 0.0);
    nrm[f] = (float)mu;
    nrm[16 + f] = sd > 0.0 ? (float)(1.0 / sd) : 0.0f;
    const float lo = unord(hm[f]), hi = unord(hm[16 + f]);
    nrm[64 + f] = lo - margin * (hi - lo);
    nrm[80 + f] = hi + margin * (hi - lo);
  }
  for (int o = 0; o < MLP_OUTP; o++) nrm[48 + o] = 1.0f;
  for (int o = 0; o < MLP_NOUT; o++) {
    const double mu = hs[32 + o] / (double)c, var = hs[48 + o] / (double)c - mu * mu;
    const double sd = sqrt(var > 0.0 ? var : 0.0);
    nrm[32 + o] = (float)mu;
    nrm[48 + o] = sd > 0.0 ? (float)sd : 1.0f;
  }
  return cress_mlp_set(m, nullptr, nrm.data());
}

// The fused inference kernel on n host rows of raw features; out gets n
// rows of 7 raw tendencies.
int cress_mlp_forward(cress_mlp *m, const float *in, long n, float *out) {
  if (!m || !in || !out || n < 1 || n > m->bmax) return CRESS_EINVAL;
  int rc = upload_rows(m, in, nullptr, (int)n);
  if (rc) return rc;
  int *list = nullptr;
  unsigned *cnt = nullptr;
  float *o = nullptr;
  rc = dmalloc(&list, (size_t)n);
  if (!rc) rc = dmalloc(&cnt, 1);
  if (!rc) rc = dmalloc(&o, (size_t)n * MLP_NOUT);
  if (!rc) {
    std::vector<int> idx((size_t)n);
    for (long b = 0; b < n; b++) idx[b] = (int)b;
    const unsigned c = (unsigned)n;
    if (hipMemcpy(list, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(cnt, &c, 4, hipMemcpyHostToDevice) != hipSuccess)
      rc = CRESS_EHIP;
  }
  if (!rc) {
    MlpNet net{m->w, m->p, m->norm};
    MlpStep s{nullptr, list, m->X, cnt, o};
    hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_infer_blocks(n)), dim3(64 * MLP_WAVES), 0, m->st, net, s, n);
    std::vector<float> t((size_t)n * MLP_NOUT);
    if (hipStreamSynchronize(m->st) != hipSuccess ||
        hipMemcpy(t.data(), o, (size_t)n * MLP_NOUT * 4, hipMemcpyDeviceToHost) != hipSuccess)
      rc = CRESS_EHIP;
    for (long b = 0; b < n && !rc; b++)
      for (int k = 0; k < MLP_NOUT; k++) out[b * MLP_NOUT + k] = t[(size_t)k * n + b];
  }
  hipFree(list);
  hipFree(cnt);
  hipFree(o);
  return rc;
}

// One forward and backward pass of the training kernels on n host rows
// (raw features and targets), no update: the gradient of the mean squared
// error of the normalised outputs and the mean squared error itself.
int cress_mlp_grad(cress_mlp *m, const float *in, const float *tgt, int n, float *grad, double *loss) {
  if (!m || !in || !tgt || n < 1 || n > m->bmax) return CRESS_EINVAL;
  int rc = upload_rows(m, in, tgt, n);
  if (rc) return rc;
  if (hipMemsetAsync(m->loss, 0, 8, m->st) != hipSuccess) return CRESS_EHIP;
  fwd_bwd(m, n, m->loss);
  double l = 0.0;
  if (hipStreamSynchronize(m->st) != hipSuccess || hipGetLastError() != hipSuccess ||
      hipMemcpy(&l, m->loss, 8, hipMemcpyDeviceToHost) != hipSuccess)
    return CRESS_EHIP;
  if (grad && hipMemcpy(grad, m->g, MLP_NPARAM * 4, hipMemcpyDeviceToHost) != hipSuccess) return CRESS_EHIP;
  if (loss) *loss = l / ((double)n * MLP_NOUT);
  return CRESS_OK;
}

// steps Adam steps (learning rate lr, beta 0.9/0.999) on batches drawn from
// the captured pairs; losses (may be NULL) gets each step's mean squared
// error.
int cress_mlp_train(cress_mlp *m, int steps, int batch, float lr, unsigned seed, double *losses) {
  if (!m || steps < 1 || batch < 16 || batch > m->bmax || !(lr > 0.0f) || steps > m->nloss) return CRESS_EINVAL;
  long c = 0;
  int rc = cress_mlp_samples(m, &c);
  if (rc) return rc;
  if (c < 1) return CRESS_EINVAL;
  if (hipMemsetAsync(m->loss, 0, (size_t)steps * 8, m->st) != hipSuccess) return CRESS_EHIP;
  const float b1 = 0.9f, b2 = 0.999f;
  for (int s = 0; s < steps; s++) {
    hipLaunchKernelGGL(k_mlp_gather, dim3((batch + 255) / 256), dim3(256), 0, m->st, m->din, m->dtgt, c, m->norm,
                       seed * 2654435761u + (unsigned)s, 0, batch, m->X, m->T);
    fwd_bwd(m, batch, m->loss + s);
    m->adam_t++;
    const float c1 = 1.0f / (1.0f - powf(b1, (float)m->adam_t));
    const float c2 = 1.0f / (1.0f - powf(b2, (float)m->adam_t));
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->st, m->p, m->g, m->m, m->v,
                       m->w, MLP_NPARAM, lr, b1, b2, 1e-8f, c1, c2, m->trainable);
  }
  if (hipStreamSynchronize(m->st) != hipSuccess || hipGetLastError() != hipSuccess) return CRESS_EHIP;
  if (losses) {
    std::vector<double> l((size_t)steps);
    if (hipMemcpy(l.data(), m->loss, (size_t)steps * 8, hipMemcpyDeviceToHost) != hipSuccess) return CRESS_EHIP;
    for (int s = 0; s < steps; s++) losses[s] = l[s] / ((double)batch * MLP_NOUT);
  }
  return CRESS_OK;
}

}  // extern "C"
