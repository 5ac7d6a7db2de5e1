// Start-vector network: a small BF16 MLP (bits as +-1 -> 256 -> 256 -> 2,
// GELU) trained on ground vectors that the FP64 LOBPCG produces; it supplies
// the LOBPCG starting vector (log-amplitude and sign per bitstring). No
// counterpart in the reference, which starts from a one-hot vector
// (rqutils/sqd.py:373-398).
//
// The dense products run on the matrix cores: gemm_bf16_kernel gives every
// wavefront one 16x16 output tile and accumulates in FP32 with
// v_mfma_f32_16x16x16_bf16 over K in steps of 16. Operand transposes are
// strides, so the same kernel serves the forward pass (X W), the input
// gradients (dZ W^T) and the weight gradients (H^T dZ, K-split over the batch
// with a fixed-order reduction of the partial tiles). Elementwise kernels do
// the encoding, GELU and its derivative, the loss and Adam.
#include "internal.hpp"

namespace sqdk {

namespace {

constexpr int kBlock = 256;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (NaN stays NaN)
__device__ inline uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
  const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}

__device__ inline float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// C = op(A) op(B) (+ bias) ; op(A)(i,k) = transA ? A[k*lda+i] : A[i*lda+k],
// op(B)(k,j) = transB ? B[j*ldb+k] : B[k*ldb+j]. Wave w handles tile
// (tm, tn) and K range ks of ksplit. With ksplit > 1 partial tiles go to
// part[ks][M][N] and gemm_reduce_kernel finishes.
__global__ void __launch_bounds__(kBlock)
gemm_bf16_kernel(int64_t M, int64_t N, int64_t K, const uint16_t* __restrict__ A, int64_t lda,
                 int transA, const uint16_t* __restrict__ B, int64_t ldb, int transB,
                 float* __restrict__ C, int64_t ldc, const float* __restrict__ bias, int ksplit,
                 float* __restrict__ part) {
  const int64_t tiles_n = (N + 15) / 16;
  const int64_t tiles = ((M + 15) / 16) * tiles_n;
  const int64_t wave = (int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
  if (wave >= tiles * ksplit) return;
  const int ks = (int)(wave / tiles);
  const int64_t t = wave % tiles;
  const int64_t tm = t / tiles_n, tn = t % tiles_n;
  const int lane = threadIdx.x & 63;
  const int64_t ksteps = (K + 15) / 16;
  const int64_t per = (ksteps + ksplit - 1) / ksplit;
  const int64_t k_lo = ks * per * 16;
  const int64_t k_hi = min(K, (ks + 1) * per * 16);
  const int64_t ai = tm * 16 + (lane & 15);   // row of op(A) this lane feeds
  const int64_t bj = tn * 16 + (lane & 15);   // column of op(B) this lane feeds
  const int kq = (lane >> 4) * 4;             // first of this lane's 4 k values
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int64_t k0 = k_lo; k0 < k_hi; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t k = k0 + kq + j;
      uint16_t va = 0, vb = 0;
      if (k < k_hi) {
        if (ai < M) va = transA ? A[k * lda + ai] : A[ai * lda + k];
        if (bj < N) vb = transB ? B[bj * ldb + k] : B[k * ldb + bj];
      }
      a[j] = (short)va;
      b[j] = (short)vb;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  // D layout: lane holds column (lane & 15), rows 4 * (lane >> 4) + j.
  const int64_t cj = tn * 16 + (lane & 15);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t ci = tm * 16 + (lane >> 4) * 4 + j;
    if (ci < M && cj < N) {
      if (ksplit == 1)
        C[ci * ldc + cj] = acc[j] + (bias ? bias[cj] : 0.f);
      else
        part[((int64_t)ks * M + ci) * N + cj] = acc[j];
    }
  }
}

__global__ void __launch_bounds__(kBlock)
gemm_reduce_kernel(int64_t M, int64_t N, int ksplit, const float* __restrict__ part,
                   float* __restrict__ C, int64_t ldc, const float* __restrict__ bias) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= M * N) return;
  const int64_t i = idx / N, j = idx % N;
  float s = bias ? bias[j] : 0.f;
  for (int k = 0; k < ksplit; ++k) s += part[(int64_t)k * M * N + idx];
  C[i * ldc + j] = s;
}

// out[i][q] = 1 - 2 bit_q(key_i) for q < nq, extra[i][q - nq] for the next
// n_extra columns, 0 up to kpad (bf16).
__global__ void __launch_bounds__(kBlock)
encode_kernel(const uint64_t* __restrict__ keys, int64_t n, int nq, const float* __restrict__ extra,
              int n_extra, int kpad, uint16_t* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n * kpad) return;
  const int64_t i = idx / kpad;
  const int q = (int)(idx % kpad);
  float v = 0.f;
  if (q < nq) v = ((keys[i] >> q) & 1ull) ? -1.f : 1.f;
  else if (q < nq + n_extra) v = extra[i * n_extra + (q - nq)];
  out[idx] = f2bf(v);
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }

__device__ inline float gelu_grad(float z) {
  const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// op 0: h = bf16(gelu(z)); op 1: dz = g * gelu'(z) -> f32 and bf16; op 2: bf16(z)
__global__ void __launch_bounds__(kBlock)
act_kernel(int op, int64_t n, const float* __restrict__ z, const float* __restrict__ g,
           uint16_t* __restrict__ out16, float* __restrict__ out32) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float r;
  if (op == 0) r = gelu(z[i]);
  else if (op == 1) r = g[i] * gelu_grad(z[i]);
  else r = z[i];
  if (out16) out16[i] = f2bf(r);
  if (out32) out32[i] = r;
}

// Weighted loss over the outputs y (log-amplitude y0, sign logit y1) and the
// targets (a, p, w): L = sum_i w_i [(y0 - a_i)^2 + lambda * BCE(sigmoid(y1), p_i)] / sum_i w_i.
// Writes dL/dy (f32 and bf16, n x 2) and per-block partial sums of L and W.
__global__ void __launch_bounds__(kBlock)
loss_kernel(int64_t n, const float* __restrict__ y, int64_t ldy, const float* __restrict__ tgt,
            double wsum, double lambda, float* __restrict__ dy32, uint16_t* __restrict__ dy16,
            double* __restrict__ partial) {
  __shared__ double red[kBlock / kWave];
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const float y0 = y[i * ldy], y1 = y[i * ldy + 1];
    const float a = tgt[3 * i], p = tgt[3 * i + 1], w = tgt[3 * i + 2];
    const float sg = 1.f / (1.f + expf(-y1));
    const float d0 = y0 - a;
    // log(1 + e^-|y|) form of the cross entropy, stable for large |y1|
    const float bce = fmaxf(y1, 0.f) - y1 * p + log1pf(expf(-fabsf(y1)));
    acc[0] += (double)w * ((double)d0 * d0 + lambda * bce);
    const float g0 = (float)(2.0 * w * d0 / wsum);
    const float g1 = (float)(lambda * w * (sg - p) / wsum);
    dy32[2 * i] = g0;
    dy32[2 * i + 1] = g1;
    dy16[2 * i] = f2bf(g0);
    dy16[2 * i + 1] = f2bf(g1);
  }
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc[0] / wsum;
}

// out[j] = sum_i x[i*cols + j] over rows, in row order (one thread per column).
__global__ void __launch_bounds__(kBlock)
colsum_kernel(int64_t n, int64_t cols, const float* __restrict__ x, float* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i * cols + j];
  out[j] = (float)s;
}

// Adam (Kingma & Ba) on FP32 master weights, refreshing their bf16 copy.
__global__ void __launch_bounds__(kBlock)
adam_kernel(int64_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
            float* __restrict__ v, uint16_t* __restrict__ w16, float lr, float b1, float b2,
            float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = f2bf(wi);
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, kBlock)); }

// ---- fused inference -----------------------------------------------------------
//
// y = the network's outputs for n bitstrings in one kernel: encoding, three
// layers, biases and GELUs, no intermediate in HBM. The layers are computed
// transposed, H^T = W^T X^T, so that the MFMA accumulator layout of one layer
// (lane l holds feature 4 (l >> 4) + j of row l & 15) is exactly the B
// operand the next layer needs: activations stay in registers. A wave takes 16
// rows at a time; W2^T (256 x 256 bf16) sits in LDS for the workgroup's whole
// run, each row padded by 16 bytes so that the 16 feature rows a fragment
// read touches fall in different banks; W1^T and W3^T (small) are read
// through the caches. Same bf16 roundings, GELU and k order as the training
// path (encode_kernel, gemm_bf16_kernel, act_kernel).
constexpr int kInferH = 256;
constexpr int kInferThreads = 512;
constexpr int kW2Stride = kInferH + 8;   // bf16 per LDS row of W2^T

__global__ void __launch_bounds__(kInferThreads)
mlp_infer_kernel(const uint64_t* __restrict__ keys, int64_t n, int nq,
                 const float* __restrict__ extra, int n_extra, int kpad,
                 const uint16_t* __restrict__ w1t, const float* __restrict__ b1,
                 const uint16_t* __restrict__ w2t, const float* __restrict__ b2,
                 const uint16_t* __restrict__ w3t, const float* __restrict__ b3,
                 float* __restrict__ y) {
  extern __shared__ __align__(16) uint16_t lw2[];    // [256][kW2Stride]
  __shared__ float lb1[kInferH], lb2[kInferH];
  const int tid = threadIdx.x;
  for (int i = tid; i < kInferH * kInferH / 4; i += kInferThreads) {
    const int f = i / (kInferH / 4), k4 = i % (kInferH / 4);
    *reinterpret_cast<uint2*>(lw2 + f * kW2Stride + 4 * k4) =
        reinterpret_cast<const uint2*>(w2t)[i];
  }
  for (int i = tid; i < kInferH; i += kInferThreads) {
    lb1[i] = b1[i];
    lb2[i] = b2[i];
  }
  __syncthreads();
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int r = lane & 15;          // row of the 16-row block (B / D column)
  const int kq = 4 * (lane >> 4);   // this lane's first k (B) / feature (D) within a step
  const int64_t nblk = (n + 15) / 16;
  const int ks1 = kpad / 16;
  for (int64_t blk = (int64_t)blockIdx.x * (kInferThreads / 64) + wave; blk < nblk;
       blk += (int64_t)gridDim.x * (kInferThreads / 64)) {
    const int64_t row = blk * 16 + r;
    const bool valid = row < n;
    const uint64_t key = valid ? keys[row] : 0ull;
    f32x4 acc[16];
    bf16x4 hb[16];
    // layer 1: K = kpad (bits as +-1, then the extra features, then zeros)
#pragma unroll
    for (int ft = 0; ft < 16; ++ft) acc[ft] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < ks1; ++ks) {
      bf16x4 xb;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = 16 * ks + kq + j;
        float v = 0.f;
        if (valid) {
          if (k < nq) v = ((key >> k) & 1ull) ? -1.f : 1.f;
          else if (k < nq + n_extra) v = extra[row * n_extra + (k - nq)];
        }
        xb[j] = (short)f2bf(v);
      }
#pragma unroll
      for (int ft = 0; ft < 16; ++ft) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(w1t + (int64_t)(16 * ft + r) * kpad +
                                                           16 * ks + kq);
        acc[ft] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, xb, acc[ft], 0, 0, 0);
      }
    }
#pragma unroll
    for (int ft = 0; ft < 16; ++ft) {
#pragma unroll
      for (int j = 0; j < 4; ++j) hb[ft][j] = (short)f2bf(gelu(acc[ft][j] + lb1[16 * ft + kq + j]));
    }
    // layer 2: 256 x 256 from LDS
#pragma unroll
    for (int ft = 0; ft < 16; ++ft) acc[ft] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
      // one k step's 16 fragments live at a time (keeps the registers in bounds)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int ft = 0; ft < 16; ++ft) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(lw2 + (16 * ft + r) * kW2Stride +
                                                           16 * ks + kq);
        acc[ft] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, hb[ks], acc[ft], 0, 0, 0);
      }
    }
#pragma unroll
    for (int ft = 0; ft < 16; ++ft) {
#pragma unroll
      for (int j = 0; j < 4; ++j) hb[ft][j] = (short)f2bf(gelu(acc[ft][j] + lb2[16 * ft + kq + j]));
    }
    // layer 3: 2 outputs (W3^T padded to 16 rows of zeros)
    f32x4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(w3t + r * kInferH + 16 * ks + kq);
      o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, hb[ks], o, 0, 0, 0);
    }
    if (kq == 0 && valid) {
      y[row * 2] = o[0] + b3[0];
      y[row * 2 + 1] = o[1] + b3[1];
    }
  }
}

}  // namespace

}  // namespace sqdk

using namespace sqdk;

extern "C" {

int sqdk_gemm_bf16_workspace(int64_t M, int64_t N, int32_t ksplit, size_t* bytes) {
  SQDK_REQUIRE(bytes, "NULL output");
  SQDK_REQUIRE(M >= 0 && N >= 0 && ksplit >= 1, "bad GEMM shape");
  *bytes = ksplit > 1 ? sizeof(float) * (size_t)ksplit * (size_t)M * (size_t)N : 0;
  return SQDK_OK;
}

int sqdk_gemm_bf16(int64_t M, int64_t N, int64_t K, const uint16_t* A, int64_t lda,
                   int32_t trans_a, const uint16_t* B, int64_t ldb, int32_t trans_b, float* C,
                   int64_t ldc, const float* bias, int32_t ksplit, float* work,
                   size_t work_bytes, void* stream) {
  SQDK_REQUIRE(A && B && C, "NULL matrix");
  SQDK_REQUIRE(M >= 1 && N >= 1 && K >= 1, "empty GEMM %lld x %lld x %lld", (long long)M,
               (long long)N, (long long)K);
  SQDK_REQUIRE(ksplit >= 1 && ksplit <= 4096, "ksplit %d out of range", ksplit);
  SQDK_REQUIRE(ksplit == 1 || (work && work_bytes >= sizeof(float) * (size_t)ksplit * M * N),
               "GEMM workspace too small for ksplit %d", ksplit);
  const int64_t waves = ((M + 15) / 16) * ((N + 15) / 16) * ksplit;
  const int64_t blocks = ceil_div(waves, kBlock / 64);
  SQDK_REQUIRE(blocks < (1ll << 31), "GEMM grid too large");
  hipStream_t s = as_stream(stream);
  gemm_bf16_kernel<<<(unsigned)blocks, kBlock, 0, s>>>(M, N, K, A, lda, trans_a, B, ldb, trans_b,
                                                       C, ldc, bias, ksplit, work);
  SQDK_LAUNCHED("gemm_bf16_kernel");
  if (ksplit > 1) {
    gemm_reduce_kernel<<<grid_for(M * N), kBlock, 0, s>>>(M, N, ksplit, work, C, ldc, bias);
    SQDK_LAUNCHED("gemm_reduce_kernel");
  }
  return SQDK_OK;
}

int sqdk_mlp_infer(const uint64_t* keys, int64_t n, int32_t num_qubits, const float* extra,
                   int32_t n_extra, int32_t kpad, const uint16_t* w1t, const float* b1,
                   const uint16_t* w2t, const float* b2, const uint16_t* w3t, const float* b3,
                   float* y, void* stream) {
  SQDK_REQUIRE(keys && w1t && b1 && w2t && b2 && w3t && b3 && y, "NULL pointer argument");
  SQDK_REQUIRE(n_extra >= 0 && (n_extra == 0 || extra), "extra features missing");
  SQDK_REQUIRE(n >= 1 && num_qubits >= 1 && num_qubits <= 64, "bad inference shape");
  SQDK_REQUIRE(kpad % 16 == 0 && kpad >= num_qubits + n_extra && kpad <= 128,
               "kpad %d must be a multiple of 16 covering the inputs (at most 128)", kpad);
  const size_t lds = sizeof(uint16_t) * kInferH * kW2Stride;
  SQDK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(mlp_infer_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int64_t blocks = std::min<int64_t>(ceil_div(ceil_div(n, 16), kInferThreads / 64), 256);
  mlp_infer_kernel<<<(unsigned)blocks, kInferThreads, lds, as_stream(stream)>>>(
      keys, n, num_qubits, extra, n_extra, kpad, w1t, b1, w2t, b2, w3t, b3, y);
  SQDK_LAUNCHED("mlp_infer_kernel");
  return SQDK_OK;
}

int sqdk_mlp_encode(const uint64_t* keys, int64_t n, int32_t num_qubits, const float* extra,
                    int32_t n_extra, int32_t kpad, uint16_t* out, void* stream) {
  SQDK_REQUIRE(keys && out, "NULL pointer argument");
  SQDK_REQUIRE(n_extra >= 0 && (n_extra == 0 || extra), "extra features missing");
  SQDK_REQUIRE(n >= 1 && num_qubits >= 1 && num_qubits <= 64 && kpad >= num_qubits + n_extra,
               "bad encode shape");
  encode_kernel<<<grid_for(n * kpad), kBlock, 0, as_stream(stream)>>>(keys, n, num_qubits, extra,
                                                                     n_extra, kpad, out);
  SQDK_LAUNCHED("encode_kernel");
  return SQDK_OK;
}

int sqdk_mlp_act(int32_t op, int64_t n, const float* z, const float* g, uint16_t* out_bf16,
                 float* out_f32, void* stream) {
  SQDK_REQUIRE(z && (out_bf16 || out_f32), "NULL pointer argument");
  SQDK_REQUIRE(op >= 0 && op <= 2, "unknown activation op %d", op);
  SQDK_REQUIRE(op != 1 || g, "gradient input missing");
  SQDK_REQUIRE(n >= 1, "empty activation");
  act_kernel<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(op, n, z, g, out_bf16, out_f32);
  SQDK_LAUNCHED("act_kernel");
  return SQDK_OK;
}

int sqdk_mlp_loss(int64_t n, const float* y, int64_t ldy, const float* target, double wsum,
                  double lambda, float* dy_f32, uint16_t* dy_bf16, double* partial,
                  int32_t nblocks, void* stream) {
  SQDK_REQUIRE(y && target && dy_f32 && dy_bf16 && partial, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1 && nblocks >= 1 && wsum > 0.0, "bad loss arguments");
  loss_kernel<<<(unsigned)nblocks, kBlock, 0, as_stream(stream)>>>(n, y, ldy, target, wsum, lambda,
                                                                   dy_f32, dy_bf16, partial);
  SQDK_LAUNCHED("loss_kernel");
  return SQDK_OK;
}

int sqdk_colsum(int64_t n, int64_t cols, const float* x, float* out, void* stream) {
  SQDK_REQUIRE(x && out && n >= 1 && cols >= 1, "bad colsum arguments");
  colsum_kernel<<<grid_for(cols), kBlock, 0, as_stream(stream)>>>(n, cols, x, out);
  SQDK_LAUNCHED("colsum_kernel");
  return SQDK_OK;
}

int sqdk_adam(int64_t n, float* w, const float* g, float* m, float* v, uint16_t* w_bf16,
              float lr, float beta1, float beta2, float eps, int32_t step, void* stream) {
  SQDK_REQUIRE(w && g && m && v && n >= 1 && step >= 1, "bad Adam arguments");
  const float c1 = 1.f - powf(beta1, (float)step);
  const float c2 = 1.f - powf(beta2, (float)step);
  adam_kernel<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(n, w, g, m, v, w_bf16, lr, beta1,
                                                             beta2, eps, c1, c2);
  SQDK_LAUNCHED("adam_kernel");
  return SQDK_OK;
}

}  // extern "C"
