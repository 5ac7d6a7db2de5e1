// Start-vector network: a small BF16 MLP (bits as +-1 -> 256 -> 256 -> 2,
// GELU) trained on ground vectors that the FP64 LOBPCG produces; it supplies
// the LOBPCG starting vector (log-amplitude and sign per bitstring). No
// counterpart in the reference, which starts from a one-hot vector
// (rqutils/sqd.py:373-398).
//
// The dense products run on the matrix cores: gemm_bf16_kernel gives every
// workgroup a 64x64 output tile, staged through LDS 32 k at a time and
// multiplied with gfx950's v_mfma_f32_16x16x32_bf16 (FP32 accumulation).
// Operand transposes are strides, so the same kernel serves the forward pass
// (X W), the input gradients (dZ W^T) and the weight gradients (H^T dZ,
// K-split over the batch with a fixed-order reduction of the partial tiles). Elementwise kernels do
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
// op(B)(k,j) = transB ? B[j*ldb+k] : B[k*ldb+j].
//
// A workgroup (4 waves in a 2 x 2 grid) computes a (32 WM) x (32 WN) tile of
// C over its share ks of the K range (ksplit shares; with ksplit > 1 the
// partial tile goes to part[ks][M][N] and gemm_reduce_kernel finishes in a
// fixed order). K advances 32 at a time through two LDS buffers: while the
// waves multiply the slices in one buffer, every thread already holds the
// next slices' global data in registers (16-byte loads along each operand's
// contiguous dimension, whatever the transpose) and stores them into the
// other buffer, so one barrier per step suffices. LDS holds both operands
// k-contiguous (row stride 40 bf16 = 80 bytes): every lane's 8 k values are
// one aligned 16-byte read. Each wave multiplies its (16 WM) x (16 WN)
// quarter as WM x WN tiles of gfx950's v_mfma_f32_16x16x32_bf16 (FP32
// accumulation). Operand layout of that instruction: lane l feeds row (or
// column) l % 16 with 8 k values, the same k for both operands (which 8 does
// not change the sum); the result holds column l % 16, rows 4 (l / 16) + j.
constexpr int kTileK = 32, kLdsStride = 40;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// 16-byte loads of an R x 32 panel, element (r, k) at p[r*ld + k] when
// kcontig, else at p[k*ld + r]; rows >= rmax and k >= kmax read as 0. vec: the
// 16-byte path is aligned (ld a multiple of 8, p 16-byte aligned).
template <int R>
__device__ inline void gemm_load_panel(uint4 (&v)[R / 64], const uint16_t* __restrict__ p,
                                       int64_t ld, bool kcontig, bool vec, int64_t r0,
                                       int64_t rmax, int64_t k0, int64_t kmax, int tid) {
#pragma unroll
  for (int t = 0; t < R / 64; ++t) {
    const int q = tid + kBlock * t;
    const int64_t gr = r0 + (kcontig ? (q >> 2) : 8 * (q % (R / 8)));
    const int64_t gk = k0 + (kcontig ? 8 * (q & 3) : q / (R / 8));
    const bool full = kcontig ? (gr < rmax && gk + 8 <= kmax) : (gk < kmax && gr + 8 <= rmax);
    if (vec && full) {
      v[t] = *reinterpret_cast<const uint4*>(kcontig ? p + gr * ld + gk : p + gk * ld + gr);
    } else {
      uint16_t e[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int64_t r = kcontig ? gr : gr + j, k = kcontig ? gk + j : gk;
        e[j] = (r < rmax && k < kmax) ? p[kcontig ? r * ld + k : k * ld + r] : (uint16_t)0;
      }
      v[t] = make_uint4(e[0] | ((uint32_t)e[1] << 16), e[2] | ((uint32_t)e[3] << 16),
                        e[4] | ((uint32_t)e[5] << 16), e[6] | ((uint32_t)e[7] << 16));
    }
  }
}

// Stores a panel loaded by gemm_load_panel as lds[r][k] (stride kLdsStride).
template <int R>
__device__ inline void gemm_store_panel(const uint4 (&v)[R / 64], uint16_t* lds, bool kcontig,
                                        int tid) {
#pragma unroll
  for (int t = 0; t < R / 64; ++t) {
    const int q = tid + kBlock * t;
    if (kcontig) {
      *reinterpret_cast<uint4*>(&lds[(q >> 2) * kLdsStride + 8 * (q & 3)]) = v[t];
    } else {
      const int r = 8 * (q % (R / 8)), k = q / (R / 8);
      const uint32_t w[4] = {v[t].x, v[t].y, v[t].z, v[t].w};
#pragma unroll
      for (int j = 0; j < 8; ++j)
        lds[(r + j) * kLdsStride + k] = (uint16_t)(w[j >> 1] >> (16 * (j & 1)));
    }
  }
}

template <int WM, int WN>
__global__ void __launch_bounds__(kBlock)
gemm_bf16_kernel(int64_t M, int64_t N, int64_t K, const uint16_t* __restrict__ A, int64_t lda,
                 int transA, const uint16_t* __restrict__ B, int64_t ldb, int transB,
                 float* __restrict__ C, int64_t ldc, const float* __restrict__ bias, int ksplit,
                 float* __restrict__ part) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  __shared__ __align__(16) uint16_t As[2][BM * kLdsStride];   // [buffer][i][k]
  __shared__ __align__(16) uint16_t Bs[2][BN * kLdsStride];   // [buffer][j][k]
  const int64_t tiles_n = (N + BN - 1) / BN;
  const int64_t tiles = ((M + BM - 1) / BM) * tiles_n;
  const int64_t blk = blockIdx.x;
  const int ks = (int)(blk / tiles);
  const int64_t t = blk % tiles;
  const int64_t m0 = (t / tiles_n) * BM, n0 = (t % tiles_n) * BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 16 * WM, wn = (wave & 1) * 16 * WN;
  const int64_t ksteps = (K + kTileK - 1) / kTileK;
  const int64_t per = (ksteps + ksplit - 1) / ksplit;
  const int64_t k_lo = (int64_t)ks * per * kTileK;
  const int64_t k_hi = min(K, ((int64_t)ks + 1) * per * kTileK);
  // op(A) rows are k-contiguous in memory unless transA; op(B) columns unless !transB
  const bool a_kc = !transA, b_kc = transB != 0;
  const bool a_vec = (lda % 8 == 0) && ((uintptr_t)A % 16 == 0);
  const bool b_vec = (ldb % 8 == 0) && ((uintptr_t)B % 16 == 0);
  f32x4 acc[WM][WN];
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  uint4 ra[BM / 64], rb[BN / 64];
  if (k_lo < k_hi) {
    gemm_load_panel<BM>(ra, A, lda, a_kc, a_vec, m0, M, k_lo, k_hi, tid);
    gemm_load_panel<BN>(rb, B, ldb, b_kc, b_vec, n0, N, k_lo, k_hi, tid);
    gemm_store_panel<BM>(ra, As[0], a_kc, tid);
    gemm_store_panel<BN>(rb, Bs[0], b_kc, tid);
  }
  __syncthreads();
  const int kq = 8 * (lane >> 4);
  int cur = 0;
  for (int64_t k0 = k_lo; k0 < k_hi; k0 += kTileK) {
    const bool more = k0 + kTileK < k_hi;
    if (more) {
      gemm_load_panel<BM>(ra, A, lda, a_kc, a_vec, m0, M, k0 + kTileK, k_hi, tid);
      gemm_load_panel<BN>(rb, B, ldb, b_kc, b_vec, n0, N, k0 + kTileK, k_hi, tid);
    }
    bf16x8 av[WM], bv[WN];
#pragma unroll
    for (int a = 0; a < WM; ++a)
      av[a] = *reinterpret_cast<const bf16x8*>(
          &As[cur][(wm + 16 * a + (lane & 15)) * kLdsStride + kq]);
#pragma unroll
    for (int b = 0; b < WN; ++b)
      bv[b] = *reinterpret_cast<const bf16x8*>(
          &Bs[cur][(wn + 16 * b + (lane & 15)) * kLdsStride + kq]);
#pragma unroll
    for (int a = 0; a < WM; ++a)
#pragma unroll
      for (int b = 0; b < WN; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    if (more) {
      // the other buffer was last read before the previous step's barrier
      gemm_store_panel<BM>(ra, As[cur ^ 1], a_kc, tid);
      gemm_store_panel<BN>(rb, Bs[cur ^ 1], b_kc, tid);
    }
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) {
      const int64_t cj = n0 + wn + 16 * b + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t ci = m0 + wm + 16 * a + 4 * (lane >> 4) + j;
        if (ci < M && cj < N) {
          if (ksplit == 1)
            C[ci * ldc + cj] = acc[a][b][j] + (bias ? bias[cj] : 0.f);
          else
            part[((int64_t)ks * M + ci) * N + cj] = acc[a][b][j];
        }
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

// out[j] = sum_i x[i*cols + j]. A workgroup of 1024 threads owns cw columns
// (cw = 16, or the next power of two >= cols when cols < 16) and splits the
// rows into 1024 / cw interleaved groups; each thread sums its group in FP64
// (four partial sums in a fixed order) and the groups are combined by a fixed
// tree in LDS, so the result does not depend on timing.
constexpr int kColsumThreads = 1024;

__global__ void __launch_bounds__(kColsumThreads)
colsum_kernel(int64_t n, int64_t cols, int cw, const float* __restrict__ x,
              float* __restrict__ out) {
  __shared__ double part[kColsumThreads];
  const int tid = threadIdx.x;
  const int c = tid % cw, g = tid / cw, groups = kColsumThreads / cw;
  const int64_t j = (int64_t)blockIdx.x * cw + c;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (j < cols) {
    int64_t i = g;
    for (; i + 3 * groups < n; i += 4 * groups) {
#pragma unroll
      for (int u = 0; u < 4; ++u) s[u] += x[(i + u * groups) * cols + j];
    }
    if (i < n) s[0] += x[i * cols + j];
    if (i + groups < n) s[1] += x[(i + groups) * cols + j];
    if (i + 2 * groups < n) s[2] += x[(i + 2 * groups) * cols + j];
  }
  part[tid] = (s[0] + s[1]) + (s[2] + s[3]);
  __syncthreads();
  for (int half = groups / 2; half >= 1; half /= 2) {
    if (g < half) part[tid] += part[tid + half * cw];
    __syncthreads();
  }
  if (g == 0 && j < cols) out[j] = (float)part[tid];
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
  // 128 x 128 tiles where they still give at least four workgroups per CU, else 64 x 64
  const int64_t big = ceil_div(M, 128) * ceil_div(N, 128) * ksplit;
  const int64_t blocks = big >= 1024 ? big : ceil_div(M, 64) * ceil_div(N, 64) * ksplit;
  SQDK_REQUIRE(blocks < (1ll << 31), "GEMM grid too large");
  hipStream_t s = as_stream(stream);
  if (big >= 1024)
    gemm_bf16_kernel<4, 4><<<(unsigned)blocks, kBlock, 0, s>>>(
        M, N, K, A, lda, trans_a, B, ldb, trans_b, C, ldc, bias, ksplit, work);
  else
    gemm_bf16_kernel<2, 2><<<(unsigned)blocks, kBlock, 0, s>>>(
        M, N, K, A, lda, trans_a, B, ldb, trans_b, C, ldc, bias, ksplit, work);
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
  int cw = 1;
  while (cw < 16 && cw < cols) cw *= 2;
  colsum_kernel<<<(unsigned)ceil_div(cols, cw), kColsumThreads, 0, as_stream(stream)>>>(
      n, cols, cw, x, out);
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
