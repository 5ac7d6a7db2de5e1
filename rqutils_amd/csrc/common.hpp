// Shared helpers of the SQD kernel library: status/error plumbing, complex
// arithmetic on HIP vector types, compensated (two-sum) accumulation and
// wavefront/block reductions for gfx950 (64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/sqdk.h"

namespace sqdk {

// ---- error plumbing ------------------------------------------------------

std::string& last_error_slot();

inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  last_error_slot() = buf;
  return code;
}

#define SQDK_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::sqdk::fail(SQDK_EHIP, "%s failed: %s", #expr,                  \
                          hipGetErrorString(e_));                             \
  } while (0)

#define SQDK_LAUNCHED(name)                                                   \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess)                                                     \
      return ::sqdk::fail(SQDK_EHIP, "launch of %s failed: %s", name,         \
                          hipGetErrorString(e_));                             \
  } while (0)

#define SQDK_TRY(expr)                                                        \
  do {                                                                        \
    int s_ = (expr);                                                          \
    if (s_ != SQDK_OK) return s_;                                             \
  } while (0)

#define SQDK_REQUIRE(cond, ...)                                               \
  do {                                                                        \
    if (!(cond)) return ::sqdk::fail(SQDK_EINVAL, __VA_ARGS__);               \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

constexpr int kWave = 64;

// ---- complex helpers -----------------------------------------------------

__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ inline double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline bool czero(double2 a) { return a.x == 0.0 && a.y == 0.0; }

// Storage traits: vectors live in double2 (complex128) or float2 (complex64);
// loads widen to double2, stores round once.
template <typename VT> struct Vec;
template <> struct Vec<double2> {
  __device__ static inline double2 load(const double2* p, int64_t i) { return p[i]; }
  __device__ static inline void store(double2* p, int64_t i, double2 v) { p[i] = v; }
  // the value as a store and a load would give it back
  __device__ static inline double2 round(double2 v) { return v; }
};
template <> struct Vec<float2> {
  __device__ static inline double2 load(const float2* p, int64_t i) {
    float2 f = p[i];
    return make_double2((double)f.x, (double)f.y);
  }
  __device__ static inline void store(float2* p, int64_t i, double2 v) {
    p[i] = make_float2((float)v.x, (float)v.y);
  }
  __device__ static inline double2 round(double2 v) {
    return make_double2((double)(float)v.x, (double)(float)v.y);
  }
};

// ---- compensated accumulation ---------------------------------------------

// Knuth's TwoSum: s_new + err == s + x exactly in round-to-nearest; the error
// term is accumulated separately. Needs IEEE evaluation order (no fast-math).
__device__ inline void two_sum(float& s, float& e, float x) {
  const float t = s + x;
  const float bp = t - s;
  const float err = (s - (t - bp)) + (x - bp);
  s = t;
  e += err;
}

// Partial sum of the products of one batch of terms (the groups of one class
// of the matvec): p += c v with two FMAs per part. A zero coefficient leaves
// p unchanged bit for bit (p is never -0: it starts at +0 and only a sum of
// two -0 gives -0).
__device__ inline void part_fma(float2& p, float cr, float ci, float2 v) {
  p.x = __builtin_fmaf(cr, v.x, p.x);
  p.x = __builtin_fmaf(-ci, v.y, p.x);
  p.y = __builtin_fmaf(cr, v.y, p.y);
  p.y = __builtin_fmaf(ci, v.x, p.y);
}
__device__ inline void part_fma(double2& p, double cr, double ci, double2 v) {
  p.x = __builtin_fma(cr, v.x, p.x);
  p.x = __builtin_fma(-ci, v.y, p.x);
  p.y = __builtin_fma(cr, v.y, p.y);
  p.y = __builtin_fma(ci, v.x, p.y);
}

// Complex accumulator: FP64 for complex128 storage, FP32 two-sum for complex64.
// `fma` adds one product; `fold` adds a partial (part_fma) and clears it.
template <typename VT> struct Acc;
template <> struct Acc<double2> {
  double sr, si;
  __device__ inline Acc() : sr(0.0), si(0.0) {}
  __device__ inline void fma(double cr, double ci, double2 v) {
    sr += cr * v.x - ci * v.y;
    si += cr * v.y + ci * v.x;
  }
  __device__ inline void fold(double2& p) {
    sr += p.x;
    si += p.y;
    p = make_double2(0.0, 0.0);
  }
  __device__ inline double2 result() const { return make_double2(sr, si); }
};
template <> struct Acc<float2> {
  float sr, er, si, ei;
  __device__ inline Acc() : sr(0.f), er(0.f), si(0.f), ei(0.f) {}
  // Each part of the complex product c v is formed in FP32 with one FMA over
  // one rounded product; the running sums are compensated (two-sum).
  __device__ inline void fma(float cr, float ci, float2 v) {
    const float pr = __builtin_fmaf(cr, v.x, -(ci * v.y));
    const float pi = __builtin_fmaf(cr, v.y, ci * v.x);
    two_sum(sr, er, pr);
    two_sum(si, ei, pi);
  }
  // The partial enters the running sums through one two-sum per part.
  __device__ inline void fold(float2& p) {
    two_sum(sr, er, p.x);
    two_sum(si, ei, p.y);
    p = make_float2(0.f, 0.f);
  }
  __device__ inline float2 result() const { return make_float2(sr + er, si + ei); }
};

// ---- reductions --------------------------------------------------------------

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int m = kWave / 2; m > 0; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

constexpr int pow2_at_least(int v) { return v <= 1 ? 1 : 2 * pow2_at_least((v + 1) / 2); }

// Block-wide sum of NV doubles in a fixed order; result valid in thread 0.
// `red` must hold NV * (blockDim.x / 64) doubles of shared memory.
//
// Within a wave the NV values (padded to P, a power of two) are summed by
// recursive halving: at the exchange over lane bit m a lane keeps half of
// its values (the lower half if that bit is 0) and receives its partner's
// share of that half, so P/2 + P/4 + ... + 1 shuffles leave lane L with the
// wave's total of value L / (64 / P), followed by log2(64 / P) plain
// exchanges; about P shuffles instead of 6 NV. Lanes L with L % (64 / P) == 0
// then store their value, and thread 0 adds the waves' values in wave order.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* red) {
  static_assert(NV >= 1 && NV <= kWave, "block_sum handles 1..64 values");
  constexpr int P = pow2_at_least(NV);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  double w[P];
#pragma unroll
  for (int k = 0; k < P; ++k) w[k] = k < NV ? v[k] : 0.0;
  int m = kWave / 2;
#pragma unroll
  for (int h = P / 2; h >= 1; h /= 2, m /= 2) {
    const bool upper = (lane & m) != 0;
#pragma unroll
    for (int k = 0; k < h; ++k) {
      const double send = upper ? w[k] : w[k + h];
      const double keep = upper ? w[k + h] : w[k];
      w[k] = keep + __shfl_xor(send, m);
    }
  }
  for (; m > 0; m /= 2) w[0] += __shfl_xor(w[0], m);
  __syncthreads();
  constexpr int kLanesPer = kWave / P;
  if (lane % kLanesPer == 0) {
    const int k = lane / kLanesPer;
    if (k < NV) red[k * nw + wid] = w[0];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int q = 0; q < nw; ++q) s += red[k * nw + q];
      v[k] = s;
    }
  }
}

// ---- subspace search -----------------------------------------------------------

// First index in keys[lo, hi) whose key is >= t.
template <typename KT>
__device__ inline int64_t lower_bound_range(const KT* __restrict__ keys, int64_t lo,
                                            int64_t hi, KT t) {
  int64_t len = hi - lo;
  int64_t base = lo;
  while (len > 0) {
    const int64_t half = len >> 1;
    if (keys[base + half] < t) {
      base += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return base;
}

__device__ inline bool popcount_present(uint64_t t, const uint64_t pc[2]) {
  const int w = __popcll(t);
  return w < 64 ? ((pc[0] >> w) & 1ull) != 0 : (pc[1] & 1ull) != 0;
}

// Index of key t in the subspace (global-memory path through the bucket
// directory), or -1.
__device__ inline int32_t find_key_global(const uint64_t* __restrict__ keys,
                                          const uint32_t* __restrict__ bstart,
                                          int nbits, int shift, uint64_t t) {
  if (nbits < 64 && (t >> nbits) != 0) return -1;
  const uint64_t b = t >> shift;
  const int64_t lo = bstart[b];
  const int64_t hi = bstart[b + 1];
  const int64_t i = lower_bound_range<uint64_t>(keys, lo, hi, t);
  return (i < hi && keys[i] == t) ? (int32_t)i : -1;
}

// Coefficient of one X group at a bitstring:
// sum_k c_k (-1)^{popcount(z_k & s)} for terms t0 <= k < t1.
template <typename CT>
__device__ inline void group_coef(const uint64_t* __restrict__ zmask,
                                  const CT* __restrict__ coef, int t0, int t1,
                                  uint64_t s, decltype(CT::x)& cr,
                                  decltype(CT::x)& ci) {
  cr = 0;
  ci = 0;
  for (int k = t0; k < t1; ++k) {
    const uint64_t z = zmask[k];
    const CT c = coef[k];
    const bool odd = (__popcll(z & s) & 1) != 0;
    cr += odd ? -c.x : c.x;
    ci += odd ? -c.y : c.y;
  }
}

}  // namespace sqdk
