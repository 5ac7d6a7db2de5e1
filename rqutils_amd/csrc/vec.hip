// Vector kernels of the LOBPCG: multi-pair inner products with FP64
// accumulation (the jnp.sum(conj(a) * b) reductions of
// rqutils/ground_locg.py:214-421), the fused Gram sums, linear combinations
// with device-resident coefficients, the fused subspace update and
// precision conversion. All reductions use a fixed grid and a fixed
// combination order, so results are bitwise reproducible run to run.
#include "internal.hpp"

#include <mutex>

namespace sqdk {

namespace {

constexpr int kBlock = 256;

// NP pairs exactly (1..4: the solvers' calls), or NP = kMaxPairs for any
// count up to it. The grid is small and fixed (kDotBlocks, for reproducible
// partial sums), so each thread keeps U elements' loads in flight per step,
// all issued before the first is added; they are added in index order, as
// one at a time.
template <typename VT, int NP>
__global__ void __launch_bounds__(kBlock)
dots_partial_kernel(int64_t n, int npairs, PairList pl, double* __restrict__ partials,
                    const int32_t* __restrict__ skip) {
  if (skip && skip[0] != 0) return;
  constexpr int U = NP <= 2 ? 4 : 2;
  __shared__ double red[2 * NP * (kBlock / kWave)];
  double acc[2 * NP];
#pragma unroll
  for (int k = 0; k < 2 * NP; ++k) acc[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  if constexpr (NP == kMaxPairs) {
    // any count: one element per step, each pair's loads just before its use
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += stride) {
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        if (k < npairs) {
          const double2 x = Vec<VT>::load(reinterpret_cast<const VT*>(pl.a[k]), i);
          const double2 y = Vec<VT>::load(reinterpret_cast<const VT*>(pl.b[k]), i);
          const double2 p = cmulc(x, y);
          acc[2 * k] += p.x;
          acc[2 * k + 1] += p.y;
        }
      }
    }
  } else {
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += U * stride) {
      double2 x[NP][U], y[NP][U];
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const VT* a = reinterpret_cast<const VT*>(pl.a[k]);
        const VT* b = reinterpret_cast<const VT*>(pl.b[k]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t j = i + u * stride < n ? i + u * stride : i;
          x[k][u] = Vec<VT>::load(a, j);
          y[k][u] = Vec<VT>::load(b, j);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (i + u * stride >= n) break;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          const double2 p = cmulc(x[k][u], y[k][u]);
          acc[2 * k] += p.x;
          acc[2 * k + 1] += p.y;
        }
      }
    }
  }
  block_sum<2 * NP>(acc, red);
  if (threadIdx.x == 0) {
    // a constant bound: indexed by the runtime pair count, acc would live in
    // scratch memory, and every accumulation above with it
#pragma unroll
    for (int k = 0; k < 2 * NP; ++k)
      if (k < 2 * npairs) partials[(int64_t)k * gridDim.x + blockIdx.x] = acc[k];
  }
}

// One block per output value: sum of the per-block partials in a fixed order.
__global__ void __launch_bounds__(kBlock)
dots_final_kernel(int nblocks, const double* __restrict__ partials, double* __restrict__ out,
                  const int32_t* __restrict__ skip) {
  if (skip && skip[0] != 0) return;
  __shared__ double red[kBlock / kWave];
  double v[1] = {0.0};
  // eight loads in flight per thread, added in the same order as one at a time
  const double* __restrict__ p = partials + (int64_t)blockIdx.x * nblocks;
  constexpr int U = 8;
  for (int i0 = threadIdx.x; i0 < nblocks; i0 += U * kBlock) {
    double t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * kBlock;
      t[u] = i < nblocks ? p[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * kBlock < nblocks) v[0] += t[u];
  }
  block_sum<1>(v, red);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

constexpr int kGramSums = 14;

// Sums over V = (X, R, P), AV = (AX, AR, AP):
//  0 XX 1 RR 2 PP 3 XR 4 XP 5 RP          (Gram, conj(first) * second)
//  6 X.AX 7 R.AR 8 P.AP 9 X.AR 10 X.AP 11 R.AP   (projected operator)
// 12 AX.AX
template <typename VT>
__global__ void __launch_bounds__(kBlock)
gram_partial_kernel(int64_t n, const VT* __restrict__ X, const VT* __restrict__ R,
                    const VT* __restrict__ P, const VT* __restrict__ AX,
                    const VT* __restrict__ AR, const VT* __restrict__ AP,
                    const VT* __restrict__ RES, double* __restrict__ partials,
                    const int32_t* __restrict__ skip) {
  if (skip && skip[0] != 0) return;
  __shared__ double red[2 * kGramSums * (kBlock / kWave)];
  double acc[2 * kGramSums];
#pragma unroll
  for (int k = 0; k < 2 * kGramSums; ++k) acc[k] = 0.0;
  // (the grid is small and fixed: two elements' loads per thread in flight,
  // added in index order)
  constexpr int U = 2;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += U * stride) {
    double2 x[U], r[U], p[U], ax[U], ar[U], ap[U], res[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t j = i + u * stride < n ? i + u * stride : i;
      x[u] = Vec<VT>::load(X, j);
      r[u] = Vec<VT>::load(R, j);
      p[u] = Vec<VT>::load(P, j);
      ax[u] = Vec<VT>::load(AX, j);
      ar[u] = Vec<VT>::load(AR, j);
      ap[u] = Vec<VT>::load(AP, j);
      res[u] = RES ? Vec<VT>::load(RES, j) : r[u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i + u * stride >= n) break;
      const double2 prods[kGramSums] = {
          cmulc(x[u], x[u]),   cmulc(r[u], r[u]),   cmulc(p[u], p[u]),     cmulc(x[u], r[u]),
          cmulc(x[u], p[u]),   cmulc(r[u], p[u]),   cmulc(x[u], ax[u]),    cmulc(r[u], ar[u]),
          cmulc(p[u], ap[u]),  cmulc(x[u], ar[u]),  cmulc(x[u], ap[u]),    cmulc(r[u], ap[u]),
          cmulc(ax[u], ax[u]), cmulc(res[u], res[u])};
#pragma unroll
      for (int k = 0; k < kGramSums; ++k) {
        acc[2 * k] += prods[k].x;
        acc[2 * k + 1] += prods[k].y;
      }
    }
  }
  block_sum<2 * kGramSums>(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 2 * kGramSums; ++k) partials[(int64_t)k * gridDim.x + blockIdx.x] = acc[k];
  }
}

struct LinIn {
  const void* in[4];
  int cidx[4];
};

template <typename VT>
__global__ void __launch_bounds__(kBlock)
lincomb_kernel(int64_t n, int k, LinIn li, const double2* __restrict__ slots, VT* __restrict__ out,
               const int32_t* __restrict__ skip) {
  if (skip && skip[0] != 0) return;
  double2 c[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) c[j] = j < k ? slots[li.cidx[j]] : make_double2(0.0, 0.0);
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j < k) acc = cadd(acc, cmul(c[j], Vec<VT>::load(reinterpret_cast<const VT*>(li.in[j]), i)));
    }
    Vec<VT>::store(out, i, acc);
  }
}

// Diagonal (Jacobi) preconditioner for the lowest eigenpair at Ritz value
// theta: 1 / max(d - sigma, floor) with the shift sigma = min(theta, min d)
// (so d - sigma >= 0: the scaling stays positive while theta is still above
// the smallest diagonal entries, and it favours those) and floor =
// max(st.y, 1e-3 |theta|), st.y = 0.02 (mean d - min d) (a small fraction of
// the diagonal's spread; st.x = min d).
__device__ inline double precond_scale(double d, double theta, double2 st) {
  const double sigma = fmin(theta, st.x);
  return 1.0 / fmax(d - sigma, fmax(fmax(st.y, 1e-3 * fabs(theta)), 1e-300));
}

// W (the search direction in the basis) may be R itself (no preconditioner,
// diag == NULL) or a separate vector that receives diag-preconditioned R'
// (R is then not stored).
// With gpart, the pass also sums, per block, the next iteration's Gram sums
// over the values it stores: every sum of gram_partial_kernel except R.AR
// (the next product's, summed by gram_rar_kernel), with X.AR taken as
// AX.R (A Hermitian).
// With the hit table's diagonal cdiag (the preconditioner then reads its real
// part instead of diag): OFF_IN, AW is the product without the diagonal, and
// A W = AW + cdiag W is formed here from the W it reads anyway; OFF_NEXT, the
// next product is without it too, and the diagonal's share of R.AR, sum
// cdiag |W'|^2, becomes R.AR's partial (gram_rar_kernel adds the rest). The
// product then reads neither the diagonal nor v's own entries, 2 s_v bytes
// per row less, and this pass reads as many bytes as with diag.
template <typename VT, bool OFF_IN, bool OFF_NEXT>
__global__ void __launch_bounds__(kBlock)
fused_update_kernel(int64_t n, VT* __restrict__ X, VT* __restrict__ AX, VT* R, VT* W,
                    const VT* __restrict__ AW, VT* __restrict__ P, VT* __restrict__ AP,
                    const double* __restrict__ diag, const VT* __restrict__ cdiag,
                    const double2* __restrict__ c, const double2* __restrict__ dstat,
                    double* __restrict__ gpart, const int32_t* __restrict__ skip) {
  if (skip && skip[0] != 0) return;
  __shared__ double red[2 * kGramSums * (kBlock / kWave)];
  // c[0..2] = c0, c1, c2; c[3..4] = d1, d2; c[5].x = theta
  const double2 c0 = c[0], c1 = c[1], c2 = c[2], d1 = c[3], d2 = c[4];
  const double theta = c[5].x;
  const double2 st = diag ? *dstat : make_double2(0.0, 0.0);
  double acc[2 * kGramSums];
#pragma unroll
  for (int k = 0; k < 2 * kGramSums; ++k) acc[k] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double2 x = Vec<VT>::load(X, i), w = Vec<VT>::load(W, i), p = Vec<VT>::load(P, i);
    const double2 ax = Vec<VT>::load(AX, i), ap = Vec<VT>::load(AP, i);
    double2 aw = Vec<VT>::load(AW, i);
    double dreal = 0.0;
    double2 dc = make_double2(0.0, 0.0);
    if (cdiag) {
      dc = Vec<VT>::load(cdiag, i);
      dreal = dc.x;
    } else if (diag) {
      dreal = diag[i];
    }
    if constexpr (OFF_IN) aw = cadd(aw, cmul(dc, w));
    const double2 xn = cadd(cadd(cmul(c0, x), cmul(c1, w)), cmul(c2, p));
    const double2 axn = cadd(cadd(cmul(c0, ax), cmul(c1, aw)), cmul(c2, ap));
    const double2 pn = cadd(cmul(d1, w), cmul(d2, p));
    const double2 apn = cadd(cmul(d1, aw), cmul(d2, ap));
    const double2 rn = csub(axn, cscale(xn, theta));
    const double2 wn = diag ? cscale(rn, precond_scale(dreal, theta, st)) : rn;
    Vec<VT>::store(X, i, xn);
    Vec<VT>::store(AX, i, axn);
    Vec<VT>::store(P, i, pn);
    Vec<VT>::store(AP, i, apn);
    // With the preconditioner the basis holds W; R itself is read again only
    // after being rebuilt (refresh, verification), so only |R|^2 is kept.
    if (diag)
      Vec<VT>::store(W, i, wn);
    else
      Vec<VT>::store(R, i, rn);
    if (gpart) {
      const double2 xs = Vec<VT>::round(xn), ws = Vec<VT>::round(wn), ps = Vec<VT>::round(pn);
      const double2 axs = Vec<VT>::round(axn), aps = Vec<VT>::round(apn);
      const double2 rs = Vec<VT>::round(rn);
      // squared norms (sums 0, 1, 2, 12, 13) are real: only their real parts
      auto nrm = [](double2 a) { return make_double2(a.x * a.x + a.y * a.y, 0.0); };
      const double ww = ws.x * ws.x + ws.y * ws.y;
      const double2 prods[kGramSums] = {
          nrm(xs), nrm(ws), nrm(ps), cmulc(xs, ws), cmulc(xs, ps),
          cmulc(ws, ps), cmulc(xs, axs), make_double2(dc.x * ww, dc.y * ww), cmulc(ps, aps),
          cmulc(axs, ws), cmulc(xs, aps), cmulc(ws, aps), nrm(axs), nrm(rs)};
      constexpr unsigned kReal = (1u << 0) | (1u << 1) | (1u << 2) | (1u << 12) | (1u << 13);
#pragma unroll
      for (int k = 0; k < kGramSums; ++k) {
        if (k == 7 && !OFF_NEXT) continue;
        acc[2 * k] += prods[k].x;
        if (!((kReal >> k) & 1u)) acc[2 * k + 1] += prods[k].y;
      }
    }
  }
  if (!gpart) return;
  block_sum<2 * kGramSums>(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 2 * kGramSums; ++k) {
      if (!OFF_NEXT && (k == 14 || k == 15)) continue;   // R.AR: gram_rar_kernel's
      gpart[(int64_t)k * gridDim.x + blockIdx.x] = acc[k];
    }
  }
}

// The one Gram sum that needs the new product: R.AR (W.AW) per block, into
// the slots the fused update left free (same grid as that update).
template <typename VT>
__global__ void __launch_bounds__(kBlock)
gram_rar_kernel(int64_t n, const VT* __restrict__ W, const VT* __restrict__ AW,
                double* __restrict__ gpart, const int32_t* __restrict__ skip, int accumulate) {
  if (skip && skip[0] != 0) return;
  __shared__ double red[2 * (kBlock / kWave)];
  double acc[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double2 pr = cmulc(Vec<VT>::load(W, i), Vec<VT>::load(AW, i));
    acc[0] += pr.x;
    acc[1] += pr.y;
  }
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) {
    // (accumulate: the update pass left the diagonal's share there)
    double* g0 = gpart + (int64_t)14 * gridDim.x + blockIdx.x;
    double* g1 = gpart + (int64_t)15 * gridDim.x + blockIdx.x;
    *g0 = accumulate ? *g0 + acc[0] : acc[0];
    *g1 = accumulate ? *g1 + acc[1] : acc[1];
  }
}

// W = R * precond_scale(diag, theta), theta = slots[theta_slot].x, the
// diagonal's statistics at slots[stat_slot].
template <typename VT>
__global__ void __launch_bounds__(kBlock)
precond_kernel(int64_t n, const VT* __restrict__ R, const double* __restrict__ diag,
               const double2* __restrict__ slots, int theta_slot, int stat_slot,
               VT* __restrict__ W, const int32_t* __restrict__ skip) {
  if (skip && skip[0] != 0) return;
  const double theta = slots[theta_slot].x;
  const double2 st = slots[stat_slot];
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    Vec<VT>::store(W, i, cscale(Vec<VT>::load(R, i), precond_scale(diag[i], theta, st)));
}

// Statistics of the diagonal for the preconditioner: per-block min and sum,
// then one block writes out = (min, 0.02 (mean - min)).
__global__ void __launch_bounds__(kBlock)
diag_stats_partial_kernel(int64_t n, const double* __restrict__ diag,
                          double* __restrict__ partials) {
  __shared__ double smin[kBlock], ssum[kBlock];
  double mn = INFINITY, sm = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    mn = fmin(mn, diag[i]);
    sm += diag[i];
  }
  smin[threadIdx.x] = mn;
  ssum[threadIdx.x] = sm;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + h]);
      ssum[threadIdx.x] += ssum[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = smin[0];
    partials[2 * blockIdx.x + 1] = ssum[0];
  }
}

__global__ void __launch_bounds__(kBlock)
diag_stats_final_kernel(int64_t n, int blocks, const double* __restrict__ partials,
                        double2* __restrict__ out) {
  __shared__ double smin[kBlock], ssum[kBlock];
  double mn = INFINITY, sm = 0.0;
  for (int b = threadIdx.x; b < blocks; b += kBlock) {
    mn = fmin(mn, partials[2 * b]);
    sm += partials[2 * b + 1];
  }
  smin[threadIdx.x] = mn;
  ssum[threadIdx.x] = sm;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + h]);
      ssum[threadIdx.x] += ssum[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double dmin = smin[0];
    *out = make_double2(dmin, 0.02 * (ssum[0] / (double)n - dmin));
  }
}

template <typename VI, typename VO>
__global__ void __launch_bounds__(kBlock)
convert_kernel(int64_t n, const VI* __restrict__ in, VO* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    Vec<VO>::store(out, i, Vec<VI>::load(in, i));
}

inline unsigned stream_grid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBlock), 8192));
}

inline unsigned dot_grid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBlock), kDotBlocks));
}

}  // namespace

int dots_impl(int64_t n, int precision, int npairs, const PairList& pl, double* out,
              double* partials, hipStream_t s, const int32_t* skip) {
  SQDK_REQUIRE(npairs >= 1 && npairs <= kMaxPairs, "npairs %d out of range", npairs);
  SQDK_REQUIRE(n >= 1, "empty vectors");
  const unsigned g = dot_grid(n);
  if (precision != SQDK_C128 && precision != SQDK_C64)
    return fail(SQDK_EINVAL, "unknown precision %d", precision);
  auto launch = [&](auto vt) {
    using VT = decltype(vt);
    switch (npairs) {
      case 1: dots_partial_kernel<VT, 1><<<g, kBlock, 0, s>>>(n, npairs, pl, partials, skip); break;
      case 2: dots_partial_kernel<VT, 2><<<g, kBlock, 0, s>>>(n, npairs, pl, partials, skip); break;
      case 3: dots_partial_kernel<VT, 3><<<g, kBlock, 0, s>>>(n, npairs, pl, partials, skip); break;
      case 4: dots_partial_kernel<VT, 4><<<g, kBlock, 0, s>>>(n, npairs, pl, partials, skip); break;
      default:
        dots_partial_kernel<VT, kMaxPairs><<<g, kBlock, 0, s>>>(n, npairs, pl, partials, skip);
    }
  };
  if (precision == SQDK_C128) launch(double2{});
  else launch(float2{});
  SQDK_LAUNCHED("dots_partial_kernel");
  dots_final_kernel<<<2 * npairs, kBlock, 0, s>>>((int)g, partials, out, skip);
  SQDK_LAUNCHED("dots_final_kernel");
  return SQDK_OK;
}

int gram_impl(int64_t n, int precision, const void* const* v, const void* const* av,
              const void* res, double* out, double* partials, hipStream_t s,
              const int32_t* skip) {
  const unsigned g = dot_grid(n);
  if (precision == SQDK_C128)
    gram_partial_kernel<double2><<<g, kBlock, 0, s>>>(
        n, (const double2*)v[0], (const double2*)v[1], (const double2*)v[2],
        (const double2*)av[0], (const double2*)av[1], (const double2*)av[2],
        (const double2*)res, partials, skip);
  else
    gram_partial_kernel<float2><<<g, kBlock, 0, s>>>(
        n, (const float2*)v[0], (const float2*)v[1], (const float2*)v[2],
        (const float2*)av[0], (const float2*)av[1], (const float2*)av[2],
        (const float2*)res, partials, skip);
  SQDK_LAUNCHED("gram_partial_kernel");
  dots_final_kernel<<<2 * kGramSums, kBlock, 0, s>>>((int)g, partials, out, skip);
  SQDK_LAUNCHED("dots_final_kernel");
  return SQDK_OK;
}

int lincomb_impl(int64_t n, int precision, void* out, int k, const void* const* in,
                 const int* cidx, const double* slots, hipStream_t s, const int32_t* skip) {
  SQDK_REQUIRE(k >= 1 && k <= 4, "lincomb of %d inputs", k);
  LinIn li{};
  for (int j = 0; j < k; ++j) {
    li.in[j] = in[j];
    li.cidx[j] = cidx[j];
  }
  const unsigned g = stream_grid(n);
  const double2* sl = reinterpret_cast<const double2*>(slots);
  if (precision == SQDK_C128)
    lincomb_kernel<double2><<<g, kBlock, 0, s>>>(n, k, li, sl, (double2*)out, skip);
  else
    lincomb_kernel<float2><<<g, kBlock, 0, s>>>(n, k, li, sl, (float2*)out, skip);
  SQDK_LAUNCHED("lincomb_kernel");
  return SQDK_OK;
}

template <typename VT>
void launch_fused_update(unsigned g, hipStream_t s, int64_t n, void* X, void* AX, void* R,
                         void* W, const void* AW, void* P, void* AP, const double* diag,
                         const void* cdiag, int mode, const double2* c, const double2* st,
                         double* gpart, const int32_t* skip) {
  auto run = [&](auto kern) {
    kern<<<g, kBlock, 0, s>>>(n, (VT*)X, (VT*)AX, (VT*)R, (VT*)W, (const VT*)AW, (VT*)P, (VT*)AP,
                              diag, (const VT*)cdiag, c, st, gpart, skip);
  };
  switch (mode) {
    case 1: run(fused_update_kernel<VT, true, false>); break;
    case 2: run(fused_update_kernel<VT, false, true>); break;
    case 3: run(fused_update_kernel<VT, true, true>); break;
    default: run(fused_update_kernel<VT, false, false>); break;
  }
}

int fused_update_impl(int64_t n, int precision, void* X, void* AX, void* R, void* W,
                      const void* AW, void* P, void* AP, const double* diag,
                      const double* slots, int c_base, int stat_slot, double* gpart,
                      hipStream_t s, const int32_t* skip, const void* cdiag, int cdiag_mode) {
  SQDK_REQUIRE(cdiag_mode >= 0 && cdiag_mode <= 3, "cdiag_mode %d", cdiag_mode);
  SQDK_REQUIRE(!cdiag_mode || (cdiag && diag), "the diagonal split needs the preconditioner's "
               "diagonal and the hit table's");
  SQDK_REQUIRE(!(cdiag_mode & 2) || gpart, "the diagonal's R.AR share needs gpart");
  const unsigned g = stream_grid(n);
  static_assert(kGramPartialDoubles >= (int64_t)2 * kGramSums * 8192, "split Gram partials");
  const double2* c = reinterpret_cast<const double2*>(slots) + c_base;
  const double2* st = reinterpret_cast<const double2*>(slots) + stat_slot;
  if (precision == SQDK_C128)
    launch_fused_update<double2>(g, s, n, X, AX, R, W, AW, P, AP, diag, cdiag, cdiag_mode, c, st,
                                 gpart, skip);
  else
    launch_fused_update<float2>(g, s, n, X, AX, R, W, AW, P, AP, diag, cdiag, cdiag_mode, c, st,
                                gpart, skip);
  SQDK_LAUNCHED("fused_update_kernel");
  return SQDK_OK;
}

int gram_rar_impl(int64_t n, int precision, const void* W, const void* AW, double* gpart,
                  double* out, hipStream_t s, const int32_t* skip, bool accumulate) {
  const unsigned g = stream_grid(n);   // the fused update's grid: its partials share gpart
  if (precision == SQDK_C128)
    gram_rar_kernel<double2><<<g, kBlock, 0, s>>>(n, (const double2*)W, (const double2*)AW,
                                                  gpart, skip, accumulate ? 1 : 0);
  else
    gram_rar_kernel<float2><<<g, kBlock, 0, s>>>(n, (const float2*)W, (const float2*)AW, gpart,
                                                 skip, accumulate ? 1 : 0);
  SQDK_LAUNCHED("gram_rar_kernel");
  dots_final_kernel<<<2 * kGramSums, kBlock, 0, s>>>((int)g, gpart, out, skip);
  SQDK_LAUNCHED("dots_final_kernel");
  return SQDK_OK;
}

int precond_impl(int64_t n, int precision, const void* R, const double* diag,
                 const double* slots, int theta_slot, int stat_slot, void* W, hipStream_t s,
                 const int32_t* skip) {
  const unsigned g = stream_grid(n);
  const double2* sl = reinterpret_cast<const double2*>(slots);
  if (precision == SQDK_C128)
    precond_kernel<double2><<<g, kBlock, 0, s>>>(n, (const double2*)R, diag, sl, theta_slot,
                                                 stat_slot, (double2*)W, skip);
  else
    precond_kernel<float2><<<g, kBlock, 0, s>>>(n, (const float2*)R, diag, sl, theta_slot,
                                                stat_slot, (float2*)W, skip);
  SQDK_LAUNCHED("precond_kernel");
  return SQDK_OK;
}

int diag_stats_impl(int64_t n, const double* diag, double* partials, int max_blocks,
                    double* out, hipStream_t s) {
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBlock), max_blocks));
  diag_stats_partial_kernel<<<g, kBlock, 0, s>>>(n, diag, partials);
  SQDK_LAUNCHED("diag_stats_partial_kernel");
  diag_stats_final_kernel<<<1, kBlock, 0, s>>>(n, g, partials,
                                               reinterpret_cast<double2*>(out));
  SQDK_LAUNCHED("diag_stats_final_kernel");
  return SQDK_OK;
}

int convert_impl(int64_t n, int from_p, const void* in, int to_p, void* out, hipStream_t s) {
  SQDK_REQUIRE(in && out, "NULL vector");
  SQDK_REQUIRE(n >= 1, "empty vector");
  const unsigned g = stream_grid(n);
  if (from_p == SQDK_C128 && to_p == SQDK_C64)
    convert_kernel<double2, float2><<<g, kBlock, 0, s>>>(n, (const double2*)in, (float2*)out);
  else if (from_p == SQDK_C64 && to_p == SQDK_C128)
    convert_kernel<float2, double2><<<g, kBlock, 0, s>>>(n, (const float2*)in, (double2*)out);
  else if (from_p == SQDK_C128 && to_p == SQDK_C128)
    convert_kernel<double2, double2><<<g, kBlock, 0, s>>>(n, (const double2*)in, (double2*)out);
  else if (from_p == SQDK_C64 && to_p == SQDK_C64)
    convert_kernel<float2, float2><<<g, kBlock, 0, s>>>(n, (const float2*)in, (float2*)out);
  else
    return fail(SQDK_EINVAL, "unknown precision pair %d -> %d", from_p, to_p);
  SQDK_LAUNCHED("convert_kernel");
  return SQDK_OK;
}

}  // namespace sqdk

using namespace sqdk;

extern "C" {

int sqdk_dots_workspace(int64_t n, int32_t npairs, size_t* bytes) {
  SQDK_REQUIRE(bytes, "NULL output");
  SQDK_REQUIRE(npairs >= 1 && npairs <= kMaxPairs, "npairs %d out of range", npairs);
  (void)n;
  *bytes = sizeof(double) * 2 * (size_t)npairs * kDotBlocks;
  return SQDK_OK;
}

int sqdk_dots(int64_t n, int32_t precision, int32_t npairs, const void* const* a,
              const void* const* b, double* out, void* work, size_t work_bytes, void* stream) {
  SQDK_REQUIRE(a && b && out && work, "NULL pointer argument");
  SQDK_REQUIRE(npairs >= 1 && npairs <= kMaxPairs, "npairs %d out of range", npairs);
  SQDK_REQUIRE(work_bytes >= sizeof(double) * 2 * (size_t)npairs * kDotBlocks,
               "dots workspace of %zu bytes too small", work_bytes);
  PairList pl{};
  for (int k = 0; k < npairs; ++k) {
    SQDK_REQUIRE(a[k] && b[k], "NULL vector in pair %d", k);
    pl.a[k] = a[k];
    pl.b[k] = b[k];
  }
  return dots_impl(n, precision, npairs, pl, out, reinterpret_cast<double*>(work),
                   as_stream(stream));
}

// ---- sqdk_dots without a caller workspace ----------------------------------
//
// The per-block partial sums live in this file-scope device array instead of
// a caller workspace: 2 x 16 pairs x 512 blocks doubles (128 KiB) on every GPU
// the library runs on, held until the process ends. Host code must not pass
// the array's name to a kernel (on the host it is a stand-in, not a device
// address): each device's address comes from hipGetSymbolAddress, taken once
// per device with that device current and kept in dots_scratch_ptr[]. Every
// partial a call reads was written by the same call (block b of pair k at
// k * grid + b), so nothing needs clearing between calls; calls that overlap
// on one GPU (two streams at once) share the array and race.
}  // extern "C"

static __device__ __align__(256) double scratch[2 * sqdk::kMaxPairs * sqdk::kDotBlocks];

namespace sqdk {
namespace {

constexpr int kMaxDevices = 64;
std::mutex dots_scratch_mutex;
double* dots_scratch_ptr[kMaxDevices] = {};

int dots_scratch(double** out) {
  int dev = -1;
  SQDK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(SQDK_ERANGE, "device %d beyond the %d the static scratch tracks", dev,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(dots_scratch_mutex);
  if (!dots_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SQDK_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                  hipGetErrorString(e));
    dots_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = dots_scratch_ptr[dev];
  return SQDK_OK;
}

}  // namespace
}  // namespace sqdk

extern "C" {

int sqdk_dots_noalloc(int64_t n, int32_t precision, int32_t npairs, const void* const* a,
                      const void* const* b, double* out, void* stream) {
  SQDK_REQUIRE(a && b && out, "NULL pointer argument");
  SQDK_REQUIRE(npairs >= 1, "npairs %d out of range", npairs);
  SQDK_REQUIRE(npairs <= kMaxPairs,
               "npairs %d exceeds the %d pairs the static scratch of sqdk_dots_noalloc holds",
               npairs, kMaxPairs);
  PairList pl{};
  for (int k = 0; k < npairs; ++k) {
    SQDK_REQUIRE(a[k] && b[k], "NULL vector in pair %d", k);
    pl.a[k] = a[k];
    pl.b[k] = b[k];
  }
  double* partials = nullptr;
  SQDK_TRY(dots_scratch(&partials));
  return dots_impl(n, precision, npairs, pl, out, partials, as_stream(stream));
}

int sqdk_convert(int64_t n, int32_t from_precision, const void* in, int32_t to_precision,
                 void* out, void* stream) {
  return convert_impl(n, from_precision, in, to_precision, out, as_stream(stream));
}

}  // extern "C"
