// Batched fused LOBPCG: many independent SQD instances solved together.
//
// Instances are laid end to end (sqdk_batch). Every vector kernel runs one
// workgroup per row chunk of the matvec's chunk table (chunks never cross an
// instance), so per-instance inner products are segmented sums: one partial
// per chunk, then one fixed-order sum over the instance's contiguous chunks.
// Each instance keeps its own device state (Gram sums, Rayleigh-Ritz
// coefficients, theta, flags); one device thread per instance runs the FP64
// Rayleigh-Ritz step of the fused iteration (the same algebra as the single
// solver). A converged instance sets its flag, after which the matvec and all
// vector kernels skip its chunks. refresh_every iterations plus the refresh
// of A X and A P form one chunk of work, captured once as a HIP graph and
// replayed until every instance has converged (then verified with explicit
// products) or maxiter is reached.
#include "internal.hpp"
#include "small_eig.hpp"

#include <vector>

namespace sqdk {

namespace {

constexpr int kBlock = 256;
constexpr int kGram = 13;

struct alignas(16) InstState {
  double2 gram[kGram];
  double2 upd[6];
  double2 dots[2];
  double theta;
  double rn, axn, thr;
  double coef;        // real coefficient of the last scale / axpy step
  int32_t niter, pvalid, status, matvecs;
};

enum BStep : int { kInitNorm = 0, kRho, kRR, kRefresh, kVerify };

struct BParams {
  double tol;
  int criterion;
  int precision;
};

__device__ inline void item_rows(const uint32_t* __restrict__ items, int n_items, int64_t n, int k,
                                 int64_t& r0, int64_t& r1) {
  r0 = items[k];
  r1 = (k + 1 < n_items) ? (int64_t)items[k + 1] : n;
}

// Per-chunk partial sums. GRAM: the 13 fused sums over (X,R,P | AX,AR,AP).
// Otherwise two pairs: conj(a0) b0 and conj(a1) b1.
template <typename VT, bool GRAM>
__global__ void __launch_bounds__(kBlock)
bsum_partial_kernel(const uint32_t* __restrict__ items, const uint16_t* __restrict__ item_inst,
                    int n_items, int64_t n, const int32_t* __restrict__ conv,
                    const VT* __restrict__ v0, const VT* __restrict__ v1,
                    const VT* __restrict__ v2, const VT* __restrict__ v3,
                    const VT* __restrict__ v4, const VT* __restrict__ v5,
                    double* __restrict__ partials) {
  constexpr int NS = GRAM ? kGram : 2;
  const int k = blockIdx.x;
  if (conv && conv[item_inst[k]] != 0) return;
  __shared__ double red[2 * NS * (kBlock / kWave)];
  double acc[2 * NS];
#pragma unroll
  for (int j = 0; j < 2 * NS; ++j) acc[j] = 0.0;
  int64_t r0, r1;
  item_rows(items, n_items, n, k, r0, r1);
  for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
    if (GRAM) {
      const double2 x = Vec<VT>::load(v0, i), r = Vec<VT>::load(v1, i), p = Vec<VT>::load(v2, i);
      const double2 ax = Vec<VT>::load(v3, i), ar = Vec<VT>::load(v4, i), ap = Vec<VT>::load(v5, i);
      const double2 pr[kGram] = {cmulc(x, x),  cmulc(r, r),  cmulc(p, p),  cmulc(x, r),  cmulc(x, p),
                                 cmulc(r, p),  cmulc(x, ax), cmulc(r, ar), cmulc(p, ap), cmulc(x, ar),
                                 cmulc(x, ap), cmulc(r, ap), cmulc(ax, ax)};
#pragma unroll
      for (int j = 0; j < kGram; ++j) {
        acc[2 * j] += pr[j].x;
        acc[2 * j + 1] += pr[j].y;
      }
    } else {
      const double2 p0 = cmulc(Vec<VT>::load(v0, i), Vec<VT>::load(v1, i));
      const double2 p1 = cmulc(Vec<VT>::load(v2, i), Vec<VT>::load(v3, i));
      acc[0] += p0.x;
      acc[1] += p0.y;
      acc[2] += p1.x;
      acc[3] += p1.y;
    }
  }
  block_sum<2 * NS>(acc, red);
  if (threadIdx.x == 0) {
    for (int j = 0; j < 2 * NS; ++j) partials[(int64_t)k * (2 * NS) + j] = acc[j];
  }
}

// Sum of the partials of instance i's chunks, in chunk order; one block per
// (instance, value). Writes the instance's gram[] (GRAM) or dots[].
template <bool GRAM>
__global__ void __launch_bounds__(kBlock)
bsum_final_kernel(const int32_t* __restrict__ item_base, const int32_t* __restrict__ conv,
                  const double* __restrict__ partials, InstState* __restrict__ st) {
  constexpr int NS = GRAM ? kGram : 2;
  const int inst = blockIdx.x / (2 * NS);
  const int j = blockIdx.x % (2 * NS);
  if (conv && conv[inst] != 0) return;
  __shared__ double red[kBlock / kWave];
  double v[1] = {0.0};
  for (int k = item_base[inst] + threadIdx.x; k < item_base[inst + 1]; k += kBlock)
    v[0] += partials[(int64_t)k * (2 * NS) + j];
  block_sum<1>(v, red);
  if (threadIdx.x == 0) {
    double* dst = reinterpret_cast<double*>(GRAM ? st[inst].gram : st[inst].dots);
    dst[j] = v[0];
  }
}

__device__ inline bool bcriterion(const BParams& p, double rn, double axn, double theta,
                                  int64_t nlen, double& thr) {
  if (p.criterion == SQDK_CRIT_REFERENCE) {
    thr = (axn - theta) * (double)nlen * 10.0 * p.tol;
    return rn < thr;
  }
  thr = p.tol * fabs(theta);
  return rn <= thr;
}

// One thread per instance.
__global__ void bscalar_kernel(int step, int n_inst, const int64_t* __restrict__ row_base,
                               InstState* __restrict__ st, int32_t* __restrict__ conv,
                               int32_t* __restrict__ counters, BParams prm, int use_skip) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  if (use_skip && conv[i] != 0) return;
  InstState& s = st[i];
  switch (step) {
    case kInitNorm:
      s.coef = 1.0 / sqrt(s.dots[0].x);
      s.matvecs += 1;  // the A X that follows
      break;
    case kRho:
      s.theta = s.dots[0].x;
      s.coef = -s.theta;
      break;
    case kRR: {
      // NaN or Inf in the sums ends this instance with status 3.
      bool finite = isfinite(s.theta);
      for (int k = 0; k < 13; ++k) finite = finite && isfinite(s.gram[k].x) && isfinite(s.gram[k].y);
      if (!finite) {
        s.status = 3;
        s.matvecs += 1;
        conv[i] = 1;
        atomicAdd(&counters[0], 1);
        break;
      }
      const double rn = sqrt(fmax(s.gram[1].x, 0.0));
      const double axn = sqrt(fmax(s.gram[12].x, 0.0));
      double thr;
      // a zero residual (an exact eigenvector, e.g. a one-state instance) has
      // converged even where the threshold is 0
      const bool ok = bcriterion(prm, rn, axn, s.theta, row_base[i + 1] - row_base[i], thr) ||
                      rn == 0.0;
      s.rn = rn;
      s.axn = axn;
      s.thr = thr;
      s.matvecs += 1;  // the A R of this iteration ran
      if (ok) {
        // a zero residual before the first update: the start spans an
        // invariant subspace of a multi-state instance (status 4)
        if (rn == 0.0 && s.niter == 0 && row_base[i + 1] - row_base[i] > 1) s.status = 4;
        conv[i] = 1;
        atomicAdd(&counters[0], 1);
        break;
      }
      double2 U[6];
      bool pv = false;
      if (fused_rr(s.gram, s.pvalid != 0, prm.precision, U, pv) != 0) {
        s.status = 2;
        conv[i] = 1;
        atomicAdd(&counters[0], 1);
        break;
      }
      for (int j = 0; j < 6; ++j) s.upd[j] = U[j];
      s.theta = U[5].x;
      s.pvalid = pv ? 1 : 0;
      s.niter += 1;
      break;
    }
    case kRefresh:
      s.theta = s.dots[1].x / s.dots[0].x;
      s.coef = -s.theta;
      s.matvecs += 2;  // A X and A P
      break;
    case kVerify: {
      const double axn = sqrt(s.dots[0].x), rn = sqrt(s.dots[1].x);
      double thr;
      const bool ok = bcriterion(prm, rn, axn, s.theta, row_base[i + 1] - row_base[i], thr) ||
                      rn == 0.0;
      s.rn = rn;
      s.axn = axn;
      s.thr = thr;
      conv[i] = (ok || s.status >= 2) ? 1 : 0;
      break;
    }
    default:
      break;
  }
}

__global__ void bcount_kernel(int n_inst, const int32_t* __restrict__ conv,
                              int32_t* __restrict__ counters) {
  __shared__ int total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  int c = 0;
  for (int i = threadIdx.x; i < n_inst; i += blockDim.x) c += conv[i] != 0;
  atomicAdd(&total, c);
  __syncthreads();
  if (threadIdx.x == 0) counters[0] = total;
}

// x *= coef (real, per instance)
template <typename VT>
__global__ void __launch_bounds__(kBlock)
bscale_kernel(const uint32_t* __restrict__ items, const uint16_t* __restrict__ item_inst,
              int n_items, int64_t n, const int32_t* __restrict__ conv,
              const InstState* __restrict__ st, VT* __restrict__ x) {
  const int k = blockIdx.x;
  const int inst = item_inst[k];
  if (conv && conv[inst] != 0) return;
  const double c = st[inst].coef;
  int64_t r0, r1;
  item_rows(items, n_items, n, k, r0, r1);
  for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) Vec<VT>::store(x, i, cscale(Vec<VT>::load(x, i), c));
}

// out = a + coef * b (real coefficient per instance)
template <typename VT>
__global__ void __launch_bounds__(kBlock)
baxpy_kernel(const uint32_t* __restrict__ items, const uint16_t* __restrict__ item_inst,
             int n_items, int64_t n, const int32_t* __restrict__ conv,
             const InstState* __restrict__ st, const VT* __restrict__ a, const VT* __restrict__ b,
             VT* __restrict__ out) {
  const int k = blockIdx.x;
  const int inst = item_inst[k];
  if (conv && conv[inst] != 0) return;
  const double c = st[inst].coef;
  int64_t r0, r1;
  item_rows(items, n_items, n, k, r0, r1);
  for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock)
    Vec<VT>::store(out, i, cadd(Vec<VT>::load(a, i), cscale(Vec<VT>::load(b, i), c)));
}

template <typename VT>
__global__ void __launch_bounds__(kBlock)
bupdate_kernel(const uint32_t* __restrict__ items, const uint16_t* __restrict__ item_inst,
               int n_items, int64_t n, const int32_t* __restrict__ conv,
               const InstState* __restrict__ st, VT* __restrict__ X, VT* __restrict__ AX,
               VT* __restrict__ R, const VT* __restrict__ AR, VT* __restrict__ P,
               VT* __restrict__ AP) {
  const int k = blockIdx.x;
  const int inst = item_inst[k];
  if (conv[inst] != 0) return;
  const InstState& s = st[inst];
  const double2 c0 = s.upd[0], c1 = s.upd[1], c2 = s.upd[2], d1 = s.upd[3], d2 = s.upd[4];
  const double theta = s.upd[5].x;
  int64_t r0, r1;
  item_rows(items, n_items, n, k, r0, r1);
  for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
    const double2 x = Vec<VT>::load(X, i), r = Vec<VT>::load(R, i), p = Vec<VT>::load(P, i);
    const double2 ax = Vec<VT>::load(AX, i), ar = Vec<VT>::load(AR, i), ap = Vec<VT>::load(AP, i);
    const double2 xn = cadd(cadd(cmul(c0, x), cmul(c1, r)), cmul(c2, p));
    const double2 axn = cadd(cadd(cmul(c0, ax), cmul(c1, ar)), cmul(c2, ap));
    Vec<VT>::store(X, i, xn);
    Vec<VT>::store(AX, i, axn);
    Vec<VT>::store(P, i, cadd(cmul(d1, r), cmul(d2, p)));
    Vec<VT>::store(AP, i, cadd(cmul(d1, ar), cmul(d2, ap)));
    Vec<VT>::store(R, i, csub(axn, cscale(xn, theta)));
  }
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct BSolver {
  const sqdk_batch* b;
  const sqdk_tune* tune;
  int prec;
  hipStream_t s;
  InstState* st;
  int32_t* conv;
  int32_t* counters;
  double* partials;
  char* X;
  char* R;
  char* P;
  char* AX;
  char* AR;
  char* AP;
  void* mv_work;
  size_t mv_bytes;
  BParams prm;

  unsigned items() const { return (unsigned)b->n_items; }

  int matvec(const void* in, void* out, bool skip) {
    if (b->cache)
      return hit_matvec_impl(b->cache, prec, in, out, s, nullptr, nullptr, b->item_inst,
                             skip ? conv : nullptr, b->chunk_rows / kWave);
    return batch_matvec_impl(b, prec, in, out, tune, mv_work, mv_bytes, s, nullptr,
                             skip ? conv : nullptr, nullptr);
  }

  template <typename VT>
  int sums(bool gram, const void* a0, const void* b0, const void* a1, const void* b1, bool skip) {
    const int32_t* cv = skip ? conv : nullptr;
    if (gram) {
      bsum_partial_kernel<VT, true><<<items(), kBlock, 0, s>>>(
          b->items, b->item_inst, b->n_items, b->n, cv, (const VT*)X, (const VT*)R,
          (const VT*)P, (const VT*)AX, (const VT*)AR, (const VT*)AP, partials);
      SQDK_LAUNCHED("bsum_partial_kernel");
      bsum_final_kernel<true><<<(unsigned)b->n_inst * 2 * kGram, kBlock, 0, s>>>(
          b->item_base, cv, partials, st);
    } else {
      bsum_partial_kernel<VT, false><<<items(), kBlock, 0, s>>>(
          b->items, b->item_inst, b->n_items, b->n, cv, (const VT*)a0, (const VT*)b0,
          (const VT*)a1, (const VT*)b1, nullptr, nullptr, partials);
      SQDK_LAUNCHED("bsum_partial_kernel");
      bsum_final_kernel<false><<<(unsigned)b->n_inst * 4, kBlock, 0, s>>>(b->item_base, cv,
                                                                          partials, st);
    }
    SQDK_LAUNCHED("bsum_final_kernel");
    return SQDK_OK;
  }

  int scalar(int step, bool skip) {
    bscalar_kernel<<<(unsigned)ceil_div(b->n_inst, 64), 64, 0, s>>>(
        step, b->n_inst, b->row_base, st, conv, counters, prm, skip ? 1 : 0);
    SQDK_LAUNCHED("bscalar_kernel");
    return SQDK_OK;
  }

  template <typename VT>
  int scale(void* x, bool skip) {
    bscale_kernel<VT><<<items(), kBlock, 0, s>>>(b->items, b->item_inst, b->n_items, b->n,
                                                skip ? conv : nullptr, st, (VT*)x);
    SQDK_LAUNCHED("bscale_kernel");
    return SQDK_OK;
  }

  template <typename VT>
  int axpy(const void* a, const void* bb, void* out, bool skip) {
    baxpy_kernel<VT><<<items(), kBlock, 0, s>>>(b->items, b->item_inst, b->n_items, b->n,
                                               skip ? conv : nullptr, st, (const VT*)a,
                                               (const VT*)bb, (VT*)out);
    SQDK_LAUNCHED("baxpy_kernel");
    return SQDK_OK;
  }

  template <typename VT>
  int update() {
    bupdate_kernel<VT><<<items(), kBlock, 0, s>>>(b->items, b->item_inst, b->n_items, b->n,
                                                 conv, st, (VT*)X, (VT*)AX, (VT*)R,
                                                 (const VT*)AR, (VT*)P, (VT*)AP);
    SQDK_LAUNCHED("bupdate_kernel");
    return SQDK_OK;
  }

  // refresh_every iterations, then A X and A P recomputed and R rebuilt.
  template <typename VT>
  int chunk(int iters) {
    for (int it = 0; it < iters; ++it) {
      SQDK_TRY(matvec(R, AR, true));
      SQDK_TRY(sums<VT>(true, nullptr, nullptr, nullptr, nullptr, true));
      SQDK_TRY(scalar(kRR, true));
      SQDK_TRY(update<VT>());
    }
    SQDK_TRY(matvec(X, AX, true));
    SQDK_TRY(matvec(P, AP, true));
    SQDK_TRY(sums<VT>(false, X, X, X, AX, true));
    SQDK_TRY(scalar(kRefresh, true));
    SQDK_TRY(axpy<VT>(AX, X, R, true));
    return SQDK_OK;
  }

  // Explicit A X, theta and residual for every instance, then the criterion.
  template <typename VT>
  int verify() {
    SQDK_TRY(matvec(X, AX, false));
    SQDK_TRY(matvec(P, AP, false));
    SQDK_TRY(sums<VT>(false, X, X, X, AX, false));
    SQDK_TRY(scalar(kRefresh, false));
    SQDK_TRY(axpy<VT>(AX, X, R, false));
    SQDK_TRY(sums<VT>(false, AX, AX, R, R, false));
    SQDK_TRY(scalar(kVerify, false));
    bcount_kernel<<<1, 256, 0, s>>>(b->n_inst, conv, counters);
    SQDK_LAUNCHED("bcount_kernel");
    return SQDK_OK;
  }
};

template <typename VT>
int run_batch(BSolver& sv, const sqdk_locg_opts& o, int32_t* total_iters) {
  const int every = o.refresh_every > 0 ? o.refresh_every : 16;
  // Set-up: normalise X, A X, theta, R = A X - theta X.
  SQDK_TRY(sv.sums<VT>(false, sv.X, sv.X, sv.X, sv.X, false));
  SQDK_TRY(sv.scalar(kInitNorm, false));
  SQDK_TRY(sv.scale<VT>(sv.X, false));
  SQDK_TRY(sv.matvec(sv.X, sv.AX, false));
  SQDK_TRY(sv.sums<VT>(false, sv.X, sv.AX, sv.X, sv.AX, false));
  SQDK_TRY(sv.scalar(kRho, false));
  SQDK_TRY(sv.axpy<VT>(sv.AX, sv.X, sv.R, false));

  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int rc = SQDK_OK;
  if (o.use_graph) {
    SQDK_HIP(hipStreamBeginCapture(sv.s, hipStreamCaptureModeThreadLocal));
    rc = sv.chunk<VT>(every);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(sv.s, &g);
    if (rc != SQDK_OK) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(SQDK_EHIP, "graph capture failed: %s", hipGetErrorString(e));
    graph = g;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      hipGraphDestroy(graph);
      return fail(SQDK_EHIP, "graph instantiation failed: %s", hipGetErrorString(ei));
    }
  }
  int done = 0;
  int32_t host_counters[4] = {0, 0, 0, 0};
  while (true) {
    if (done < o.maxiter) {
      if (exec) {
        const hipError_t el = hipGraphLaunch(exec, sv.s);
        if (el != hipSuccess) {
          rc = fail(SQDK_EHIP, "graph launch failed: %s", hipGetErrorString(el));
          break;
        }
      } else {
        rc = sv.chunk<VT>(every);
        if (rc != SQDK_OK) break;
      }
      done += every;
    }
    hipError_t e = hipMemcpyAsync(host_counters, sv.counters, sizeof(host_counters),
                                  hipMemcpyDeviceToHost, sv.s);
    if (e == hipSuccess) e = hipStreamSynchronize(sv.s);
    if (e != hipSuccess) {
      rc = fail(SQDK_EHIP, "reading convergence counters failed: %s", hipGetErrorString(e));
      break;
    }
    if (host_counters[0] >= sv.b->n_inst || done >= o.maxiter) {
      rc = sv.verify<VT>();
      if (rc != SQDK_OK) break;
      e = hipMemcpyAsync(host_counters, sv.counters, sizeof(host_counters),
                         hipMemcpyDeviceToHost, sv.s);
      if (e == hipSuccess) e = hipStreamSynchronize(sv.s);
      if (e != hipSuccess) {
        rc = fail(SQDK_EHIP, "reading convergence counters failed: %s", hipGetErrorString(e));
        break;
      }
      if (host_counters[0] >= sv.b->n_inst || done >= o.maxiter) break;
    }
  }
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  *total_iters = done;
  return rc;
}

}  // namespace

size_t batch_locg_bytes(const sqdk_batch* b, const sqdk_tune* tune, int precision) {
  const size_t es = precision == SQDK_C128 ? 16 : 8;
  const size_t vb = align256((size_t)b->n * es);
  size_t mv = 0;
  const int split = (tune && tune->group_split > 1) ? tune->group_split : 1;
  if (split > 1) mv = align256((size_t)(split - 1) * (size_t)b->n * es);
  return align256(sizeof(InstState) * (size_t)b->n_inst) + align256(4 * (size_t)b->n_inst) +
         256 + align256(sizeof(double) * 2 * kGram * (size_t)b->n_items) + 5 * vb + mv;
}

}  // namespace sqdk

using namespace sqdk;

extern "C" {

int sqdk_batch_locg_workspace(const sqdk_batch* batch, const sqdk_tune* tune,
                              const sqdk_locg_opts* opts, size_t* bytes) {
  SQDK_REQUIRE(batch && opts && bytes, "NULL pointer argument");
  SQDK_REQUIRE(opts->precision == SQDK_C128 || opts->precision == SQDK_C64,
               "unknown precision %d", opts->precision);
  *bytes = batch_locg_bytes(batch, tune, opts->precision);
  return SQDK_OK;
}

int sqdk_batch_locg_solve(const sqdk_batch* batch, const sqdk_tune* tune, void* x,
                          const sqdk_locg_opts* opts, void* work, size_t work_bytes,
                          sqdk_locg_result* results, void* stream) {
  SQDK_REQUIRE(batch && x && opts && work && results, "NULL pointer argument");
  SQDK_REQUIRE(opts->mode == SQDK_LOCG_FUSED, "batched solves use the fused mode");
  SQDK_REQUIRE(opts->precond == SQDK_PRECOND_NONE, "batched solves have no preconditioner");
  SQDK_REQUIRE(opts->criterion == SQDK_CRIT_REFERENCE || opts->criterion == SQDK_CRIT_RELATIVE,
               "unknown criterion %d", opts->criterion);
  SQDK_REQUIRE(opts->maxiter >= 0, "maxiter %d < 0", opts->maxiter);
  SQDK_REQUIRE(batch->row_base && batch->item_base, "batch lacks row_base or item_base");
  SQDK_REQUIRE(!batch->cache || (batch->cache->precision == opts->precision &&
                                 batch->cache->n == batch->n),
               "cached hit table does not match the solve (precision or size)");
  const size_t need = batch_locg_bytes(batch, tune, opts->precision);
  SQDK_REQUIRE(work_bytes >= need, "workspace of %zu bytes, %zu needed", work_bytes, need);

  BSolver sv{};
  sv.b = batch;
  sv.tune = tune;
  sv.prec = opts->precision;
  sv.prm.tol = opts->tol;
  sv.prm.criterion = opts->criterion;
  sv.prm.precision = opts->precision;
  const size_t es = opts->precision == SQDK_C128 ? 16 : 8;
  const size_t vb = align256((size_t)batch->n * es);
  char* w = static_cast<char*>(work);
  sv.st = reinterpret_cast<InstState*>(w);
  w += align256(sizeof(InstState) * (size_t)batch->n_inst);
  sv.conv = reinterpret_cast<int32_t*>(w);
  w += align256(4 * (size_t)batch->n_inst);
  sv.counters = reinterpret_cast<int32_t*>(w);
  w += 256;
  sv.partials = reinterpret_cast<double*>(w);
  w += align256(sizeof(double) * 2 * kGram * (size_t)batch->n_items);
  sv.X = static_cast<char*>(x);
  sv.R = w;
  sv.P = w + vb;
  sv.AX = w + 2 * vb;
  sv.AR = w + 3 * vb;
  sv.AP = w + 4 * vb;
  w += 5 * vb;
  sv.mv_bytes = 0;
  {
    const int split = (tune && tune->group_split > 1) ? tune->group_split : 1;
    if (split > 1) {
      sv.mv_work = w;
      sv.mv_bytes = (size_t)(split - 1) * (size_t)batch->n * es;
    }
  }

  hipStream_t caller = as_stream(stream);
  hipStream_t own = nullptr;
  if (opts->use_graph) {
    // Graph capture needs a stream other than the legacy default one.
    SQDK_HIP(hipStreamSynchronize(caller));
    SQDK_HIP(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
    sv.s = own;
  } else {
    sv.s = caller;
  }
  int rc = SQDK_OK;
  auto clean = [&]() {
    if (own) {
      hipStreamSynchronize(own);
      hipStreamDestroy(own);
    }
  };
  // Zero the state, flags, P and AP before anything is captured.
  hipError_t e = hipMemsetAsync(sv.st, 0, sizeof(InstState) * (size_t)batch->n_inst, sv.s);
  if (e == hipSuccess) e = hipMemsetAsync(sv.conv, 0, 4 * (size_t)batch->n_inst, sv.s);
  if (e == hipSuccess) e = hipMemsetAsync(sv.counters, 0, 256, sv.s);
  if (e == hipSuccess) e = hipMemsetAsync(sv.P, 0, (size_t)batch->n * es, sv.s);
  if (e == hipSuccess) e = hipMemsetAsync(sv.AP, 0, (size_t)batch->n * es, sv.s);
  if (e != hipSuccess) {
    clean();
    return fail(SQDK_EHIP, "clearing the solver state failed: %s", hipGetErrorString(e));
  }
  int32_t iters = 0;
  rc = opts->precision == SQDK_C128 ? run_batch<double2>(sv, *opts, &iters)
                                     : run_batch<float2>(sv, *opts, &iters);
  if (rc != SQDK_OK) {
    clean();
    return rc;
  }
  std::vector<InstState> host(batch->n_inst);
  std::vector<int32_t> conv(batch->n_inst);
  e = hipMemcpyAsync(host.data(), sv.st, sizeof(InstState) * host.size(), hipMemcpyDeviceToHost, sv.s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(conv.data(), sv.conv, 4 * conv.size(), hipMemcpyDeviceToHost, sv.s);
  if (e == hipSuccess) e = hipStreamSynchronize(sv.s);
  clean();
  if (e != hipSuccess) return fail(SQDK_EHIP, "reading results failed: %s", hipGetErrorString(e));
  for (int i = 0; i < batch->n_inst; ++i) {
    results[i].eigval = host[i].theta;
    results[i].resid_norm = host[i].rn;
    results[i].ax_norm = host[i].axn;
    results[i].niter = host[i].niter;
    results[i].converged = conv[i] && host[i].status != 3;
    results[i].matvecs = host[i].matvecs + 2;  // + the verification products
    results[i].status = host[i].status;
  }
  return SQDK_OK;
}

}  // extern "C"
