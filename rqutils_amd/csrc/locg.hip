// Single-vector LOBPCG on the device (rqutils/ground_locg.py:128-471).
//
// Every vector stays in HBM and every scalar (inner products, Rayleigh-Ritz
// matrices, eigenpairs, update coefficients, the convergence flag) stays in a
// small device slot area; the host only enqueues kernels and polls the flag
// every check_every iterations. Two modes:
//
//  * REFERENCE follows ground_locg.py step by step: body_iter0 (:264-272),
//    body_iter1 with the closed-form 2x2 (:274-286), then the loop body
//    (:288-341) with _project_out (:380-421), compute_sas with three matvecs
//    (:235-248), the Kopp closed-form 3x3 (:438-471) and a fourth matvec for
//    the new residual; convergence |r| < tol * 10 * N * (|Ax| - theta).
//  * FUSED keeps X, AX, R, AR, P, AP and needs one matvec (A R) per
//    iteration: one pass forms the 3x3 Gram and projected matrices of
//    (X, R, P) in FP64, one device thread solves the generalized 3x3
//    Rayleigh-Ritz problem (Cholesky + complex Jacobi), and one pass updates
//    all vectors. A X and A P are recomputed every refresh_every iterations
//    and once more to verify the residual at convergence.
#include "internal.hpp"
#include "small_eig.hpp"

#include <cmath>
#include <vector>

namespace sqdk {

namespace {

// ---- slot layout (complex double units) -----------------------------------
enum Slot : int {
  kOne = 0,
  kZero = 1,
  kD = 2,        // dot results D0.. (16 pairs)
  kGram = 20,    // 14 fused Gram sums
  kC = 40,       // lincomb coefficients C0..C7
  kUpd = 50,     // fused update coefficients (6)
  kTheta = 60,   // .x theta
  kKappa = 61,   // kappa[0..2]
  kInfo = 64,    // .x |r|, .y |Ax|
  kInfo2 = 65,   // .x reltol / tol*|theta|, .y last criterion value
  kDiagStat = 66, // preconditioner: .x min diag, .y 0.02 (mean diag - min diag)
  kNumSlots = 80
};
enum Flag : int { kConv = 0, kNiter = 1, kPValid = 2, kStatus = 3, kMatvecs = 4, kNumFlags = 16 };

constexpr size_t kSlotBytes = kNumSlots * sizeof(double2);
constexpr size_t kFlagBytes = kNumFlags * sizeof(int32_t);
constexpr size_t kPartialBytes = sizeof(double) * 2 * kMaxPairs * kDotBlocks;
constexpr size_t kGramPartialBytes = sizeof(double) * (size_t)kGramPartialDoubles;

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct ScalarParams {
  double tol;
  double nlen;     // vector length in the reference criterion
  int64_t n;       // vector length of the solve
  int criterion;
  int precision;
  double2* dbg;    // REFERENCE with diagnostics: kDbgRecord double2 per record, else NULL
};

// One diagnostics record of the reference iteration (the dict of
// ground_locg.py:211-233): (theta, rho), (reltol, converged), (|r|, 0),
// kappa[3], then S^H A S of (x, y, r) row-major.
constexpr int kDbgRecord = 15;

// Status flags (sqdk_locg_result.status).
constexpr int kStatusBreakdown = 2;    // Rayleigh-Ritz breakdown (fused)
constexpr int kStatusNonFinite = 3;    // NaN or Inf in the sums or the small eigenproblem
constexpr int kStatusInvariant = 4;    // exact eigenvector start with n > 1

__device__ inline bool finite_all(const double2* D, int k) {
  bool ok = true;
  for (int i = 0; i < k; ++i) ok = ok && isfinite(D[i].x) && isfinite(D[i].y);
  return ok;
}

// Stop the solve: the remaining kernels of the iteration see the flag.
__device__ inline void stop_nonfinite(int32_t* F) {
  F[kStatus] = kStatusNonFinite;
  F[kConv] = 1;
}

enum Step : int {
  kStepInitNorm = 0,   // C0 = 1 / sqrt(D0)
  kStepRho,            // theta = Re D0, C0 = -theta
  kStepInvNormGuard,   // C0 = 1 / (|D0| == 0 ? 1 : |D0|)
  kStepEig2,           // 2x2 Rayleigh-Ritz from D0..D2
  kStepInvNorm2,       // C0 = 1/sqrt(D0), C1 = 1/sqrt(D1)
  kStepNegTheta,       // C0 = -theta
  kStepProjNeg,        // C0 = -D0, C1 = -D1
  kStepThresh,         // C0 = sqrt(D0) >= 0.99 ? 1 : 0
  kStepEig3,           // 3x3 Rayleigh-Ritz from D0..D5; C0 = k1, C1 = k2
  kStepTU,             // from D0 = |s|^2: C0 = k0/ns, C1 = -ns, C2 = k0, C3 = 1
  kStepConv,           // D0 = |Ax|^2, D1 = |r|^2: reference criterion
  kStepFusedRR,        // fused: criterion + generalized 3x3 Rayleigh-Ritz
  kStepRefreshRho,     // fused refresh: theta = Re D1 / D0, C0 = -theta
  kStepVerify,         // D0 = |Ax|^2, D1 = |r|^2 after refresh: info only
  kStepZeroResid,      // D0 = |r|^2 of the start: exactly 0 -> converged at rho
  kStepDebug,          // D0..D5 = S^H A S of (x, y, r) as in kStepEig3: write record
};

__device__ bool criterion_met(const ScalarParams& p, double rnorm, double axnorm, double theta,
                              double2* S) {
  bool conv;
  double thr;
  if (p.criterion == SQDK_CRIT_REFERENCE) {
    thr = (axnorm - theta) * p.nlen * 10.0 * p.tol;
    conv = rnorm < thr;
  } else {
    thr = p.tol * fabs(theta);
    conv = rnorm <= thr;
  }
  S[kInfo] = make_double2(rnorm, axnorm);
  S[kInfo2].x = thr;
  return conv;
}

__global__ void scalar_kernel(int step, double2* __restrict__ S, int32_t* __restrict__ F,
                              ScalarParams prm, int use_skip, int arg) {
  if (threadIdx.x != 0) return;
  if (use_skip && F[kConv] != 0) return;
  double2* D = S + kD;
  double2* C = S + kC;
  switch (step) {
    case kStepInitNorm:
      C[0] = make_double2(1.0 / sqrt(D[0].x), 0.0);
      break;
    case kStepRho:
      S[kTheta] = make_double2(D[0].x, 0.0);
      C[0] = make_double2(-D[0].x, 0.0);
      break;
    case kStepZeroResid:
      // The start is an exact eigenvector (always so in a one-state
      // subspace): the 2x2 Rayleigh-Ritz of the reference (ground_locg.py:
      // 274-286, 424-435) would divide 0 by 0 here, so the solve stops at
      // theta = rho with the normalized start.
      // With more than one state this is an invariant subspace of the
      // start (a one-hot on a state without off-diagonal neighbours), and
      // theta need not be the lowest eigenvalue: status 4 says so.
      if (D[0].x == 0.0) {
        F[kConv] = 1;
        if (prm.n > 1) F[kStatus] = kStatusInvariant;
        S[kInfo] = make_double2(0.0, fabs(S[kTheta].x));
        S[kInfo2] = make_double2(0.0, 0.0);
      } else if (!isfinite(D[0].x) || !isfinite(S[kTheta].x)) {
        stop_nonfinite(F);
      }
      break;
    case kStepInvNormGuard: {
      const double nr = sqrt(D[0].x);
      C[0] = make_double2(1.0 / (nr == 0.0 ? 1.0 : nr), 0.0);
      break;
    }
    case kStepEig2: {
      // compute_sas(x, p): D0 = (p, Ax), D1 = (x, Ax), D2 = (p, Ap)
      double2 m[2][2];
      m[1][0] = D[0];
      m[0][1] = cconj(D[0]);
      m[0][0] = D[1];
      m[1][1] = D[2];
      if (!finite_all(D, 3)) {
        stop_nonfinite(F);
        break;
      }
      double lam;
      double2 k[2];
      eig2_reference(m, lam, k);
      S[kTheta] = make_double2(lam, 0.0);
      S[kKappa] = k[0];
      S[kKappa + 1] = k[1];
      // t = p k0 - x k1 ; u = x k0 + p k1
      C[0] = k[0];
      C[1] = make_double2(-k[1].x, -k[1].y);
      C[2] = k[0];
      C[3] = k[1];
      break;
    }
    case kStepInvNorm2:
      C[0] = make_double2(1.0 / sqrt(D[0].x), 0.0);
      // t = 0 (x already optimal in span(x, y, p)): y' = 0 instead of 0/0
      C[1] = make_double2(D[1].x > 0.0 ? 1.0 / sqrt(D[1].x) : 0.0, 0.0);
      break;
    case kStepNegTheta:
      C[0] = make_double2(-S[kTheta].x, 0.0);
      break;
    case kStepProjNeg:
      C[0] = make_double2(-D[0].x, -D[0].y);
      C[1] = make_double2(-D[1].x, -D[1].y);
      break;
    case kStepThresh:
      C[0] = make_double2(sqrt(D[0].x) >= 0.99 ? 1.0 : 0.0, 0.0);
      break;
    case kStepEig3: {
      // compute_sas(x, y, p): D0 = (y,Ax) D1 = (p,Ax) D2 = (p,Ay) D3 = (x,Ax) D4 = (y,Ay) D5 = (p,Ap)
      double2 m[3][3];
      m[1][0] = D[0];
      m[2][0] = D[1];
      m[2][1] = D[2];
      m[0][1] = cconj(D[0]);
      m[0][2] = cconj(D[1]);
      m[1][2] = cconj(D[2]);
      m[0][0] = D[3];
      m[1][1] = D[4];
      m[2][2] = D[5];
      if (!finite_all(D, 6)) {
        stop_nonfinite(F);
        break;
      }
      // A zero basis vector (p zeroed by the 0.99 rule, y = 0 after an update
      // that left x unchanged) has an all-zero row and column, which would add
      // a spurious eigenvalue 0 to the 3x3 problem: the reference then returns
      // NaN or, for an operator with a positive spectrum, theta = 0
      // (ground_locg.py:300-305, 421, 438-471). Rayleigh-Ritz runs on the
      // nonzero vectors instead.
      const bool yzero = czero(D[0]) && czero(D[2]) && czero(D[4]);
      const bool pzero = czero(D[1]) && czero(D[2]) && czero(D[5]);
      double lam;
      double2 k[3];
      if (!yzero && !pzero) {
        eig3_reference(m, lam, k);
      } else if (yzero && pzero) {
        lam = D[3].x;
        k[0] = make_double2(1.0, 0.0);
        k[1] = k[2] = make_double2(0.0, 0.0);
      } else {
        const int j = pzero ? 1 : 2;
        const double2 m2[2][2] = {{m[0][0], m[0][j]}, {m[j][0], m[j][j]}};
        double2 k2[2];
        eig2_reference(m2, lam, k2);
        k[0] = k2[0];
        k[j] = k2[1];
        k[3 - j] = make_double2(0.0, 0.0);
      }
      S[kTheta] = make_double2(lam, 0.0);
      for (int i = 0; i < 3; ++i) S[kKappa + i] = k[i];
      C[0] = k[1];
      C[1] = k[2];
      break;
    }
    case kStepTU: {
      const double ns = sqrt(D[0].x);
      const double2 k0 = S[kKappa];
      // s = 0: t = 0 (the reference divides by ns = 0 here)
      C[0] = ns > 0.0 ? cscale(k0, 1.0 / ns) : make_double2(0.0, 0.0);
      C[1] = make_double2(-ns, 0.0);
      C[2] = k0;
      C[3] = make_double2(1.0, 0.0);
      break;
    }
    case kStepConv: {
      const double axn = sqrt(D[0].x), rn = sqrt(D[1].x);
      const bool conv = criterion_met(prm, rn, axn, S[kTheta].x, S);
      F[kNiter] += 1;
      // with diagnostics every iteration runs, as the reference's scan over
      // maxiter steps does (ground_locg.py:358-364); the flag is recorded
      S[kInfo2].y = conv ? 1.0 : 0.0;
      if (conv && !prm.dbg) F[kConv] = 1;
      // The reference iterates on NaN until maxiter; a non-finite residual,
      // norm or theta ends the solve with status 3 instead.
      if (!isfinite(axn) || !isfinite(rn) || !isfinite(S[kTheta].x)) stop_nonfinite(F);
      break;
    }
    case kStepFusedRR: {
      const double2* g = S + kGram;
      const double theta = S[kTheta].x;
      // A non-finite sum (NaN or Inf in the operator or the vectors) ends the
      // solve with status 3 instead of iterating on it.
      bool finite = isfinite(theta);
      for (int i = 0; i < 14; ++i) finite = finite && isfinite(g[i].x) && isfinite(g[i].y);
      if (!finite) {
        stop_nonfinite(F);
        break;
      }
      const double rn = sqrt(fmax(g[13].x, 0.0));   // |R| (the residual itself)
      const double axn = sqrt(fmax(g[12].x, 0.0));
      if (criterion_met(prm, rn, axn, theta, S)) {
        F[kConv] = 1;
        break;
      }
      double2 U[6];
      bool pv = false;
      if (fused_rr(g, F[kPValid] != 0, prm.precision, U, pv) != 0) {
        F[kStatus] = kStatusBreakdown;
        F[kConv] = 1;
        break;
      }
      for (int i = 0; i < 6; ++i) S[kUpd + i] = U[i];
      S[kTheta] = U[5];
      F[kPValid] = pv ? 1 : 0;
      F[kNiter] += 1;
      break;
    }
    case kStepRefreshRho: {
      const double th = D[1].x / D[0].x;
      S[kTheta] = make_double2(th, 0.0);
      C[0] = make_double2(-th, 0.0);
      break;
    }
    case kStepVerify: {
      const double axn = sqrt(D[0].x), rn = sqrt(D[1].x);
      // a zero residual is an exact eigenpair even where the threshold is 0
      const bool conv = criterion_met(prm, rn, axn, S[kTheta].x, S) || rn == 0.0;
      F[kConv] = conv ? 1 : 0;
      break;
    }
    case kStepDebug: {
      // arg = record index; kappa of the record: kKappa..+2 (zeros for the
      // first record, (k0, 0, k1) after the 2x2 step)
      double2* r = prm.dbg + (int64_t)arg * kDbgRecord;
      const double theta = S[kTheta].x;
      r[0] = make_double2(theta, D[3].x);
      const double reltol = arg >= 2 ? (S[kInfo].y - theta) * prm.nlen * 10.0 : 0.0;
      r[1] = make_double2(reltol, arg >= 2 ? S[kInfo2].y : 0.0);
      r[2] = make_double2(arg >= 2 ? S[kInfo].x : 0.0, 0.0);
      for (int i = 0; i < 3; ++i)
        r[3 + i] = arg == 0 ? make_double2(0.0, 0.0)
                   : arg == 1 ? (i == 0 ? S[kKappa] : i == 1 ? make_double2(0.0, 0.0)
                                                             : S[kKappa + 1])
                              : S[kKappa + i];
      double2 m[3][3];
      m[1][0] = D[0];
      m[2][0] = D[1];
      m[2][1] = D[2];
      m[0][1] = cconj(D[0]);
      m[0][2] = cconj(D[1]);
      m[1][2] = cconj(D[2]);
      m[0][0] = D[3];
      m[1][1] = D[4];
      m[2][2] = D[5];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[6 + 3 * i + j] = m[i][j];
      break;
    }
    default:
      break;
  }
}

// diag[i] = Re sum_k c_k (-1)^{popcount(z_k & key_i)} over the X = 0 terms.
__global__ void __launch_bounds__(256)
diag_real_kernel(const uint64_t* __restrict__ zmask, const double2* __restrict__ coef, int t0,
                 int t1, const uint64_t* __restrict__ keys, int64_t n, double* __restrict__ diag) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= n) return;
  double cr, ci;
  group_coef<double2>(zmask, coef, t0, t1, keys[i], cr, ci);
  diag[i] = cr;
}

__global__ void init_slots_kernel(double2* S, int32_t* F) {
  const int i = threadIdx.x;
  if (i < kNumSlots) S[i] = make_double2(i == kOne ? 1.0 : 0.0, 0.0);
  if (i < kNumFlags) F[i] = 0;
}

__global__ void small_eig_kernel(int dim, const double2* __restrict__ mat, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double lam;
  if (dim == 2) {
    double2 m[2][2] = {{mat[0], mat[1]}, {mat[2], mat[3]}};
    double2 v[2];
    eig2_reference(m, lam, v);
    out[0] = lam;
    for (int i = 0; i < 2; ++i) { out[1 + 2 * i] = v[i].x; out[2 + 2 * i] = v[i].y; }
  } else {
    double2 m[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) m[i][j] = mat[3 * i + j];
    double2 v[3];
    eig3_reference(m, lam, v);
    out[0] = lam;
    for (int i = 0; i < 3; ++i) { out[1 + 2 * i] = v[i].x; out[2 + 2 * i] = v[i].y; }
  }
}

// ---- host driver -------------------------------------------------------------

struct Solver {
  const sqdk_op* op;
  int64_t n;
  int prec;
  hipStream_t s;
  double2* S;
  int32_t* F;
  double* partials;
  double* gpart;      // the split Gram path's partials (fused mode)
  char* vec[7];
  double* diag = nullptr;   // FUSED with the diagonal preconditioner
  void* mv_work = nullptr;
  size_t mv_bytes = 0;
  ScalarParams prm;
  int matvecs = 0;
  bool use_graph = false;   // FUSED: replay refresh periods as a captured HIP graph

  int apply(const void* in, void* out, bool skip) {
    if (op->fn) {
      ++matvecs;  // callbacks always run (the host polls every iteration)
      const int rc = op->fn(op->ctx, in, out, s);
      if (rc != 0) return fail(SQDK_ECALLBACK, "matvec callback returned %d", rc);
      return SQDK_OK;
    }
    if (op->cache)
      return hit_matvec_impl(op->cache, prec, in, out, s, skip ? F + kConv : nullptr,
                             F + kMatvecs);
    return matvec_impl(op->ham, op->space, prec, in, out, op->tune, mv_work, mv_bytes, s,
                       skip ? F + kConv : nullptr, F + kMatvecs);
  }
  // The hit table's product without its diagonal (the fused update adds it).
  int apply_offdiag(const void* in, void* out, bool skip) {
    return hit_matvec_impl(op->cache, prec, in, out, s, skip ? F + kConv : nullptr,
                           F + kMatvecs, nullptr, nullptr, 0, true);
  }
  int dots(std::initializer_list<std::pair<const void*, const void*>> pairs, bool skip,
           int slot = kD) {
    PairList pl{};
    int k = 0;
    for (auto& p : pairs) {
      pl.a[k] = p.first;
      pl.b[k] = p.second;
      ++k;
    }
    return dots_impl(n, prec, k, pl, reinterpret_cast<double*>(S + slot), partials, s,
                     skip ? F + kConv : nullptr);
  }
  int lin(void* out, std::initializer_list<const void*> in, std::initializer_list<int> cidx,
          bool skip) {
    const void* ins[4];
    int ci[4];
    int k = 0;
    for (auto p : in) ins[k++] = p;
    k = 0;
    for (auto c : cidx) ci[k++] = c;
    return lincomb_impl(n, prec, out, k, ins, ci, reinterpret_cast<const double*>(S), s,
                        skip ? F + kConv : nullptr);
  }
  int scalar(int step, bool skip, int arg = 0) {
    scalar_kernel<<<1, 64, 0, s>>>(step, S, F, prm, skip ? 1 : 0, arg);
    SQDK_LAUNCHED("scalar_kernel");
    return SQDK_OK;
  }
  int read_flags(int32_t* host) {
    SQDK_HIP(hipMemcpyAsync(host, F, kFlagBytes, hipMemcpyDeviceToHost, s));
    SQDK_HIP(hipStreamSynchronize(s));
    return SQDK_OK;
  }
  int read_slots(double2* host) {
    SQDK_HIP(hipMemcpyAsync(host, S, kSlotBytes, hipMemcpyDeviceToHost, s));
    SQDK_HIP(hipStreamSynchronize(s));
    return SQDK_OK;
  }
};

constexpr int C0 = kC, C1 = kC + 1, C2 = kC + 2, C3 = kC + 3;

int run_reference(Solver& sv, void* x, const sqdk_locg_opts& o, int32_t* flags_host) {
  char* y = sv.vec[0];
  char* r = sv.vec[1];
  char* p = sv.vec[2];
  char* ax = sv.vec[3];
  char* ay = sv.vec[4];
  char* ap = sv.vec[5];
  // Diagnostics record k (ground_locg.py:211-233): S^H A S of (x, y, r) from
  // three extra products into the scratch vectors ax, ay, ap (free at every
  // record point), then optionally the three vectors themselves.
  const size_t vb = (size_t)sv.n * elem_bytes(sv.prec);
  auto record = [&](int k, bool skip) -> int {
    if (!sv.prm.dbg) return SQDK_OK;
    SQDK_TRY(sv.apply(x, ax, skip));
    SQDK_TRY(sv.apply(y, ay, skip));
    SQDK_TRY(sv.apply(r, ap, skip));
    SQDK_TRY(sv.dots({{y, ax}, {r, ax}, {r, ay}, {x, ax}, {y, ay}, {r, ap}}, skip));
    SQDK_TRY(sv.scalar(kStepDebug, skip, k));
    if (o.debug_vectors) {
      char* dst = static_cast<char*>(o.debug_vectors) + (size_t)k * 3 * vb;
      SQDK_HIP(hipMemcpyAsync(dst, x, vb, hipMemcpyDeviceToDevice, sv.s));
      SQDK_HIP(hipMemcpyAsync(dst + vb, y, vb, hipMemcpyDeviceToDevice, sv.s));
      SQDK_HIP(hipMemcpyAsync(dst + 2 * vb, r, vb, hipMemcpyDeviceToDevice, sv.s));
    }
    return SQDK_OK;
  };
  // xinit /= |xinit|  (ground_locg.py:346)
  SQDK_TRY(sv.dots({{x, x}}, false));
  SQDK_TRY(sv.scalar(kStepInitNorm, false));
  SQDK_TRY(sv.lin(x, {x}, {C0}, false));
  // body_iter0 (:264-272): r = Ax - rho x
  SQDK_TRY(sv.apply(x, ax, false));
  SQDK_TRY(sv.dots({{x, ax}}, false));
  SQDK_TRY(sv.scalar(kStepRho, false));
  SQDK_TRY(sv.lin(r, {ax, x}, {kOne, C0}, false));
  if (sv.prm.dbg) {
    SQDK_TRY(sv.lin(y, {x}, {kZero}, false));   // y = 0 in the first record
    SQDK_TRY(record(0, false));
  }
  // body_iter1 (:274-286); skipped when the start is an exact eigenvector
  SQDK_TRY(sv.dots({{r, r}}, false));
  SQDK_TRY(sv.scalar(kStepZeroResid, false));
  SQDK_TRY(sv.scalar(kStepInvNormGuard, true));
  SQDK_TRY(sv.lin(p, {r}, {C0}, true));
  SQDK_TRY(sv.apply(x, ax, true));
  SQDK_TRY(sv.apply(p, ap, true));
  SQDK_TRY(sv.dots({{p, ax}, {x, ax}, {p, ap}}, true));
  SQDK_TRY(sv.scalar(kStepEig2, true));
  SQDK_TRY(sv.lin(y, {p, x}, {C0, C1}, true));   // t
  SQDK_TRY(sv.lin(ay, {x, p}, {C2, C3}, true));  // u
  SQDK_TRY(sv.dots({{ay, ay}, {y, y}}, true));
  SQDK_TRY(sv.scalar(kStepInvNorm2, true));
  SQDK_TRY(sv.lin(x, {ay}, {C0}, true));
  SQDK_TRY(sv.lin(y, {y}, {C1}, true));
  SQDK_TRY(sv.apply(x, ax, true));
  SQDK_TRY(sv.scalar(kStepNegTheta, true));
  SQDK_TRY(sv.lin(r, {ax, x}, {kOne, C0}, true));
  SQDK_TRY(record(1, true));

  int done = 0;
  int rec = 2;
  while (done < o.maxiter) {
    const int chunk = std::min(o.check_every, o.maxiter - done);
    for (int it = 0; it < chunk; ++it) {
      // _project_out((x, y), r) (:380-421)
      const void* src = r;
      for (int rep = 0; rep < 2; ++rep) {
        SQDK_TRY(sv.dots({{x, src}, {y, src}}, true));
        SQDK_TRY(sv.scalar(kStepProjNeg, true));
        SQDK_TRY(sv.lin(p, {src, x, y}, {kOne, C0, C1}, true));
        SQDK_TRY(sv.dots({{p, p}}, true));
        SQDK_TRY(sv.scalar(kStepInvNormGuard, true));
        SQDK_TRY(sv.lin(p, {p}, {C0}, true));
        src = p;
      }
      for (int rep = 0; rep < 2; ++rep) {
        SQDK_TRY(sv.dots({{x, p}, {y, p}}, true));
        SQDK_TRY(sv.scalar(kStepProjNeg, true));
        SQDK_TRY(sv.lin(p, {p, x, y}, {kOne, C0, C1}, true));
      }
      SQDK_TRY(sv.dots({{p, p}}, true));
      SQDK_TRY(sv.scalar(kStepThresh, true));
      SQDK_TRY(sv.lin(p, {p}, {C0}, true));
      // rayleigh_ritz(x, y, p) (:235-255)
      SQDK_TRY(sv.apply(x, ax, true));
      SQDK_TRY(sv.apply(y, ay, true));
      SQDK_TRY(sv.apply(p, ap, true));
      SQDK_TRY(sv.dots({{y, ax}, {p, ax}, {p, ay}, {x, ax}, {y, ay}, {p, ap}}, true));
      SQDK_TRY(sv.scalar(kStepEig3, true));
      // s = y k1 + p k2 -> ay ; t = s k0/ns - x ns -> ap ; u = x k0 + s -> ax
      SQDK_TRY(sv.lin(ay, {y, p}, {C0, C1}, true));
      SQDK_TRY(sv.dots({{ay, ay}}, true));
      SQDK_TRY(sv.scalar(kStepTU, true));
      SQDK_TRY(sv.lin(ap, {ay, x}, {C0, C1}, true));
      SQDK_TRY(sv.lin(ax, {x, ay}, {C2, C3}, true));
      SQDK_TRY(sv.dots({{ax, ax}, {ap, ap}}, true));
      SQDK_TRY(sv.scalar(kStepInvNorm2, true));
      SQDK_TRY(sv.lin(x, {ax}, {C0}, true));
      SQDK_TRY(sv.lin(y, {ap}, {C1}, true));
      // r' = A x' - theta x' and the convergence test (:306-334)
      SQDK_TRY(sv.apply(x, ax, true));
      SQDK_TRY(sv.scalar(kStepNegTheta, true));
      SQDK_TRY(sv.lin(r, {ax, x}, {kOne, C0}, true));
      SQDK_TRY(sv.dots({{ax, ax}, {r, r}}, true));
      SQDK_TRY(sv.scalar(kStepConv, true));
      SQDK_TRY(record(rec++, true));
      if (sv.op->fn && it + 1 < chunk) break;  // callbacks cannot be skipped on device
    }
    SQDK_TRY(sv.read_flags(flags_host));
    done = flags_host[kNiter];
    if (flags_host[kConv]) break;
  }
  return SQDK_OK;
}

// FUSED: one matvec (A W) per iteration over the basis (X, W, P), where W is
// the residual R itself or, with the diagonal preconditioner, R scaled by
// 1 / max(diag - min(theta, min diag), floor) (Jacobi preconditioned LOBPCG,
// the shift kept at or below the smallest diagonal entry).
int run_fused(Solver& sv, void* x, const sqdk_locg_opts& o, int32_t* flags_host) {
  char* X = static_cast<char*>(x);
  char* R = sv.vec[0];
  char* P = sv.vec[1];
  char* AX = sv.vec[2];
  char* AW = sv.vec[3];
  char* AP = sv.vec[4];
  char* W = sv.diag ? sv.vec[5] : R;
  const size_t vb = (size_t)sv.n * elem_bytes(sv.prec);
  auto precondition = [&](bool skip) -> int {
    if (!sv.diag) return SQDK_OK;
    return precond_impl(sv.n, sv.prec, R, sv.diag, reinterpret_cast<const double*>(sv.S),
                        kTheta, kDiagStat, W, sv.s, skip ? sv.F + kConv : nullptr);
  };
  SQDK_HIP(hipMemsetAsync(P, 0, vb, sv.s));
  SQDK_HIP(hipMemsetAsync(AP, 0, vb, sv.s));
  SQDK_TRY(sv.dots({{X, X}}, false));
  SQDK_TRY(sv.scalar(kStepInitNorm, false));
  SQDK_TRY(sv.lin(X, {X}, {C0}, false));
  SQDK_TRY(sv.apply(X, AX, false));
  SQDK_TRY(sv.dots({{X, AX}}, false));
  SQDK_TRY(sv.scalar(kStepRho, false));
  SQDK_TRY(sv.lin(R, {AX, X}, {kOne, C0}, false));
  SQDK_TRY(sv.dots({{R, R}}, false));
  SQDK_TRY(sv.scalar(kStepZeroResid, false));   // an exact eigenvector stops here
  SQDK_TRY(precondition(false));

  const void* V[3] = {X, W, P};
  const void* AV[3] = {AX, AW, AP};
  const void* res = sv.diag ? R : nullptr;
  int done = 0;
  int since_refresh = 0;
  // The Gram sums come from a full pass over the six basis vectors on the
  // first iteration and after X, AX, R or P were rebuilt; otherwise the
  // previous update pass left all of them but R.AR as partials, and only W
  // and A W are read after the product.
  bool full_gram = true;
  double* gram_out = reinterpret_cast<double*>(sv.S + kGram);
  // With the hit table and the diagonal preconditioner, the products inside a
  // refresh period skip the table's diagonal: the update pass, which reads W
  // and a diagonal anyway, adds it (and its share of R.AR), so a product reads
  // 2 vector elements less per row. Products whose result the full Gram pass
  // reads (a period's first) and the refreshes stay whole.
  const bool split_diag = sv.diag && sv.op->cache && !sv.op->fn;
  const void* cdiag = split_diag ? sv.op->cache->diag : nullptr;
  auto iterate = [&](int count) -> int {
    for (int it = 0; it < count; ++it) {
      const bool off = split_diag && !full_gram;
      if (off)
        SQDK_TRY(sv.apply_offdiag(W, AW, true));
      else
        SQDK_TRY(sv.apply(W, AW, true));
      if (full_gram)
        SQDK_TRY(gram_impl(sv.n, sv.prec, V, AV, res, gram_out, sv.partials, sv.s,
                           sv.F + kConv));
      else
        SQDK_TRY(gram_rar_impl(sv.n, sv.prec, W, AW, sv.gpart, gram_out, sv.s, sv.F + kConv,
                               off));
      SQDK_TRY(sv.scalar(kStepFusedRR, true));
      // the next iteration's product skips the diagonal unless a refresh
      // (after this update) makes it a period's first
      const bool refresh_next = o.refresh_every > 0 && since_refresh + 1 >= o.refresh_every;
      const bool off_next = split_diag && !refresh_next;
      SQDK_TRY(fused_update_impl(sv.n, sv.prec, X, AX, R, W, AW, P, AP, sv.diag,
                                 reinterpret_cast<const double*>(sv.S), kUpd, kDiagStat,
                                 sv.gpart, sv.s, sv.F + kConv, cdiag,
                                 (off ? 1 : 0) | (off_next ? 2 : 0)));
      full_gram = false;
      if (o.refresh_every > 0 && ++since_refresh >= o.refresh_every) {
        since_refresh = 0;
        full_gram = true;
        SQDK_TRY(sv.apply(X, AX, true));
        SQDK_TRY(sv.apply(P, AP, true));
        SQDK_TRY(sv.dots({{X, X}, {X, AX}}, true));
        SQDK_TRY(sv.scalar(kStepRefreshRho, true));
        SQDK_TRY(sv.lin(R, {AX, X}, {kOne, C0}, true));
        SQDK_TRY(precondition(true));
      }
      if (sv.op->fn && it + 1 < count) break;
    }
    return SQDK_OK;
  };
  // With a graph, one refresh period (refresh_every iterations, the first
  // with the full Gram pass, the last followed by the explicit refresh) is
  // captured once and replayed whenever a period starts and fits below
  // maxiter; the host then polls once per period. Every kernel in it skips
  // its work once the convergence flag is set.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  struct GraphGuard {
    hipGraph_t& g;
    hipGraphExec_t& e;
    ~GraphGuard() {
      if (e) hipGraphExecDestroy(e);
      if (g) hipGraphDestroy(g);
    }
  } guard{graph, exec};
  const bool graphs = sv.use_graph && o.refresh_every > 0 && !sv.op->fn;
  while (true) {
    const int left = o.maxiter - done;
    if (graphs && full_gram && since_refresh == 0 && left >= o.refresh_every) {
      if (!exec) {
        SQDK_HIP(hipStreamBeginCapture(sv.s, hipStreamCaptureModeThreadLocal));
        const int rc = iterate(o.refresh_every);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(sv.s, &g);
        if (rc != SQDK_OK) {
          if (g) hipGraphDestroy(g);
          return rc;
        }
        if (e != hipSuccess)
          return fail(SQDK_EHIP, "graph capture failed: %s", hipGetErrorString(e));
        graph = g;
        SQDK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      }
      SQDK_HIP(hipGraphLaunch(exec, sv.s));
      since_refresh = 0;   // the period ends with the refresh
      full_gram = true;
    } else {
      SQDK_TRY(iterate(std::min(o.check_every, left)));
    }
    SQDK_TRY(sv.read_flags(flags_host));
    done = flags_host[kNiter];
    const bool conv = flags_host[kConv] != 0;
    if (conv || done >= o.maxiter) {
      // Verify with an explicit product: A X, theta = <X, AX> / <X, X>, r = AX - theta X.
      SQDK_TRY(sv.apply(X, AX, false));
      SQDK_TRY(sv.dots({{X, X}, {X, AX}}, false));
      SQDK_TRY(sv.scalar(kStepRefreshRho, false));
      SQDK_TRY(sv.lin(R, {AX, X}, {kOne, C0}, false));
      SQDK_TRY(sv.dots({{AX, AX}, {R, R}}, false));
      SQDK_TRY(sv.scalar(kStepVerify, false));
      SQDK_TRY(sv.read_flags(flags_host));
      if (flags_host[kConv] || done >= o.maxiter || flags_host[kStatus] >= 2) break;
      // Not converged after all: A P is refreshed too and the loop goes on.
      SQDK_TRY(sv.apply(P, AP, false));
      SQDK_TRY(precondition(false));
      since_refresh = 0;
      full_gram = true;
    }
  }
  return SQDK_OK;
}

}  // namespace

}  // namespace sqdk

using namespace sqdk;

extern "C" {

int sqdk_locg_workspace(const sqdk_op* op, int64_t n, const sqdk_locg_opts* opts,
                        size_t* bytes) {
  SQDK_REQUIRE(opts && bytes, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1, "empty problem");
  SQDK_REQUIRE(opts->precision == SQDK_C128 || opts->precision == SQDK_C64,
               "unknown precision %d", opts->precision);
  const size_t vb = align256((size_t)n * elem_bytes(opts->precision));
  size_t mv = 0;
  if (op && !op->fn && op->space) mv = align256(matvec_work_bytes(op->space, op->tune, opts->precision));
  const size_t dg = opts->precond ? align256((size_t)n * sizeof(double)) : 0;
  *bytes = align256(kSlotBytes) + align256(kFlagBytes) + align256(kPartialBytes) +
           align256(kGramPartialBytes) + 7 * vb + mv + dg;
  return SQDK_OK;
}

int sqdk_locg_solve(const sqdk_op* op, int64_t n, void* x, const sqdk_locg_opts* opts,
                    void* work, size_t work_bytes, sqdk_locg_result* result, void* stream) {
  SQDK_REQUIRE(op && x && opts && work && result, "NULL pointer argument");
  SQDK_REQUIRE(op->fn || (op->ham && op->space), "operator has neither a callback nor ham+space");
  SQDK_REQUIRE(opts->mode == SQDK_LOCG_REFERENCE || opts->mode == SQDK_LOCG_FUSED,
               "unknown mode %d", opts->mode);
  SQDK_REQUIRE(opts->maxiter >= 0, "maxiter %d < 0", opts->maxiter);
  SQDK_REQUIRE(opts->criterion == SQDK_CRIT_REFERENCE || opts->criterion == SQDK_CRIT_RELATIVE,
               "unknown criterion %d", opts->criterion);
  if (!op->fn) {
    SQDK_REQUIRE(op->space->n == n, "operator size %lld != vector size %lld",
                 (long long)op->space->n, (long long)n);
  }
  SQDK_REQUIRE(!op->cache || op->fn || (op->cache->precision == opts->precision &&
                                        op->cache->n == n),
               "cached hit table does not match the solve (precision or size)");
  SQDK_REQUIRE(opts->precond == SQDK_PRECOND_NONE || opts->precond == SQDK_PRECOND_DIAG,
               "unknown preconditioner %d", opts->precond);
  SQDK_REQUIRE(opts->precond == SQDK_PRECOND_NONE ||
                   (opts->mode == SQDK_LOCG_FUSED && !op->fn),
               "the diagonal preconditioner needs the FUSED mode and the built-in operator");
  SQDK_REQUIRE(!opts->debug || opts->mode == SQDK_LOCG_REFERENCE,
               "diagnostics are recorded by the REFERENCE mode only");
  SQDK_REQUIRE(!opts->debug_vectors || opts->debug, "debug_vectors needs debug");
  size_t need = 0;
  SQDK_TRY(sqdk_locg_workspace(op, n, opts, &need));
  SQDK_REQUIRE(work_bytes >= need, "workspace of %zu bytes, %zu needed", work_bytes, need);
  sqdk_locg_opts o = *opts;
  if (o.check_every < 1) o.check_every = 1;
  if (op->fn) o.check_every = 1;

  Solver sv{};
  sv.op = op;
  sv.n = n;
  sv.prec = o.precision;
  sv.s = as_stream(stream);
  sv.use_graph = o.use_graph && o.mode == SQDK_LOCG_FUSED && !op->fn && o.refresh_every > 0 &&
                 o.maxiter >= o.refresh_every;
  // Graph capture needs a stream other than the legacy default one: a private
  // stream, after the caller's has finished what it queued (the start vector).
  struct OwnStream {
    hipStream_t s = nullptr;
    ~OwnStream() {
      if (s) {
        hipStreamSynchronize(s);
        hipStreamDestroy(s);
      }
    }
  } own;
  if (sv.use_graph) {
    SQDK_HIP(hipStreamSynchronize(sv.s));
    SQDK_HIP(hipStreamCreateWithFlags(&own.s, hipStreamNonBlocking));
    sv.s = own.s;
  }
  char* w = static_cast<char*>(work);
  sv.S = reinterpret_cast<double2*>(w);
  w += align256(kSlotBytes);
  sv.F = reinterpret_cast<int32_t*>(w);
  w += align256(kFlagBytes);
  sv.partials = reinterpret_cast<double*>(w);
  w += align256(kPartialBytes);
  sv.gpart = reinterpret_cast<double*>(w);
  w += align256(kGramPartialBytes);
  const size_t vb = align256((size_t)n * elem_bytes(o.precision));
  for (int i = 0; i < 7; ++i) sv.vec[i] = w + i * vb;
  if (!op->fn) {
    sv.mv_bytes = matvec_work_bytes(op->space, op->tune, o.precision);
    sv.mv_work = sv.mv_bytes ? w + 7 * vb : nullptr;
    if (o.precond == SQDK_PRECOND_DIAG) {
      sv.diag = reinterpret_cast<double*>(w + 7 * vb + align256(sv.mv_bytes));
      const sqdk_ham* h = op->ham;
      diag_real_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, sv.s>>>(
          h->zmask, reinterpret_cast<const double2*>(h->coef), h->diag_t0, h->diag_t1,
          op->space->keys, n, sv.diag);
      SQDK_LAUNCHED("diag_real_kernel");
    }
  }
  sv.prm.tol = o.tol;
  sv.prm.nlen = (double)(o.n_for_tol > 0 ? o.n_for_tol : n);
  sv.prm.n = n;
  sv.prm.dbg = reinterpret_cast<double2*>(o.debug);
  sv.prm.criterion = o.criterion;
  sv.prm.precision = o.precision;

  init_slots_kernel<<<1, 128, 0, sv.s>>>(sv.S, sv.F);
  SQDK_LAUNCHED("init_slots_kernel");
  if (sv.diag)
    SQDK_TRY(diag_stats_impl(n, sv.diag, sv.partials, kDotBlocks,
                             reinterpret_cast<double*>(sv.S + kDiagStat), sv.s));
  int32_t flags[kNumFlags] = {0};
  int rc = o.mode == SQDK_LOCG_REFERENCE ? run_reference(sv, x, o, flags)
                                         : run_fused(sv, x, o, flags);
  if (rc != SQDK_OK) return rc;
  double2 slots[kNumSlots];
  SQDK_TRY(sv.read_slots(slots));
  SQDK_TRY(sv.read_flags(flags));
  result->eigval = slots[kTheta].x;
  result->resid_norm = slots[kInfo].x;
  result->ax_norm = slots[kInfo].y;
  result->niter = flags[kNiter];
  // a stop on NaN or Inf sets the flag that ends the loop but is no convergence
  result->converged = flags[kConv] && flags[kStatus] != kStatusNonFinite;
  result->matvecs = sv.matvecs + flags[kMatvecs];
  result->status = flags[kStatus];
  return SQDK_OK;
}

int sqdk_small_eigenpair(int32_t dim, const double* mat, double* out, void* stream) {
  SQDK_REQUIRE(mat && out, "NULL pointer argument");
  SQDK_REQUIRE(dim == 2 || dim == 3, "dim %d not 2 or 3", dim);
  small_eig_kernel<<<1, 64, 0, as_stream(stream)>>>(dim, reinterpret_cast<const double2*>(mat), out);
  SQDK_LAUNCHED("small_eig_kernel");
  return SQDK_OK;
}

}  // extern "C"
