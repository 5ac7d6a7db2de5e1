// Internal launch helpers shared by the matvec, vector and LOBPCG sources.
// Every helper that takes `skip` turns into a no-op when *skip != 0 at kernel
// start: the LOBPCG enqueues several iterations ahead and stops work on the
// device once its convergence flag is set.
#pragma once

#include "common.hpp"

namespace sqdk {

int check_space_public(const sqdk_space* sp);

// Matvec with an optional device skip flag: when *skip != 0 at kernel start
// the launch does nothing (used for iterations after convergence inside a
// chunk enqueued ahead). `counter`, if set, counts the launches that ran.
int matvec_impl(const sqdk_ham* ham, const sqdk_space* space, int32_t precision,
                const void* v, void* out, const sqdk_tune* tune, void* work, size_t work_bytes,
                hipStream_t s, const int32_t* skip = nullptr, int32_t* counter = nullptr);

// Batched matvec; inst_skip (per instance, may be NULL) skips converged instances.
int batch_matvec_impl(const sqdk_batch* b, int32_t precision, const void* v, void* out,
                      const sqdk_tune* tune, void* work, size_t work_bytes, hipStream_t s,
                      const int32_t* skip, const int32_t* inst_skip, int32_t* counter);

// Workspace bytes the matvec needs for this tuning (partial sums of a group split).
size_t matvec_work_bytes(const sqdk_space* space, const sqdk_tune* tune, int32_t precision);

constexpr int kMaxPairs = 16;
constexpr int kDotBlocks = 512;   // partial sums per pair (fixed: deterministic)

struct PairList {
  const void* a[kMaxPairs];
  const void* b[kMaxPairs];
};

// out[k] (complex double) = sum_i conj(a_k[i]) b_k[i]; partials needs
// kDotBlocks * npairs * 2 doubles.
int dots_impl(int64_t n, int precision, int npairs, const PairList& pl, double* out,
              double* partials, hipStream_t s, const int32_t* skip = nullptr);

// out = sum_k S[cidx[k]] * in_k (complex coefficients read from device slots).
int lincomb_impl(int64_t n, int precision, void* out, int k, const void* const* in,
                 const int* cidx, const double* slots, hipStream_t s,
                 const int32_t* skip = nullptr);

// Gram and projected-operator sums of the fused LOBPCG over
// V = (X, W, P), AV = (AX, AW, AP): 13 complex sums, then |res|^2 (res ==
// NULL: |W|^2), written to out.
int gram_impl(int64_t n, int precision, const void* const* v, const void* const* av,
              const void* res, double* out, double* partials, hipStream_t s,
              const int32_t* skip = nullptr);

// Diagonal preconditioner of the fused LOBPCG (theta = the Ritz value; st =
// (min diag, 0.02 (mean diag - min diag)) in a slot): R scaled by
// 1 / max(diag - min(theta, min diag), max(st.y, 1e-3 |theta|)).
//
// Fused LOBPCG update; coefficients at slots[c_base ...]:
//   X' = c0 X + c1 W + c2 P, AX' = c0 AX + c1 AW + c2 AP,
//   P' = d1 W + d2 P, AP' = d1 AW + d2 AP, R' = AX' - theta' X',
// and with diag given W' = the preconditioned R' (the statistics at
// slots[stat_slot]; W may alias R when diag is NULL).
// With gpart (kGramPartialDoubles doubles), the pass also leaves the next
// iteration's Gram sums, except R.AR, as per-block partials there;
// gram_rar_impl adds R.AR after the next product and reduces all 14.
// cdiag (the hit table's diagonal, cache precision; with diag only) replaces
// diag in the preconditioner; cdiag_mode bit 0: AW is the product without its
// diagonal, which the pass adds (cdiag W); bit 1: the next product will be
// without it too, so the pass leaves sum cdiag |W'|^2 as R.AR's partial for
// gram_rar_impl to add to (accumulate).
int fused_update_impl(int64_t n, int precision, void* X, void* AX, void* R, void* W,
                      const void* AW, void* P, void* AP, const double* diag,
                      const double* slots, int c_base, int stat_slot, double* gpart,
                      hipStream_t s, const int32_t* skip = nullptr,
                      const void* cdiag = nullptr, int cdiag_mode = 0);
int gram_rar_impl(int64_t n, int precision, const void* W, const void* AW, double* gpart,
                  double* out, hipStream_t s, const int32_t* skip = nullptr,
                  bool accumulate = false);
// Partial sums of the split Gram path: 28 doubles per block of a streaming pass.
constexpr int64_t kGramPartialDoubles = (int64_t)28 * 8192;

// W = the preconditioned R at theta = slots[theta_slot].x.
int precond_impl(int64_t n, int precision, const void* R, const double* diag,
                 const double* slots, int theta_slot, int stat_slot, void* W, hipStream_t s,
                 const int32_t* skip = nullptr);

// out (a complex slot) = (min diag, 0.02 (mean diag - min diag)); partials
// holds 2 * max_blocks doubles.
int diag_stats_impl(int64_t n, const double* diag, double* partials, int max_blocks,
                    double* out, hipStream_t s);

// out = the cached hit table's product (precision must be the cache's), or
// with offdiag its part off the diagonal (the table's diagonal and v's own
// entries are not read). In a batch, item_inst / inst_skip / slices_per_item
// leave converged instances' slices alone.
int hit_matvec_impl(const sqdk_cache* c, int32_t precision, const void* v, void* out,
                    hipStream_t s, const int32_t* skip = nullptr, int32_t* counter = nullptr,
                    const uint16_t* item_inst = nullptr, const int32_t* inst_skip = nullptr,
                    int slices_per_item = 0, bool offdiag = false);

int convert_impl(int64_t n, int from_p, const void* in, int to_p, void* out, hipStream_t s);

inline size_t elem_bytes(int precision) { return precision == SQDK_C128 ? 16 : 8; }

}  // namespace sqdk
