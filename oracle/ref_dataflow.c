/*
 * ref_dataflow.c -- plain-C restatement of the reference's matvec dataflow
 * (test and CPU-baseline infrastructure only; the product never loads it).
 *
 * Follows rqutils/sqd.py with the subspace as sorted uint64 keys (ascending
 * keys are the reference's lexicographic order of padded packed rows,
 * sqd.py:237, 424-426):
 *   ref_xsource     get_xsource (sqd.py:448-503): a stable sort of the 2N
 *                   rows [S; S ^ x] with their indices (an LSD radix sort of
 *                   (key, index) pairs: stable, as lax.sort is), then each
 *                   original row takes the index of its sorted successor when
 *                   the two keys are equal, else -1.
 *   ref_diagonal    get_diagonal (sqd.py:554-575): sum over the group's terms,
 *                   in order, of c_k (-1)^{popcount(z_k & s)}; the sum stops
 *                   at the first zero coefficient, as the reference's does.
 *   ref_matvec      apply_h (sqd.py:590-605; cache_level (0, 0)) or, with
 *                   source indices given, apply_h_x_cached (sqd.py:641-654):
 *                   out += C_g * v[xsource_g] (0 where the source is -1) for
 *                   the requested groups.
 * Groups run in parallel over OpenMP threads, each accumulating its own
 * partial output; the partials are summed in thread order at the end.
 * Complex numbers are interleaved (re, im) doubles.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef struct {
  uint64_t key;
  uint32_t idx;
  uint32_t pad;
} pair_t;

/* stable LSD radix sort of n pairs on the low `bits` bits of key, 8 at a
 * time, ping-ponging between a and tmp; the result ends in a */
static void radix_sort(pair_t* a, pair_t* tmp, int64_t n, int bits) {
  pair_t* src = a;
  pair_t* dst = tmp;
  for (int shift = 0; shift < bits; shift += 8) {
    int64_t count[257];
    memset(count, 0, sizeof(count));
    for (int64_t i = 0; i < n; ++i) count[((src[i].key >> shift) & 255) + 1]++;
    for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
    for (int64_t i = 0; i < n; ++i) dst[count[(src[i].key >> shift) & 255]++] = src[i];
    pair_t* t = src;
    src = dst;
    dst = t;
  }
  if (src != a) memcpy(a, src, (size_t)n * sizeof(pair_t));
}

/* xsource[i] for one X mask; work holds 4 n pairs */
static void xsource_one(const uint64_t* keys, int64_t n, int bits, uint64_t x, pair_t* work,
                        int32_t* out) {
  pair_t* p = work;
  pair_t* tmp = work + 2 * n;
  for (int64_t i = 0; i < n; ++i) {
    p[i].key = keys[i];
    p[i].idx = (uint32_t)i;
    p[n + i].key = keys[i] ^ x;
    p[n + i].idx = (uint32_t)(n + i);
  }
  radix_sort(p, tmp, 2 * n, bits);
  for (int64_t k = 0; k < 2 * n; ++k) {
    if (p[k].idx >= (uint32_t)n) continue;
    const int eq = k + 1 < 2 * n && p[k + 1].key == p[k].key;
    out[p[k].idx] = eq ? (int32_t)(p[k + 1].idx - n) : -1;
  }
}

static void diagonal_one(const uint64_t* keys, int64_t n, const uint64_t* zmask,
                         const double* coef, int32_t t0, int32_t t1, double* d) {
  for (int64_t i = 0; i < n; ++i) d[2 * i] = d[2 * i + 1] = 0.0;
  for (int32_t k = t0; k < t1; ++k) {
    const double cr = coef[2 * k], ci = coef[2 * k + 1];
    if (cr == 0.0 && ci == 0.0) break;
    const uint64_t z = zmask[k];
    for (int64_t i = 0; i < n; ++i) {
      const double s = (__builtin_popcountll(z & keys[i]) & 1) ? -1.0 : 1.0;
      d[2 * i] += cr * s;
      d[2 * i + 1] += ci * s;
    }
  }
}

int ref_xsource(const uint64_t* keys, int64_t n, int bits, uint64_t x, int32_t* out) {
  pair_t* work = (pair_t*)malloc((size_t)4 * n * sizeof(pair_t));
  if (!work) return -1;
  xsource_one(keys, n, bits, x, work, out);
  free(work);
  return 0;
}

int ref_diagonal(const uint64_t* keys, int64_t n, const uint64_t* zmask, const double* coef,
                 int32_t t0, int32_t t1, double* out) {
  diagonal_one(keys, n, zmask, coef, t0, t1, out);
  return 0;
}

/* out (2n doubles) = sum over groups[0..ngroups) of C_g * v[xsource_g].
 * xmask[g], grp_off[g]..grp_off[g+1] are group g's mask and term range.
 * xsources (may be NULL) holds precomputed source indices, row-major
 * [ngroups][n] in the order of `groups`. Returns the thread count used. */
int ref_matvec(const uint64_t* keys, int64_t n, int bits, const uint64_t* xmask,
               const int32_t* grp_off, const uint64_t* zmask, const double* coef,
               const int32_t* groups, int32_t ngroups, const int32_t* xsources,
               const double* v, double* out, int32_t nthreads) {
  if (nthreads < 1) nthreads = 1;
  double* partial = (double*)calloc((size_t)nthreads * 2 * n, sizeof(double));
  if (!partial) return -1;
  int failed = 0;
#pragma omp parallel num_threads(nthreads)
  {
    const int t = omp_get_thread_num();
    double* acc = partial + (size_t)t * 2 * n;
    double* d = (double*)malloc((size_t)2 * n * sizeof(double));
    int32_t* xs = xsources ? NULL : (int32_t*)malloc((size_t)n * sizeof(int32_t));
    pair_t* work = xsources ? NULL : (pair_t*)malloc((size_t)4 * n * sizeof(pair_t));
    if (!d || (!xsources && (!xs || !work))) {
#pragma omp atomic write
      failed = 1;
    }
    /* every thread sees every failure before any of them enters the loop */
#pragma omp barrier
    int any_failed;
#pragma omp atomic read
    any_failed = failed;
    if (!any_failed) {
#pragma omp for schedule(dynamic, 1)
      for (int32_t j = 0; j < ngroups; ++j) {
        const int32_t g = groups[j];
        const int32_t* src = xs;
        if (xsources)
          src = xsources + (size_t)j * n;
        else
          xsource_one(keys, n, bits, xmask[g], work, xs);
        diagonal_one(keys, n, zmask, coef, grp_off[g], grp_off[g + 1], d);
        for (int64_t i = 0; i < n; ++i) {
          const int32_t s = src[i];
          if (s < 0) continue;
          const double vr = v[2 * (int64_t)s], vi = v[2 * (int64_t)s + 1];
          acc[2 * i] += d[2 * i] * vr - d[2 * i + 1] * vi;
          acc[2 * i + 1] += d[2 * i] * vi + d[2 * i + 1] * vr;
        }
      }
    }
    free(d);
    free(xs);
    free(work);
  }
  if (!failed) {
    for (int64_t i = 0; i < 2 * n; ++i) {
      double s = 0.0;
      for (int t = 0; t < nthreads; ++t) s += partial[(size_t)t * 2 * n + i];
      out[i] = s;
    }
  }
  free(partial);
  return failed ? -1 : nthreads;
}
